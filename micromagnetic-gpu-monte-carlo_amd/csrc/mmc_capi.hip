// mmc_capi.hip -- context, device memory, work lists and the C ABI.
//
// Host-side restatement of the reference's set-up and control flow
// (src/main.cu:254-756 allocations, padding, tensor + cache, FFT plans;
// :801-1139 the anneal loop), driving the kernels of mmc_kernels.hip on one
// HIP stream per context.
#include "mmc.h"
#include "mmc_internal.h"
#include "fft_fixed.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <vector>

using namespace mmc;

namespace {

constexpr double kPiTrunc = 3.14159265358;       // src/constant.h:21
constexpr double kBoltzmann = 1.3806488E-23;     // src/constant.h:22

const char *kKinds[] = {"tensor", "xfwd", "zfwd", "ymul", "zinv", "xinv", "metro", "energy",
                        "xinv_metro", "metro_xfwd", "yzconv", "row_sweep"};
constexpr int kNumKinds = 12;
enum Kind { K_TENSOR, K_XFWD, K_ZFWD, K_YMUL, K_ZINV, K_XINV, K_METRO, K_ENERGY, K_XINVM, K_MXFWD,
            K_YZC, K_ROWS };

struct Segment {
    int fixed_n = 0;   // common FFT length of the segment's kernel, 0 = generic
    int rows = 0;      // row pairs (x pass) or columns (y pass) per fixed-plan tile
    int begin = 0, count = 0;
    size_t lds = 0;
    double bytes = 0;  // algorithmic HBM bytes of one launch of the segment
    bool half = true;  // y pass: every instance's rows fill at most half the padded length
};

struct List {
    std::vector<Work> host;
    std::vector<int> first;   // first item of each instance, size count+1
    std::vector<Segment> seg; // x and y passes: items grouped by kernel variant
    Work *dev = nullptr;
    size_t lds = 0;
    int n() const { return (int)host.size(); }
};

bool is_smooth(int v)
{
    for (int p : {2, 3, 5}) while (v % p == 0) v /= p;
    return v == 1;
}

int padded_len(int n, int pbc)                  // src/main.cu:254-263
{
    if (pbc != 0) return n;
    int v = std::max(1, 2 * n - 1);
    while (!is_smooth(v)) ++v;
    return v;
}

bool make_plan(int n, FftPlan &pl)
{
    pl.n = n;
    pl.nst = 0;
    int v = n;
    while (v % 4 == 0) { pl.radix[pl.nst++] = 4; v /= 4; }
    for (int p : {2, 3, 5, 7})
        while (v % p == 0) {
            if (pl.nst >= kMaxStages) return false;
            pl.radix[pl.nst++] = p;
            v /= p;
        }
    return v == 1;
}

// x lengths with fixed-plan x-pass kernels, and their full tile (row pairs)
bool xfixed_size(int n)
{
    return n == 128 || n == 192 || n == 256 || n == 384 || n == 512 || n == 1024 || n == 2048;
}

int xrows_for(size_t tsize) { return tsize == 4 ? kXRows<float>::value : kXRows<double>::value; }

}  // namespace

struct mmc_ctx {
    int precision = MMC_FP64;
    int count = 0;
    size_t tsize = 8;                 // sizeof(T)
    std::vector<mmc_params> params;
    std::vector<Inst> inst;
    Inst *dInst = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    size_t lds_pref = 64 * 1024;      // preferred LDS per FFT workgroup
    size_t lds_max = 64 * 1024;       // device limit per workgroup
    int num_cu = 256;                 // compute units of the device

    // pools
    void *d_m = nullptr, *d_h = nullptr, *d_a = nullptr, *d_b = nullptr, *d_khat = nullptr;
    double *d_K = nullptr;
    void *d_tw32 = nullptr, *d_tw64 = nullptr;
    double *d_beta = nullptr;
    uint32_t *d_ctl = nullptr;
    double *d_part = nullptr, *d_eout = nullptr;
    int *d_efirst = nullptr;
    double *d_H = nullptr;            // field output staging (largest instance)
    size_t m_elems = 0, a_elems = 0, b_elems = 0, khat_elems = 0, k_elems = 0, ka_elems = 0;
    long maxN = 0;

    List xl, zl, yl, cl, ml, mtl, tl, txl, tzl, tyl;  // FFT passes, cells, colour slots,
                                                      // Metropolis tiles, tensor passes
    List fl;                          // fused x pass + Metropolis tiles
    List zls;                         // z passes of fused sweeps (instances
                                      // whose fused tiles do not fold z in)
    List yzl;                         // y passes with the direct z convolution
    List rl;                          // whole-sweep row-pair tiles (k_row_sweep)
    bool rows = false;                // sweeps run k_row_sweep after the y pass
    bool rows_ok = false;             // every instance admits it
    void *d_m2 = nullptr;             // second state buffer of row sweeps
    bool metro_tiled = false;         // both colours per launch (k_metro_tile)
    bool metro_tiled_ok = false;      // every instance admits the tiled kernel
    bool fuse = false;                // sweeps run the fused x + Metropolis passes
    bool fuse_ok = false;             // every instance admits them
    bool zfold = false;               // fused tiles of short-z instances fold z in
    int any_pol = -1;                 // >= 0: the y pass and both fused passes are one
                                      // launch each over every instance (tile policy)
    bool any_ok = false;              // every instance admits those launches
    int z_fixed = 0;                  // common compile-time-planned z length of the
                                      // z passes (0: runtime plan)
    bool z_half = false;              // planned z passes: every instance's z rows fill
                                      // at most half its padded length (open z)
    bool concurrent = false;          // per-length launches of a pass on side streams
    std::vector<hipStream_t> side;    // side streams (default priority) and the
    hipEvent_t fork_ev = nullptr;     // events that fork them from / join them
    std::vector<hipEvent_t> join_ev;  // back into the context's stream
    // per instance (tyx, ncol, tyf): widest tiles that fit, and the tiles
    // chosen for the launch sizes (the autotuner's starting points)
    std::vector<std::array<int, 3>> tile_full, tile_default;
    int max_colors = 2;
    bool any_demag = false;
    bool tensor_ready = false;
    std::vector<double> T;

    // options
    bool use_graph = false;
    bool timing = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_sweeps = 0;
    double tensor_evals = 2e9;        // Newell evaluations per tensor-build launch

    // energy sums of the current state (8 per instance), valid until the
    // state or the tensor changes
    std::vector<double> e_sums;
    bool e_valid = false;

    struct Rec { int kind; hipEvent_t a, b; double bytes; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    long launches[kNumKinds] = {0};
    double total_ms[kNumKinds] = {0};
    double total_bytes[kNumKinds] = {0};

    std::string err;
};

namespace {

int fail(mmc_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                               \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail((c), MMC_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ---- timing ---------------------------------------------------------------
hipEvent_t next_event(mmc_ctx *c)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

void harvest(mmc_ctx *c)
{
    if (c->recs.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->launches[r.kind]++;
            c->total_ms[r.kind] += ms;
            c->total_bytes[r.kind] += r.bytes;
        }
    }
    c->recs.clear();
    c->ev_used = 0;
}

// Bytes of a launch that covers every instance (unsegmented lists).
double all_bytes(const mmc_ctx *c, int kind);

template <typename F>
hipError_t timed(mmc_ctx *c, int kind, F &&f, double bytes = -1.0)
{
    if (!c->timing) return f();
    if (c->ev_used + 2 > 4096) harvest(c);
    hipEvent_t a = next_event(c), b = next_event(c);
    if (!a || !b) return f();
    set_launch_events(a, b);
    hipError_t e = f();
    set_launch_events(nullptr, nullptr);
    c->recs.push_back({kind, a, b, bytes >= 0 ? bytes : all_bytes(c, kind)});
    return e;
}

// ---- tiles ----------------------------------------------------------------
// Largest unit count <= want whose two ping-pong LDS buffers fit the
// preferred per-workgroup budget (two workgroups per CU), else one unit if
// it fits the device maximum, else 0.
int fit_units(int want, size_t bytes_per_unit, size_t pref, size_t maxb)
{
    int u = want;
    while (u > 1 && (size_t)u * bytes_per_unit > pref) --u;
    if ((size_t)u * bytes_per_unit <= maxb) return u;
    return 0;
}

// The FFT kernels also hold their twiddle table (n complex) in LDS.
int fit_tw(int want, size_t per_unit, size_t tw, size_t pref, size_t maxb)
{
    if (tw >= maxb) return 0;
    return fit_units(want, per_unit, pref > tw ? pref - tw : 0, maxb - tw);
}

bool choose_tiles(const mmc_ctx *c, Inst &in, size_t pref, size_t maxb)
{
    const size_t cs = 2 * c->tsize;
    const int tymax = c->tsize == 4 ? 16 : 8;
    int pairs = fit_tw(tymax / 2, 2 * (size_t)in.Px * cs, in.Px * cs, pref, maxb);
    if (xfixed_size(in.Px) && (size_t)(xrows_for(c->tsize) + 1) * in.Px * cs <= maxb)
        pairs = xrows_for(c->tsize);
    in.tyx = 2 * pairs;
    // z tiles: up to 64 (FP32) / 32 (FP64) y columns, 256-512 bytes of
    // contiguous y per z row
    in.tyz = in.Pz <= kZDirect ? 1 : fit_tw(4 * tymax, 2 * (size_t)in.Pz * cs, in.Pz * cs, pref, maxb);
    in.ncol = fixed_fft_size(in.Py) ? fixed_cols(in.Py)
                                    : fit_tw(4, 2 * 3 * (size_t)in.Py * cs, in.Py * cs, pref, maxb);
    if (fixed_fft_size(in.Py) && (3 * (size_t)fixed_cols(in.Py) + 1) * in.Py * cs > maxb) in.ncol = 0;
    in.tyx64 = 2 * fit_tw(4, 2 * (size_t)in.Px * 16, in.Px * 16, pref, maxb);
    in.tyz64 = in.Pz <= kZDirect ? 1 : fit_tw(8, 2 * (size_t)in.Pz * 16, in.Pz * 16, pref, maxb);
    const bool yt = 3 * (size_t)in.Py * 16 <= maxb;
    return pairs && in.tyz && in.ncol && in.tyx64 && in.tyz64 && yt;
}

// LDS of a Metropolis tile: m with a 2-cell halo, h with a 1-cell halo.
size_t mtile_lds(const mmc_ctx *c, int tx, int ty, int nz)
{
    return 3 * (size_t)nz * ((size_t)(tx + 4) * (ty + 4) + (size_t)(tx + 2) * (ty + 2)) * c->tsize;
}

// Widest Metropolis tile (checkerboard instances only) whose LDS fits the
// preferred budget; (0,0) when none does.
void choose_mtile(const mmc_ctx *c, Inst &in)
{
    static const int cand[][2] = {{64, 8}, {64, 4}, {32, 8}, {32, 4}, {32, 2}};
    in.mtx = in.mty = 0;
    if (in.ncolors != 2) return;
    for (auto &t : cand)
        if (mtile_lds(c, t[0], t[1], in.nz) <= c->lds_pref) {
            in.mtx = t[0];
            in.mty = t[1];
            return;
        }
}

// FP32 copies of an instance's physics constants (Inst::pf).
void round_phys(Inst &in)
{
    for (int a = 0; a < 3; ++a) in.pf.lam[a] = (float)in.lam[a];
    for (int a = 0; a < 2; ++a) in.pf.dmi[a] = (float)in.dmi[a];
    in.pf.kap = (float)in.kap;
    for (int a = 0; a < 3; ++a) in.pf.u[a] = (float)in.u[a];
    for (int a = 0; a < 3; ++a) in.pf.b[a] = (float)in.b[a];
    for (int k = 0; k < 6; ++k) in.pf.K0[k] = (float)in.K0[k];
}

int setup_inst(mmc_ctx *c, int idx, const mmc_params &p, Inst &in)
{
    if (p.nx < 1 || p.ny < 1 || p.nz < 1)
        return fail(c, MMC_EINVAL, "instance %d: grid must be at least 1x1x1", idx);
    if (!(p.dx > 0 && p.dy > 0 && p.dz > 0))
        return fail(c, MMC_EINVAL, "instance %d: cell size must be positive", idx);
    if (!(p.ms > 0))
        return fail(c, MMC_EINVAL, "instance %d: ms must be positive", idx);
    // kernels index the three components of m as i, N+i, 2N+i in 32 bits,
    // with cell indices x + nx*(y + ny*z) from 24-bit products
    if ((long)p.nx * p.ny * p.nz > ((1L << 32) - 1) / 3)
        return fail(c, MMC_EINVAL, "instance %d: more than (2^32-1)/3 cells", idx);
    if (p.nx >= (1 << 24) || (long)p.ny * p.nz >= (1L << 24))
        return fail(c, MMC_EINVAL, "instance %d: nx and ny*nz must be below 2^24", idx);
    std::memset(&in, 0, sizeof(in));
    in.nx = p.nx; in.ny = p.ny; in.nz = p.nz;
    in.pbc[0] = p.pbc_x != 0; in.pbc[1] = p.pbc_y != 0; in.pbc[2] = p.pbc_z != 0;
    in.img[0] = std::max(p.pbc_x, 0);           // src/main.cu:475-477
    in.img[1] = std::max(p.pbc_y, 0);
    in.img[2] = std::max(p.pbc_z, 0);
    in.Px = padded_len(p.nx, p.pbc_x);
    in.Py = padded_len(p.ny, p.pbc_y);
    in.Pz = padded_len(p.nz, p.pbc_z);
    in.Hx = in.Px / 2 + 1;
    in.N = (long)p.nx * p.ny * p.nz;
    in.NP = (long)in.Px * in.Py * in.Pz;
    if (!make_plan(in.Px, in.plx) || !make_plan(in.Py, in.ply) || !make_plan(in.Pz, in.plz))
        return fail(c, MMC_EUNSUP,
                    "instance %d: padded lengths (%d,%d,%d) must factor into 2,3,5,7 "
                    "(a periodic axis keeps its own length)", idx, in.Px, in.Py, in.Pz);
    if (!choose_tiles(c, in, c->lds_pref, c->lds_max))
        return fail(c, MMC_EUNSUP,
                    "instance %d: padded grid (%d,%d,%d) does not fit the %zu-byte LDS "
                    "per workgroup", idx, in.Px, in.Py, in.Pz, c->lds_max);
    in.ncolors = 2;
    for (int a = 0; a < 3; ++a) {
        int n = a == 0 ? p.nx : (a == 1 ? p.ny : p.nz);
        if (in.pbc[a] && n >= 3 && (n & 1)) in.ncolors = 3;
    }
    in.cal_demag = p.cal_demag ? 1 : 0;
    // films of a few layers: the y pass sums the z convolution directly (no
    // z transforms of the magnetisation spectrum or of the tensor)
    in.zd = in.cal_demag && in.nz > 1 && in.nz <= kZConvMax && in.Pz <= kZDirect &&
            fixed_fft_size(in.Py) &&
            (3 * (size_t)in.nz * in.Py + in.Py / 4) * 2 * c->tsize <= c->lds_max;
    for (int c6 = 0; c6 < 6; ++c6) in.kmode[c6] = 2;   // until the tensor is built
    choose_mtile(c, in);
    // physics in reduced units (fields per Ms, energies per mu0 Ms^2 V)
    const double mu0 = 4e-7 * kPiTrunc;          // src/main.cu:244
    const double ms = p.ms;
    in.d[0] = p.dx; in.d[1] = p.dy; in.d[2] = p.dz;
    const double dd[3] = {p.dx, p.dy, p.dz};
    for (int a = 0; a < 3; ++a) in.lam[a] = 2.0 * p.aex / (mu0 * ms * ms * dd[a] * dd[a]);
    for (int a = 0; a < 2; ++a) in.dmi[a] = p.dind / (mu0 * ms * ms * dd[a]);
    in.kap = p.ku1 / (mu0 * ms * ms);
    const double ul = std::sqrt(p.anis_ux * p.anis_ux + p.anis_uy * p.anis_uy + p.anis_uz * p.anis_uz);
    if (ul > 0) { in.u[0] = p.anis_ux / ul; in.u[1] = p.anis_uy / ul; in.u[2] = p.anis_uz / ul; }
    in.b[0] = p.bext_x / (mu0 * ms);
    in.b[1] = p.bext_y / (mu0 * ms);
    in.b[2] = p.bext_z / (mu0 * ms);
    in.ms = ms;
    in.e_unit = mu0 * ms * ms * p.dx * p.dy * p.dz;
    in.key[0] = (uint32_t)p.seed;
    in.key[1] = (uint32_t)(p.seed >> 32);
    round_phys(in);
    return MMC_OK;
}

int upload_list(mmc_ctx *c, List &l)
{
    if (l.dev) { (void)hipFree(l.dev); l.dev = nullptr; }
    if (l.host.empty()) return MMC_OK;
    HIPCHK(c, hipMalloc(&l.dev, sizeof(Work) * l.host.size()));
    HIPCHK(c, hipMemcpy(l.dev, l.host.data(), sizeof(Work) * l.host.size(), hipMemcpyHostToDevice));
    return MMC_OK;
}

// Algorithmic HBM bytes of one pass of kernel `k` over instance `in`
// (b = element size): each array read or written once.
double inst_bytes(const mmc_ctx *c, const Inst &in, int k)
{
    const double b = (double)c->tsize, cb = 2.0 * b;
    const double N = (double)in.N;
    const double spec_x = 3.0 * in.Hx * in.nz * in.ny * cb;   // A
    const double spec_z = 3.0 * in.Hx * in.Pz * in.ny * cb;   // B
    // zero planes are not read, real planes are half size; the direct z
    // convolution reads the slots of offsets -(nz-1)..nz-1 only
    const int kslots = in.zd ? std::min(in.Pz, 2 * in.nz - 1) : in.Pz;
    double khat = 0;
    for (int c6 = 0; c6 < 6; ++c6) khat += in.kmode[c6] * (double)in.Hx * kslots * in.Py * b;
    switch (k) {
    case K_XFWD: case K_XINV: return in.cal_demag ? 3 * N * b + spec_x : 0.0;
    case K_ZFWD: case K_ZINV: return (in.cal_demag && in.Pz > 1 && !in.zd) ? spec_x + spec_z : 0.0;
    case K_YMUL: return in.cal_demag && !in.zd ? 2 * spec_z + khat : 0.0;
    case K_YZC: return in.cal_demag && in.zd ? 2 * spec_x + khat : 0.0;
    case K_METRO: {
        // one colour pass: all of m read (cell + neighbours), h of the
        // colour's cells and their updated m
        const double K = in.ncolors;
        return 3 * N * b + (in.cal_demag ? 3 * N * b / K : 0.0) + 3 * N * b / K;
    }
    case K_ENERGY: return 3 * N * b + (in.cal_demag ? 3 * N * b : 0.0);
    // x pass (spectrum) + one checkerboard colour: m read (cell and
    // neighbours), half of m written, half of h written (xinv_metro) or
    // read (metro_xfwd)
    case K_XINVM: case K_MXFWD:
        return in.cal_demag ? (in.zf ? spec_z : spec_x) + 6 * N * b : 0.0;
    // whole sweep of row pairs (all layers): spectrum read and written, m
    // read (cells and neighbours) and written
    case K_ROWS: return in.cal_demag ? 2 * spec_x + 6 * N * b : 0.0;
    case K_TENSOR: return 6.0 * in.NP * 8;
    }
    return 0.0;
}

double all_bytes(const mmc_ctx *c, int kind)
{
    double t = 0;
    for (auto &in : c->inst) t += inst_bytes(c, in, kind);
    return t;
}


// Does instance `in` take the fixed-plan x kernels (full tile fits LDS)?
bool xfixed(const mmc_ctx *c, const Inst &in)
{
    return xfixed_size(in.Px) &&
           (size_t)(xrows_for(c->tsize) + 1) * in.Px * 2 * c->tsize <= c->lds_max;
}

// Complex rows per component of a full fused x + Metropolis tile of
// instance `in` (0 when the instance cannot take the fused passes: no demag,
// three colours, an x length without a fixed plan, or LDS too small for one
// row pair).  With the "zfold" option, sets in.zf when a tile can span every
// layer and fold the z transform in (Pz <= kZDirect, the layers divide the
// tile, and the instance has z transforms: not zd).
int fused_full(const mmc_ctx *c, Inst &in)
{
    in.zf = 0;
    if (!in.cal_demag || in.ncolors != 2 || !xfixed_size(in.Px)) return 0;
    const size_t cs = 2 * c->tsize;
    const int r = fused_rows(c->tsize == 4 ? fused_rows_base<float>() : fused_rows_base<double>(),
                             in.Px);
    const bool fits = (3 * (size_t)r + 1) * in.Px * cs <= c->lds_max;
    if (c->zfold && fits && in.Pz > 1 && in.Pz <= kZDirect && !in.zd && r % in.nz == 0) {
        in.zf = 1;
        return r;
    }
    if (fits) return r;
    return (4 * (size_t)in.Px * cs <= c->lds_max) ? 1 : 0;
}

// LDS of a whole-sweep row-pair tile: three row pairs of every layer and
// component, and the twiddles.
size_t row_lds(const mmc_ctx *c, const Inst &in)
{
    return (9 * (size_t)in.nz + 1) * in.Px * 2 * c->tsize;
}

// Can instance `in` take whole-sweep row-pair tiles (k_row_sweep)?  A
// checkerboard film of one or two layers with demag, a fused-kernel x
// length and room for its tile in LDS.
bool row_ok(const mmc_ctx *c, const Inst &in)
{
    return in.cal_demag && in.nz <= 2 && in.ncolors == 2 && xfixed_size(in.Px) &&
           row_lds(c, in) <= c->lds_max;
}

// Workgroups are dispatched to the 8 XCDs round-robin by index, and each
// XCD has its own L2.  x-pass items that touch the same 128-byte lines of a
// y-contiguous spectrum (consecutive row tiles: `share` items per line) are
// re-indexed so that a group of `share` items gets indices ≡ g (mod 8), one
// after another: they run on one XCD at about the same time, and a line
// comes from HBM once instead of once per item.
constexpr int kXcds = 8;

void xcd_group(std::vector<Work> &items, int share)
{
    if (share <= 1) return;
    const long n = (long)items.size();
    const long block = (long)kXcds * share;
    const long full = n / block * block;
    std::vector<Work> out(items);
    for (long j = 0; j < full; ++j) {
        const long g = j / share, l = j % share;
        const long sb = g / kXcds, gx = g % kXcds;
        out[sb * block + l * kXcds + gx] = items[j];
    }
    items.swap(out);
}

// Items of an x-pass tile of `rows` real rows that share one 128-byte line
// of the spectrum (complex element size cs).
int line_share(int rows, size_t cs)
{
    const int chunk = (int)(rows * cs);
    return chunk >= 128 ? 1 : 128 / chunk;
}

// One segment (one launch of an "any" kernel) holding every segment's
// items, the longest transforms first so that the short items fill the
// launch's tail; each segment's XCD grouping survives the concatenation
// (a shift of the start index shifts every item of a group alike).
void merge_segments(List &l, int pol)
{
    if (l.seg.empty()) return;
    std::vector<Segment> segs = l.seg;
    std::stable_sort(segs.begin(), segs.end(),
                     [](const Segment &a, const Segment &b) { return a.fixed_n > b.fixed_n; });
    Segment m;
    m.fixed_n = kAnyLen;
    m.rows = pol;
    std::vector<Work> items;
    for (const Segment &sg : segs) {
        items.insert(items.end(), l.host.begin() + sg.begin, l.host.begin() + sg.begin + sg.count);
        m.lds = std::max(m.lds, sg.lds);
        m.bytes += sg.bytes;
        m.half = m.half && sg.half;
    }
    m.count = (int)items.size();
    l.host.swap(items);
    l.seg.assign(1, m);
}

void build_lists(mmc_ctx *c)
{
    const size_t cs = 2 * c->tsize;   // complex element size
    auto reset = [&](List &l) { l.host.clear(); l.first.assign(1, 0); l.lds = 0; l.seg.clear(); };
    std::map<std::pair<int, int>, Segment> yclass;   // (fixed length or 0, columns) -> items
    std::map<std::pair<int, int>, std::vector<Work>> yitems;
    std::map<std::pair<int, int>, Segment> xclass;   // (fixed length, rows) -> items
    std::map<std::pair<int, int>, std::vector<Work>> xitems;
    std::map<std::pair<int, int>, Segment> rclass;   // (x length, layers) -> row-sweep items
    std::map<std::pair<int, int>, std::vector<Work>> ritems;
    c->rows_ok = !c->inst.empty();
    for (auto &in : c->inst) c->rows_ok = c->rows_ok && row_ok(c, in);
    std::map<std::pair<int, int>, Segment> fclass;   // (x length, row pairs) -> items
    std::map<std::pair<int, int>, std::vector<Work>> fitems;
    for (List *l : {&c->xl, &c->zl, &c->yl, &c->cl, &c->ml, &c->mtl, &c->tl, &c->txl, &c->tzl, &c->tyl, &c->fl, &c->zls, &c->yzl, &c->rl}) reset(*l);
    std::map<std::pair<int, int>, Segment> zdclass;  // (y length, layers) -> items
    std::map<std::pair<int, int>, std::vector<Work>> zditems;
    // z passes take the compile-time plan when every instance that has them
    // shares one planned length
    c->z_fixed = 0;
    for (const Inst &in : c->inst) {
        if (!in.cal_demag || in.Pz <= kZDirect || in.zd) continue;
        if (c->z_fixed == 0 && fixed_z_size(in.Pz)) c->z_fixed = in.Pz;
        else if (c->z_fixed != in.Pz) { c->z_fixed = -1; break; }
    }
    for (const Inst &in : c->inst)
        if (in.cal_demag && in.Pz > 1 && in.Pz <= kZDirect && !in.zd) c->z_fixed = -1;
    if (c->z_fixed < 0) c->z_fixed = 0;
    c->z_half = c->z_fixed > 0;
    for (const Inst &in : c->inst)
        if (in.cal_demag && in.Pz > kZDirect && !in.zd) c->z_half = c->z_half && 2 * in.nz <= in.Pz;
    for (int i = 0; i < c->count; ++i) {
        const Inst &in = c->inst[i];
        const bool zdirect = in.Pz <= kZDirect;
        const int zstep = zdirect ? kThreads : (c->z_fixed ? z_cols(c->tsize) : in.tyz);
        const int zstep64 = zdirect ? kThreads : in.tyz64;
        if (in.cal_demag) {
            const bool xf = xfixed(c, in);
            const std::pair<int, int> xk = xf ? std::make_pair(in.Px, in.tyx / 2) : std::make_pair(0, 0);
            for (int comp = 0; comp < 3; ++comp)
                for (int z = 0; z < in.nz; ++z)
                    for (int y0 = 0; y0 < in.ny; y0 += in.tyx) xitems[xk].push_back({i, comp, z, y0});
            const size_t xlds = xf ? ((size_t)in.tyx / 2 + 1) * in.Px * cs
                                   : (2 * (size_t)(in.tyx / 2) + 1) * in.Px * cs;
            Segment &xs = xclass[xk];
            xs.fixed_n = xk.first;
            xs.rows = xk.second;
            xs.lds = std::max(xs.lds, xlds);
            xs.bytes += inst_bytes(c, in, K_XFWD);
            c->xl.lds = std::max(c->xl.lds, xlds);
            if (in.zd) {
                const std::pair<int, int> zk(in.Py, in.nz);
                for (int kx = 0; kx < in.Hx; ++kx) zditems[zk].push_back({i, kx, 0, 0});
                Segment &zs = zdclass[zk];
                zs.fixed_n = in.Py;
                zs.rows = in.nz;
                zs.lds = std::max(zs.lds, (3 * (size_t)in.nz * in.Py + in.Py / 4) * cs);
                zs.bytes += inst_bytes(c, in, K_YZC);
            } else if (in.Pz > 1) {
                for (int comp = 0; comp < 3; ++comp)
                    for (int kx = 0; kx < in.Hx; ++kx)
                        for (int y0 = 0; y0 < in.ny; y0 += zstep) {
                            c->zl.host.push_back({i, comp, kx, y0});
                            if (!in.zf) c->zls.host.push_back({i, comp, kx, y0});
                        }
                if (!zdirect)
                    c->zl.lds = std::max(c->zl.lds, c->z_fixed ? ((size_t)z_cols(c->tsize) + 1) * in.Pz * cs
                                                               : (2 * (size_t)in.tyz + 1) * in.Pz * cs);
            }
            const int fn = fixed_fft_size(in.Py) ? in.Py : 0;
            const std::pair<int, int> yk(fn, fn ? in.ncol : 0);
            for (int kz = 0; kz < in.Pz && !in.zd; ++kz)
                for (int kx0 = 0; kx0 < in.Hx; kx0 += in.ncol)
                    yitems[yk].push_back({i, kx0, kz, std::min(in.ncol, in.Hx - kx0)});
            if (!in.zd) {
                const size_t ylds = fn ? (3 * (size_t)in.ncol + 1) * in.Py * cs
                                       : (2 * (size_t)3 * in.ncol + 1) * in.Py * cs;
                Segment &sg = yclass[yk];
                sg.fixed_n = fn;
                sg.rows = yk.second;
                sg.half = sg.half && 2 * in.ny <= in.Py;
                sg.lds = std::max(sg.lds, ylds);
                sg.bytes += inst_bytes(c, in, K_YMUL);
                c->yl.lds = std::max(c->yl.lds, ylds);
            }
            // tensor passes (FP64)
            for (int comp = 0; comp < 6; ++comp)
                for (int z = 0; z < in.Pz; ++z)
                    for (int y0 = 0; y0 < in.Py; y0 += in.tyx64) c->txl.host.push_back({i, comp, z, y0});
            c->txl.lds = std::max(c->txl.lds, (2 * (size_t)(in.tyx64 / 2) + 1) * in.Px * 16);
            if (in.Pz > 1 && !in.zd) {   // zd: the tensor's z slices stay untransformed
                for (int comp = 0; comp < 6; ++comp)
                    for (int kx = 0; kx < in.Hx; ++kx)
                        for (int y0 = 0; y0 < in.Py; y0 += zstep64) c->tzl.host.push_back({i, comp, kx, y0});
                if (!zdirect) c->tzl.lds = std::max(c->tzl.lds, (2 * (size_t)in.tyz64 + 1) * in.Pz * 16);
            }
            for (int comp = 0; comp < 6; ++comp)
                for (int kx = 0; kx < in.Hx; ++kx)
                    for (int kz = 0; kz < in.Pz; ++kz) c->tyl.host.push_back({i, comp, kx, kz});
            c->tyl.lds = std::max(c->tyl.lds, 3 * (size_t)in.Py * 16);
            for (long s = 0; s < in.NP; s += kThreads) c->tl.host.push_back({i, (int)s, 0, 0});
        }
        for (long s = 0; s < in.N; s += kThreads) c->cl.host.push_back({i, (int)s, 0, 0});
        // Metropolis: with two colours a pass visits ceil(nx/2) slots per row
        const long slots = in.ncolors == 2 ? (long)((in.nx + 1) / 2) * in.ny * in.nz : in.N;
        for (long s = 0; s < slots; s += kThreads) c->ml.host.push_back({i, (int)s, 0, 0});
        if (in.mtx) {
            for (int y0 = 0; y0 < in.ny; y0 += in.mty)
                for (int x0 = 0; x0 < in.nx; x0 += in.mtx) c->mtl.host.push_back({i, x0, y0, 0});
            c->mtl.lds = std::max(c->mtl.lds, mtile_lds(c, in.mtx, in.mty, in.nz));
        }
        if (c->rows_ok) {
            const std::pair<int, int> rk(in.Px, in.nz);
            for (int y0 = 0; y0 < in.ny; y0 += 2) ritems[rk].push_back({i, 0, 0, y0});
            Segment &rs = rclass[rk];
            rs.fixed_n = in.Px;
            rs.rows = in.nz;
            rs.lds = std::max(rs.lds, row_lds(c, in));
            rs.bytes += inst_bytes(c, in, K_ROWS);
        }
        if (in.tyf) {
            const std::pair<int, int> fk(in.Px, in.tyf);
            if (in.zf) {
                const int rows = 2 * (in.tyf / in.nz);
                for (int y0 = 0; y0 < in.ny; y0 += rows) fitems[fk].push_back({i, 0, 0, y0});
            } else {
                // the layers of one row pair next to each other: a cell's z
                // neighbours sit in the other layers' tiles, which then run on
                // the same XCD (xcd_group below) and find those rows in its L2
                for (int y0 = 0; y0 < in.ny; y0 += 2 * in.tyf)
                    for (int z = 0; z < in.nz; ++z) fitems[fk].push_back({i, 0, z, y0});
            }
            Segment &fs = fclass[fk];
            fs.fixed_n = in.Px;
            fs.rows = in.tyf;
            fs.lds = std::max(fs.lds, (3 * (size_t)in.tyf + 1) * in.Px * cs);
            fs.bytes += inst_bytes(c, in, K_XINVM);
        }
        for (List *l : {&c->xl, &c->zl, &c->yl, &c->cl, &c->ml, &c->mtl, &c->tl, &c->txl, &c->tzl, &c->tyl, &c->fl, &c->zls, &c->yzl, &c->rl})
            l->first.push_back(l->n());
    }
    c->zls.lds = c->zl.lds;
    for (auto &kv : fitems) {
        const Inst &i0 = c->inst[kv.second.front().inst];
        xcd_group(kv.second, i0.zf ? line_share(2 * kv.first.second / i0.nz, cs)
                                   : line_share(2 * kv.first.second, cs) * i0.nz);
        Segment sg = fclass[kv.first];
        sg.begin = c->fl.n();
        sg.count = (int)kv.second.size();
        c->fl.host.insert(c->fl.host.end(), kv.second.begin(), kv.second.end());
        c->fl.seg.push_back(sg);
    }
    for (auto &kv : ritems) {
        xcd_group(kv.second, line_share(2, cs));
        Segment sg = rclass[kv.first];
        sg.begin = c->rl.n();
        sg.count = (int)kv.second.size();
        c->rl.host.insert(c->rl.host.end(), kv.second.begin(), kv.second.end());
        c->rl.seg.push_back(sg);
    }
    // x and y passes: one contiguous segment (and launch) per kernel variant
    for (auto &kv : xitems) {
        xcd_group(kv.second, line_share(2 * std::max(1, kv.first.second), cs));
        Segment sg = xclass[kv.first];
        sg.begin = c->xl.n();
        sg.count = (int)kv.second.size();
        c->xl.host.insert(c->xl.host.end(), kv.second.begin(), kv.second.end());
        c->xl.seg.push_back(sg);
    }
    for (auto &kv : zditems) {
        // 8 consecutive x frequencies on one XCD: their pair-major stores
        // (16 bytes each) fill 128-byte lines together
        xcd_group(kv.second, 8);
        Segment sg = zdclass[kv.first];
        sg.begin = c->yzl.n();
        sg.count = (int)kv.second.size();
        c->yzl.host.insert(c->yzl.host.end(), kv.second.begin(), kv.second.end());
        c->yzl.seg.push_back(sg);
    }
    for (auto &kv : yitems) {
        // pair-major stores of consecutive x frequencies fill 128-byte lines
        // together when their items run on one XCD
        if (kv.first.second > 0 && kv.first.second < 8) xcd_group(kv.second, 8 / kv.first.second);
        Segment sg = yclass[kv.first];
        sg.begin = c->yl.n();
        sg.count = (int)kv.second.size();
        c->yl.host.insert(c->yl.host.end(), kv.second.begin(), kv.second.end());
        c->yl.seg.push_back(sg);
    }
    // the "any" kernels take their tile shape from the policy: a tile option
    // set since (fuse_rows, ncol, zfold) returns the batch to per-length launches
    for (auto &in : c->inst)
        if (c->any_pol >= 0 && (in.ncol != any_tile(c->any_pol, in.Py) ||
                                in.tyf != any_tile(c->any_pol, in.Px) || in.zf))
            c->any_pol = -1;
    if (c->any_pol >= 0) {
        merge_segments(c->yl, c->any_pol);
        merge_segments(c->fl, c->any_pol);
    }
}

template <typename T>
void twiddles(int n, std::vector<cpx<T>> &out)
{
    for (int t = 0; t < n; ++t) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
        out.push_back({(T)cosl(a), (T)sinl(a)});
    }
}

// Largest x-pass row tile (even, >= 2) and y-pass column count that fit the
// LDS limits of instance `in`.
int max_tyx(const mmc_ctx *c, const Inst &in)
{
    const size_t cs = 2 * c->tsize;
    return 2 * fit_tw(64, 2 * (size_t)in.Px * cs, in.Px * cs, c->lds_max, c->lds_max);
}

int max_ncol(const mmc_ctx *c, const Inst &in)
{
    if (fixed_fft_size(in.Py)) return fixed_cols(in.Py);
    const size_t cs = 2 * c->tsize;
    return fit_tw(64, 2 * 3 * (size_t)in.Py * cs, in.Py * cs, c->lds_max, c->lds_max);
}

// Shrink the tiles of every group of instances that one launch serves
// (same kernel variant) until the group has `target` work items or its
// tiles are minimal.  key(in) < 0: the instance is not in this pass.
template <typename Key, typename Items, typename Shrink>
void adapt_groups(std::vector<Inst> &inst, long target, Key key, Items items, Shrink shrink)
{
    std::map<int, std::vector<Inst *>> groups;
    for (auto &in : inst) {
        const int k = key(in);
        if (k >= 0) groups[k].push_back(&in);
    }
    for (auto &g : groups)
        for (bool changed = true; changed;) {
            long n = 0;
            for (const Inst *p : g.second) n += items(*p);
            if (n >= target) break;
            changed = false;
            for (Inst *p : g.second) changed |= shrink(*p);
        }
}

// Small grids cannot fill the chip with wide tiles: per launch (each
// fixed-plan length has its own), shrink x-pass row tiles, y-pass column
// groups, Metropolis tiles and fused tiles until the launch has about two
// workgroups per compute unit (or the tiles are minimal).
void adapt_tiles(mmc_ctx *c)
{
    const long target = 2L * c->num_cu;
    adapt_groups(
        c->inst, target,
        [&](const Inst &in) { return in.cal_demag ? (xfixed(c, in) ? in.Px : 0) : -1; },
        [](const Inst &in) { return 3L * in.nz * ((in.ny + in.tyx - 1) / in.tyx); },
        [&](Inst &in) {
            if (in.tyx <= 2) return false;
            // fixed-plan x tiles come in two shapes: full, or one row pair
            in.tyx = xfixed(c, in) ? 2 : std::max(2, (in.tyx / 2) & ~1);
            return true;
        });
    adapt_groups(
        c->inst, target,
        [](const Inst &in) { return in.cal_demag ? (fixed_fft_size(in.Py) ? in.Py : 0) : -1; },
        [](const Inst &in) { return (long)in.Pz * ((in.Hx + in.ncol - 1) / in.ncol); },
        [](Inst &in) {
            if (in.ncol <= 1) return false;
            // fixed-plan y items come in two widths: fixed_cols(Py), or one
            in.ncol = fixed_fft_size(in.Py) ? 1 : std::max(1, in.ncol / 2);
            return true;
        });
    adapt_groups(
        c->inst, target, [](const Inst &in) { return in.mtx ? 0 : -1; },
        [](const Inst &in) {
            return (long)((in.nx + in.mtx - 1) / in.mtx) * ((in.ny + in.mty - 1) / in.mty);
        },
        [](Inst &in) {
            if (in.mty > 2) { in.mty /= 2; return true; }
            if (in.mtx > 16) { in.mtx /= 2; return true; }
            return false;
        });
    c->metro_tiled_ok = true;
    for (auto &in : c->inst) if (!in.mtx) c->metro_tiled_ok = false;
    c->metro_tiled = false;           // per-colour launches unless asked
    // fused x + Metropolis tiles: full, or one row pair for small grids
    c->fuse_ok = !c->inst.empty();
    for (auto &in : c->inst) {
        in.tyf = fused_full(c, in);
        if (!in.tyf) c->fuse_ok = false;
    }
    if (!c->fuse_ok)
        for (auto &in : c->inst) in.tyf = 0;
    adapt_groups(
        c->inst, target, [](const Inst &in) { return in.tyf ? in.Px : -1; },
        [](const Inst &in) { return (long)in.nz * ((in.ny + 2 * in.tyf - 1) / (2 * in.tyf)); },
        [](Inst &in) {
            if (in.tyf <= 1 || in.zf) return false;   // z-folding tiles keep all layers
            in.tyf = 1;
            return true;
        });
    c->fuse = c->fuse_ok;
    c->any_ok = c->fuse_ok;
    for (auto &in : c->inst)
        c->any_ok = c->any_ok && any_x_len(in.Px) && any_y_len(in.Py) && !in.zd;
}

int alloc_all(mmc_ctx *c)
{
    const size_t ts = c->tsize, cs = 2 * ts;
    size_t m = 0, a = 0, b = 0, kh = 0, k = 0, ka = 0, tw = 0;
    std::vector<cpx<float>> tw32;
    std::vector<cpx<double>> tw64;
    // A and B share one pool; an instance with Pz == 1 has no z pass and its
    // B region is its A region
    for (auto &in : c->inst) {
        in.m_off = (long)m;
        in.h_off = (long)m;
        m += 3 * in.N;
        c->maxN = std::max(c->maxN, in.N);
        if (in.cal_demag) {
            in.a_off = (long)a;
            a += (size_t)3 * in.Hx * in.nz * in.ny;
            if (in.Pz > 1) {
                in.b_off = (long)b;           // rebased past the A regions below
                b += (size_t)3 * in.Hx * in.Pz * in.ny;
            } else {
                in.b_off = -1;
            }
            // a film without z passes whose y pass has a compile-time plan
            // gets a region of its own for the pair-major spectrum
            in.pm_off = -1;
            if (in.Pz == 1 && !in.zd && (in.ny % 2) == 0 && fixed_fft_size(in.Py)) {
                b = (b + 1) & ~(size_t)1;
                in.pm_off = (long)b;
                b += (size_t)3 * in.Hx * in.ny;
            }
            in.khat_off = (long)kh;
            kh += (size_t)6 * in.Hx * in.Pz * in.Py;
            in.k_off = (long)k;
            k += (size_t)6 * in.NP;
            in.ka_off = (long)ka;
            ka += (size_t)6 * in.Hx * in.Pz * in.Py;
            in.twx = (long)tw; twiddles(in.Px, tw32); twiddles(in.Px, tw64); tw += in.Px;
            in.twy = (long)tw; twiddles(in.Py, tw32); twiddles(in.Py, tw64); tw += in.Py;
            in.twz = (long)tw; twiddles(in.Pz, tw32); twiddles(in.Pz, tw64); tw += in.Pz;
        }
    }
    a = (a + 1) & ~(size_t)1;   // B regions start 16-byte aligned (FP32)
    for (auto &in : c->inst) {
        if (in.cal_demag) in.b_off = in.b_off < 0 ? in.a_off : (long)a + in.b_off;
        // the B region of a zd instance (no z transforms) is free; pair-major
        // layout of [3][nz][ny/2][Hx][2] fits it when ny is even (Pz > nz)
        if (in.cal_demag && in.zd) in.pm_off = in.b_off != in.a_off ? in.b_off : -1;
        else if (in.cal_demag && in.pm_off >= 0) in.pm_off += (long)a;
        else in.pm_off = -1;
        in.a2 = in.cal_demag && (in.ny % 2) == 0 && in.pm_off >= 0 && (in.pm_off % 2) == 0;
    }
    c->m_elems = m; c->a_elems = a; c->b_elems = b; c->khat_elems = kh; c->k_elems = k;
    c->ka_elems = ka;
    HIPCHK(c, hipMalloc(&c->d_m, std::max<size_t>(1, m) * ts));
    HIPCHK(c, hipMalloc(&c->d_h, std::max<size_t>(1, m) * ts));
    HIPCHK(c, hipMemsetAsync(c->d_h, 0, std::max<size_t>(1, m) * ts, c->stream));
    if (c->rows_ok) HIPCHK(c, hipMalloc(&c->d_m2, std::max<size_t>(1, m) * ts));
    if (a + b) HIPCHK(c, hipMalloc(&c->d_a, (a + b) * cs));
    c->d_b = c->d_a;
    if (kh) HIPCHK(c, hipMalloc(&c->d_khat, kh * cs));
    if (k) {
        HIPCHK(c, hipMalloc(&c->d_K, k * sizeof(double)));
        HIPCHK(c, hipMemsetAsync(c->d_K, 0, k * sizeof(double), c->stream));
    }
    if (tw) {
        HIPCHK(c, hipMalloc(&c->d_tw32, tw * sizeof(cpx<float>)));
        HIPCHK(c, hipMalloc(&c->d_tw64, tw * sizeof(cpx<double>)));
        HIPCHK(c, hipMemcpy(c->d_tw32, tw32.data(), tw * sizeof(cpx<float>), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(c->d_tw64, tw64.data(), tw * sizeof(cpx<double>), hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMalloc(&c->d_beta, sizeof(double) * c->count));
    HIPCHK(c, hipMalloc(&c->d_ctl, sizeof(uint32_t) * 4));
    HIPCHK(c, hipMemsetAsync(c->d_ctl, 0, sizeof(uint32_t) * 4, c->stream));
    HIPCHK(c, hipMalloc(&c->d_part, sizeof(double) * 16 * std::max(1, c->cl.n())));
    HIPCHK(c, hipMalloc(&c->d_eout, sizeof(double) * 8 * c->count));
    HIPCHK(c, hipMalloc(&c->d_efirst, sizeof(int) * (c->count + 1)));
    HIPCHK(c, hipMemcpy(c->d_efirst, c->cl.first.data(), sizeof(int) * (c->count + 1),
                        hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->d_H, sizeof(double) * 3 * std::max<long>(1, c->maxN)));
    HIPCHK(c, hipMalloc(&c->dInst, sizeof(Inst) * c->count));
    HIPCHK(c, hipMemcpy(c->dInst, c->inst.data(), sizeof(Inst) * c->count, hipMemcpyHostToDevice));
    for (List *l : {&c->xl, &c->zl, &c->yl, &c->cl, &c->ml, &c->mtl, &c->tl, &c->txl, &c->tzl, &c->tyl, &c->fl, &c->zls, &c->yzl, &c->rl}) {
        int rc = upload_list(c, *l);
        if (rc) return rc;
    }
    c->T.assign(c->count, 300.0);
    std::vector<double> beta(c->count);
    for (int i = 0; i < c->count; ++i) beta[i] = c->inst[i].e_unit / (kBoltzmann * c->T[i]);
    HIPCHK(c, hipMemcpy(c->d_beta, beta.data(), sizeof(double) * c->count, hipMemcpyHostToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MMC_OK;
}

// ---- fixed scratch of mmc_energy_noalloc ------------------------------------
// The energy reduction's scratch in a device array of the library instead of
// memory a context allocates: the per-instance sums (8 doubles for each of
// up to kNoAllocMaxInst instances) and then the block partials (16 doubles
// per block of kThreads cells).  Each GPU holds its own copy until the
// process ends; every call on a GPU shares it.
constexpr size_t kScratchBytes = size_t(2) << 20;
constexpr int kNoAllocMaxInst = 1024;
constexpr size_t kScratchSums = size_t(8) * kNoAllocMaxInst * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices];   // device address of `scratch` per device
std::mutex g_scratch_mu;

// Look up the device address of `scratch` on the context's device (current
// when the context is created), once per device.  In host code the array's
// name is only a stand-in, never a device address.
int scratch_setup(mmc_ctx *c)
{
    if (c->device < 0 || c->device >= kMaxDevices)
        return fail(c, MMC_EUNSUP, "device %d: at most %d devices", c->device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return MMC_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, MMC_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[c->device] = p;
    return MMC_OK;
}

int create_common(const mmc_params *p, int count, int precision, mmc_ctx **out)
{
    if (!out) return MMC_EINVAL;
    *out = nullptr;
    if (!p || count < 1) return MMC_EINVAL;
    if (precision != MMC_FP64 && precision != MMC_FP32) return MMC_EINVAL;
    mmc_ctx *c = new mmc_ctx();
    c->precision = precision;
    c->tsize = precision == MMC_FP32 ? 4 : 8;
    c->count = count;
    c->params.assign(p, p + count);
    c->inst.resize(count);
    int rc;
    if (hipGetDevice(&c->device) != hipSuccess) { *out = c; return fail(c, MMC_EHIP, "no HIP device"); }
    if ((rc = scratch_setup(c))) { *out = c; return rc; }
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess &&
            v > 0) {
            c->lds_max = (size_t)v;
            if (c->lds_max > 64 * 1024 && set_lds_limit(c->lds_max) != hipSuccess) c->lds_max = 64 * 1024;
        }
        c->lds_pref = std::min(c->lds_pref, c->lds_max);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess &&
            v > 0)
            c->num_cu = v;
    }
    for (int i = 0; i < count; ++i) {
        if ((rc = setup_inst(c, i, p[i], c->inst[i]))) { *out = c; return rc; }
        c->max_colors = std::max(c->max_colors, c->inst[i].ncolors);
        if (c->inst[i].cal_demag) c->any_demag = true;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        *out = c;
        return fail(c, MMC_EHIP, "hipStreamCreate failed");
    }
    for (auto &in : c->inst) {
        in.tyf = fused_full(c, in);
        c->tile_full.push_back({in.tyx, in.ncol, in.tyf});
    }
    adapt_tiles(c);
    for (auto &in : c->inst) c->tile_default.push_back({in.tyx, in.ncol, in.tyf});
    build_lists(c);
    rc = alloc_all(c);
    *out = c;
    return rc;
}

// Set every instance's x-pass row tile / y-pass column group (clamped to
// what LDS allows), rebuild and upload the work lists, drop any graph.
int retile(mmc_ctx *c, int tyx, int ncol)
{
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &in : c->inst) {
        if (tyx > 0) {
            in.tyx = std::max(2, std::min(tyx & ~1, max_tyx(c, in)));
            if (xfixed(c, in)) in.tyx = in.tyx >= 2 * xrows_for(c->tsize) ? 2 * xrows_for(c->tsize) : 2;
        }
        if (ncol > 0) {
            in.ncol = std::max(1, std::min(ncol, max_ncol(c, in)));
            if (fixed_fft_size(in.Py) && in.ncol < fixed_cols(in.Py)) in.ncol = 1;
        }
    }
    build_lists(c);
    for (List *l : {&c->xl, &c->zl, &c->yl, &c->cl, &c->ml, &c->mtl, &c->tl, &c->txl, &c->tzl, &c->tyl, &c->fl, &c->zls, &c->yzl, &c->rl}) {
        int rc = upload_list(c, *l);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemcpy(c->dInst, c->inst.data(), sizeof(Inst) * c->count, hipMemcpyHostToDevice));
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    return MMC_OK;
}

// ---- demag chain -------------------------------------------------------------
template <typename T>
struct Bufs {
    const cpx<T> *tw;
    T *m, *h;
    cpx<T> *A, *B;
    const cpx<T> *Kh;
    explicit Bufs(const mmc_ctx *c)
        : tw((const cpx<T> *)(c->tsize == 4 ? c->d_tw32 : c->d_tw64)), m((T *)c->d_m),
          h((T *)c->d_h), A((cpx<T> *)c->d_a), B((cpx<T> *)c->d_b),
          Kh((const cpx<T> *)c->d_khat) {}
};

constexpr int kSideStreams = 3;   // with the context's own: four hardware queues

int ensure_side_streams(mmc_ctx *c)
{
    if (!c->side.empty()) return MMC_OK;
    HIPCHK(c, hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
    for (int k = 0; k < kSideStreams; ++k) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        HIPCHK(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        c->side.push_back(s);
        HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->join_ev.push_back(e);
    }
    return MMC_OK;
}

// Launch one kernel per segment of `l` (kernel variants serving disjoint
// instances): in order on the context's stream, or with "concurrent" spread
// over side streams forked from and joined back into it, so small
// per-length launches of a batch overlap.  Captures into graphs as
// parallel branches.
template <typename F>
int run_segments(mmc_ctx *c, const List &l, int kind, F &&launch)
{
    const size_t n = l.seg.size();
    if (!c->concurrent || n <= 1 || !c->stream) {
        for (const Segment &sg : l.seg)
            HIPCHK(c, timed(c, kind, [&] { return launch(sg, c->stream); }, sg.bytes));
        return MMC_OK;
    }
    int rc = ensure_side_streams(c);
    if (rc) return rc;
    const size_t ns = std::min(n - 1, c->side.size());
    HIPCHK(c, hipEventRecord(c->fork_ev, c->stream));
    for (size_t k = 0; k < ns; ++k) HIPCHK(c, hipStreamWaitEvent(c->side[k], c->fork_ev, 0));
    for (size_t i = 0; i < n; ++i) {
        hipStream_t s = i == 0 ? c->stream : c->side[(i - 1) % ns];
        const Segment &sg = l.seg[i];
        HIPCHK(c, timed(c, kind, [&] { return launch(sg, s); }, sg.bytes));
    }
    for (size_t k = 0; k < ns; ++k) {
        HIPCHK(c, hipEventRecord(c->join_ev[k], c->side[k]));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->join_ev[k], 0));
    }
    return MMC_OK;
}

// m -> A (x pass of every instance with demag)
template <typename T>
int pass_xfwd(mmc_ctx *c)
{
    const Bufs<T> b(c);
    return run_segments(c, c->xl, K_XFWD, [&](const Segment &sg, hipStream_t st) {
        return launch_xfwd<T>(st, c->dInst, c->xl.dev + sg.begin, sg.count, 0, sg.fixed_n, sg.rows,
                              b.m, b.A, b.tw, sg.lds);
    });
}

// z pass bytes over the instances list `l` serves (all with Pz > 1, or in
// fused sweeps those whose tiles do not fold z in)
double zbytes(const mmc_ctx *c, bool sweep)
{
    double t = 0;
    for (auto &in : c->inst)
        if (!(sweep && in.zf)) t += inst_bytes(c, in, K_ZFWD);
    return t;
}

template <typename T>
int pass_zfwd(mmc_ctx *c, bool sweep)
{
    const Bufs<T> b(c);
    const List &l = sweep ? c->zls : c->zl;
    if (l.n())
        HIPCHK(c, timed(c, K_ZFWD, [&] { return launch_zfwd<T>(c->stream, c->dInst, l.dev, l.n(), 0, c->z_fixed, b.A, b.B, b.tw, l.lds, c->z_half); },
                        zbytes(c, sweep)));
    return MMC_OK;
}

// A -> (z) -> y, Khat product, y -> (z) -> A; zd instances (a few layers)
// take one y pass over A that sums the z convolution directly.  In fused
// sweeps (`sweep`) instances whose fused tiles fold z in skip the z passes:
// their tiles read and write B directly.
template <typename T>
int pass_zyz(mmc_ctx *c, bool sweep = false, bool a2 = false)
{
    const Bufs<T> b(c);
    hipStream_t st = c->stream;
    const List &l = sweep ? c->zls : c->zl;
    int rc = pass_zfwd<T>(c, sweep);
    if (rc) return rc;
    rc = run_segments(c, c->yl, K_YMUL, [&](const Segment &sg, hipStream_t s) {
        return launch_ymul<T>(s, c->dInst, c->yl.dev + sg.begin, sg.count, sg.fixed_n, sg.rows, b.B,
                              b.Kh, b.tw, sg.lds, sg.half, a2);
    });
    if (rc) return rc;
    rc = run_segments(c, c->yzl, K_YZC, [&](const Segment &sg, hipStream_t s) {
        return launch_yzconv<T>(s, c->dInst, c->yzl.dev + sg.begin, sg.count, sg.fixed_n, sg.rows,
                                b.A, b.Kh, b.tw, sg.lds, a2);
    });
    if (rc) return rc;
    if (l.n())
        HIPCHK(c, timed(c, K_ZINV, [&] { return launch_zinv<T>(st, c->dInst, l.dev, l.n(), c->z_fixed, b.B, b.A, b.tw, l.lds, c->z_half); },
                        zbytes(c, sweep)));
    return MMC_OK;
}

// Demag field of the current state: m -> h (every instance with demag).
template <typename T>
int demag_chain(mmc_ctx *c)
{
    if (!c->any_demag) return MMC_OK;
    if (!c->tensor_ready) return fail(c, MMC_ESTATE, "demag tensor not built (call mmc_build_tensor)");
    int rc = pass_xfwd<T>(c);
    if (rc) return rc;
    rc = pass_zyz<T>(c);
    if (rc) return rc;
    const Bufs<T> b(c);
    return run_segments(c, c->xl, K_XINV, [&](const Segment &sg, hipStream_t st) {
        return launch_xinv<T>(st, c->dInst, c->xl.dev + sg.begin, sg.count, sg.fixed_n, sg.rows, b.A,
                              b.h, b.tw, sg.lds);
    });
}

bool fused_sweeps(const mmc_ctx *c) { return c->fuse && !c->metro_tiled; }

// One sweep.  Fused mode expects A (B for z-folding tiles) to hold the
// spectrum of the current state (run_sweeps prepares it) and leaves it
// holding the spectrum of the new state: z/y/z passes, then inverse x +
// colour 0, then colour 1 + forward x.
template <typename T>
int enqueue_sweep(mmc_ctx *c, uint32_t sweep_off)
{
    hipStream_t st = c->stream;
    if (fused_sweeps(c) && c->rows) {
        // sweep s of a run reads state buffer s & 1 and writes the other
        int rc = pass_zyz<T>(c, true);
        if (rc) return rc;
        const Bufs<T> b(c);
        const T *m_in = (const T *)((sweep_off & 1) ? c->d_m2 : c->d_m);
        T *m_out = (T *)((sweep_off & 1) ? c->d_m : c->d_m2);
        return run_segments(c, c->rl, K_ROWS, [&](const Segment &sg, hipStream_t s) {
            return launch_row_sweep<T>(s, c->dInst, c->rl.dev + sg.begin, sg.count, sg.fixed_n,
                                       sg.rows, b.A, m_in, m_out, c->d_beta, c->d_ctl, sweep_off,
                                       b.tw, sg.lds);
        });
    }
    if (fused_sweeps(c)) {
        // films whose y pass sums z directly hand the inverse x pass their
        // spectrum pair-major (Inst::a2)
        int rc = pass_zyz<T>(c, true, true);
        if (rc) return rc;
        const Bufs<T> b(c);
        bool zf = false;
        for (const Inst &in : c->inst) zf = zf || in.zf;
        rc = run_segments(c, c->fl, K_XINVM, [&](const Segment &sg, hipStream_t s) {
            return launch_xinv_metro<T>(s, c->dInst, c->fl.dev + sg.begin, sg.count, sg.fixed_n,
                                        sg.rows, zf, b.A, b.m, b.h, c->d_beta, c->d_ctl, sweep_off,
                                        b.tw, sg.lds);
        });
        if (rc) return rc;
        return run_segments(c, c->fl, K_MXFWD, [&](const Segment &sg, hipStream_t s) {
            return launch_metro_xfwd<T>(s, c->dInst, c->fl.dev + sg.begin, sg.count, sg.fixed_n,
                                        sg.rows, zf, b.m, b.h, b.A, c->d_beta, c->d_ctl, sweep_off,
                                        b.tw, sg.lds);
        });
    }
    int rc = demag_chain<T>(c);
    if (rc) return rc;
    if (c->metro_tiled) {
        // one launch: m read, h read and m written once per sweep
        double bytes = 0;
        for (auto &in : c->inst) bytes += (in.cal_demag ? 9.0 : 6.0) * in.N * c->tsize;
        HIPCHK(c, timed(c, K_METRO, [&] {
                   return launch_metro_tile<T>(st, c->dInst, c->mtl.dev, c->mtl.n(), (T *)c->d_m,
                                               (const T *)c->d_h, c->d_beta, c->d_ctl, sweep_off,
                                               c->any_demag ? 1 : 0, c->mtl.lds);
               }, bytes));
        return MMC_OK;
    }
    for (int col = 0; col < c->max_colors; ++col)
        HIPCHK(c, timed(c, K_METRO, [&] {
                   return launch_metro<T>(st, c->dInst, c->ml.dev, c->ml.n(), (T *)c->d_m,
                                          (const T *)c->d_h, c->d_beta, c->d_ctl, sweep_off, col,
                                          c->any_demag ? 1 : 0);
               }));
    return MMC_OK;
}

template <typename T>
int run_sweeps(mmc_ctx *c, uint32_t stage, uint32_t sweep0, uint32_t count)
{
    hipStream_t st = c->stream;
    if (count) c->e_valid = false;
    if (c->any_demag && !c->tensor_ready)
        return fail(c, MMC_ESTATE, "demag tensor not built (call mmc_build_tensor)");
    if (count && fused_sweeps(c)) {
        // x spectrum of the current state (xz spectrum for z-folding tiles)
        int rc = pass_xfwd<T>(c);
        if (!rc && c->zls.n() != c->zl.n()) rc = pass_zfwd<T>(c, false);
        if (rc) return rc;
    }
    // row sweeps alternate two state buffers: an odd run ends in the second
    const bool pingpong = fused_sweeps(c) && c->rows;
    auto settle = [&](uint32_t eager) {
        if (pingpong && (eager & 1))
            return hipMemcpyAsync(c->d_m, c->d_m2, c->m_elems * c->tsize, hipMemcpyDeviceToDevice, st);
        return hipSuccess;
    };
    if (!c->use_graph || c->timing) {
        HIPCHK(c, launch_set_ctl(st, c->d_ctl, stage, sweep0));
        for (uint32_t s = 0; s < count; ++s) {
            int rc = enqueue_sweep<T>(c, s);
            if (rc) return rc;
        }
        HIPCHK(c, settle(count));
        return MMC_OK;
    }
    int G = c->graph_sweeps > 0 ? c->graph_sweeps : 16;
    if (pingpong) G += G & 1;   // a graph of row sweeps starts and ends in the first buffer
    if (!c->gexec) {
        HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = MMC_OK;
        for (int s = 0; s < G && rc == MMC_OK; ++s) rc = enqueue_sweep<T>(c, (uint32_t)s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        HIPCHK(c, e);
        c->graph = g;
        HIPCHK(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
        c->graph_sweeps = G;
    }
    uint32_t s = 0;
    for (; s + G <= count; s += G) {
        HIPCHK(c, launch_set_ctl(st, c->d_ctl, stage, sweep0 + s));
        HIPCHK(c, hipGraphLaunch(c->gexec, st));
    }
    if (s < count) {
        HIPCHK(c, launch_set_ctl(st, c->d_ctl, stage, sweep0 + s));
        for (uint32_t k = 0; s + k < count; ++k) {
            int rc = enqueue_sweep<T>(c, k);
            if (rc) return rc;
        }
        HIPCHK(c, settle(count - s));
    }
    return MMC_OK;
}

// ---- tensor build --------------------------------------------------------
std::string cache_name(const Inst &in, const mmc_params &p)
{
    // src/main.cu:421-423 with the reference's padded sizes (2n-1 on open axes)
    const int rx = p.pbc_x ? p.nx : 2 * p.nx - 1;
    const int ry = p.pbc_y ? p.ny : 2 * p.ny - 1;
    const int rz = p.pbc_z ? p.nz : 2 * p.nz - 1;
    char buf[256];
    snprintf(buf, sizeof(buf), "cache_%d_%d_%d-%g_%g_%g-%d_%d_%d.ovf", rx, ry, rz, p.dx, p.dy,
             p.dz, p.pbc_x, p.pbc_y, p.pbc_z);
    (void)in;
    return buf;
}

// Map reference cache index (z fastest, reference padding) <-> our index.
int ref_to_ours(int iref, int n, int Pref, int P, int periodic)
{
    if (periodic) return iref;
    const int off = (iref + (Pref + 1) / 2 - 1) % Pref - (Pref + 1) / 2 + 1;  // src/demag_cu.cu:149
    return (off % P + P) % P;
}

bool load_cache(const std::string &path, const Inst &in, const mmc_params &p, std::vector<double> &K)
{
    const int Pref[3] = {p.pbc_x ? p.nx : 2 * p.nx - 1, p.pbc_y ? p.ny : 2 * p.ny - 1,
                         p.pbc_z ? p.nz : 2 * p.nz - 1};
    const long NR = (long)Pref[0] * Pref[1] * Pref[2];
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    if ((long)f.tellg() != 6 * NR * 8) return false;
    f.seekg(0);
    std::vector<double> raw(6 * NR);
    f.read(reinterpret_cast<char *>(raw.data()), 6 * NR * 8);
    if (!f) return false;
    K.assign(6 * in.NP, 0.0);
    for (int comp = 0; comp < 6; ++comp)
        for (int xr = 0; xr < Pref[0]; ++xr)
            for (int yr = 0; yr < Pref[1]; ++yr)
                for (int zr = 0; zr < Pref[2]; ++zr) {
                    const long ridx = zr + (long)yr * Pref[2] + (long)xr * Pref[2] * Pref[1];
                    const int x = ref_to_ours(xr, p.nx, Pref[0], in.Px, in.pbc[0]);
                    const int y = ref_to_ours(yr, p.ny, Pref[1], in.Py, in.pbc[1]);
                    const int z = ref_to_ours(zr, p.nz, Pref[2], in.Pz, in.pbc[2]);
                    K[comp * in.NP + x + (long)in.Px * (y + (long)in.Py * z)] = raw[comp * NR + ridx];
                }
    return true;
}

bool save_cache(const std::string &path, const Inst &in, const mmc_params &p, const std::vector<double> &K)
{
    const int Pref[3] = {p.pbc_x ? p.nx : 2 * p.nx - 1, p.pbc_y ? p.ny : 2 * p.ny - 1,
                         p.pbc_z ? p.nz : 2 * p.nz - 1};
    const long NR = (long)Pref[0] * Pref[1] * Pref[2];
    std::vector<double> raw(6 * NR);
    for (int comp = 0; comp < 6; ++comp)
        for (int xr = 0; xr < Pref[0]; ++xr)
            for (int yr = 0; yr < Pref[1]; ++yr)
                for (int zr = 0; zr < Pref[2]; ++zr) {
                    const long ridx = zr + (long)yr * Pref[2] + (long)xr * Pref[2] * Pref[1];
                    const int x = ref_to_ours(xr, p.nx, Pref[0], in.Px, in.pbc[0]);
                    const int y = ref_to_ours(yr, p.ny, Pref[1], in.Py, in.pbc[1]);
                    const int z = ref_to_ours(zr, p.nz, Pref[2], in.Pz, in.pbc[2]);
                    raw[comp * NR + ridx] = K[comp * in.NP + x + (long)in.Px * (y + (long)in.Py * z)];
                }
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char *>(raw.data()), 6 * NR * 8);
    return (bool)f;
}

// FP32 bulk mode: Khat planes whose imaginary part (or whole value) is below
// this fraction of the instance's largest |Khat| are stored real (or not at
// all).  The exact tensor of an open axis is circularly even/odd, so its
// spectrum is real; the Newell cancellation noise of distant offsets leaves
// imaginary parts of 1e-10..1e-8 of the largest value on open grids up to a
// few hundred cells (periodic image sums leave far more, and keep complex
// planes).  Dropping parts below 1e-8 changes fields by less than FP32
// rounding.  FP64 mode drops only exact zeros.
constexpr double kKhatDropTolFp32 = 1e-8;

// Tensor FFT (FP64) for all instances with demag: K -> KA -> FP64 Khat,
// then each component plane is classified zero / real / complex and stored
// in the bulk precision (round to nearest) in that form.
template <typename T>
int tensor_fft(mmc_ctx *c)
{
    hipStream_t st = c->stream;
    cpx<double> *KA = nullptr, *KH = nullptr;
    HIPCHK(c, hipMalloc(&KA, std::max<size_t>(1, c->ka_elems) * sizeof(cpx<double>)));
    hipError_t e = hipMalloc(&KH, std::max<size_t>(1, c->khat_elems) * sizeof(cpx<double>));
    if (e != hipSuccess) {
        (void)hipFree(KA);
        return fail(c, MMC_EHIP, "tensor FFT: %s", hipGetErrorString(e));
    }
    const cpx<double> *tw = (const cpx<double> *)c->d_tw64;
    int rc = MMC_OK;
    e = launch_xfwd<double>(st, c->dInst, c->txl.dev, c->txl.n(), 1, 0, 0, c->d_K, KA, tw, c->txl.lds);
    if (e == hipSuccess && c->tzl.n())
        e = launch_zfwd<double>(st, c->dInst, c->tzl.dev, c->tzl.n(), 1, 0, KA, KA, tw, c->tzl.lds, false);
    if (e == hipSuccess)
        e = launch_yfwd_tensor<double>(st, c->dInst, c->tyl.dev, c->tyl.n(), KA, KH, tw, c->tyl.lds);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    std::vector<cpx<double>> h64;
    if (e == hipSuccess) {
        h64.resize(c->khat_elems);
        e = hipMemcpy(h64.data(), KH, c->khat_elems * sizeof(cpx<double>), hipMemcpyDeviceToHost);
    }
    (void)hipFree(KA);
    (void)hipFree(KH);
    if (e != hipSuccess) return fail(c, MMC_EHIP, "tensor FFT failed: %s", hipGetErrorString(e));
    std::vector<T> packed(2 * c->khat_elems, T(0));
    for (auto &in : c->inst) {
        if (!in.cal_demag) continue;
        const long S = (long)in.Hx * in.Pz * in.Py;
        double mre[6] = {0}, mim[6] = {0}, M = 0;
        for (int c6 = 0; c6 < 6; ++c6) {
            const cpx<double> *p = h64.data() + in.khat_off + c6 * S;
            for (long i = 0; i < S; ++i) {
                mre[c6] = std::max(mre[c6], std::fabs(p[i].x));
                mim[c6] = std::max(mim[c6], std::fabs(p[i].y));
            }
            M = std::max(M, std::max(mre[c6], mim[c6]));
        }
        const double lim = (c->tsize == 4 ? kKhatDropTolFp32 : 0.0) * M;
        for (int c6 = 0; c6 < 6; ++c6) {
            in.kmode[c6] = std::max(mre[c6], mim[c6]) <= lim ? 0 : (mim[c6] <= lim ? 1 : 2);
            const cpx<double> *p = h64.data() + in.khat_off + c6 * S;
            T *o = packed.data() + 2 * (in.khat_off + c6 * S);
            if (in.kmode[c6] == 1)
                for (long i = 0; i < S; ++i) o[i] = (T)p[i].x;
            else if (in.kmode[c6] == 2)
                for (long i = 0; i < S; ++i) { o[2 * i] = (T)p[i].x; o[2 * i + 1] = (T)p[i].y; }
        }
    }
    if (c->khat_elems)
        HIPCHK(c, hipMemcpy(c->d_khat, packed.data(), packed.size() * sizeof(T), hipMemcpyHostToDevice));
    // new plane modes: device copy of the instances, per-launch bytes
    const int r2 = retile(c, 0, 0);
    return r2 ? r2 : rc;
}

int refresh_mirror(mmc_ctx *c)
{
    for (int i = 0; i < c->count; ++i) {
        Inst &in = c->inst[i];
        if (!in.cal_demag) continue;
        for (int k = 0; k < 6; ++k)
            HIPCHK(c, hipMemcpy(&in.K0[k], c->d_K + in.k_off + k * in.NP, sizeof(double),
                                hipMemcpyDeviceToHost));
        round_phys(in);
    }
    HIPCHK(c, hipMemcpy(c->dInst, c->inst.data(), sizeof(Inst) * c->count, hipMemcpyHostToDevice));
    return MMC_OK;
}

int finish_tensor(mmc_ctx *c)
{
    c->e_valid = false;
    int rc = refresh_mirror(c);
    if (rc) return rc;
    rc = c->tsize == 4 ? tensor_fft<float>(c) : tensor_fft<double>(c);
    if (rc) return rc;
    c->tensor_ready = true;
    return MMC_OK;
}

int check_inst(const mmc_ctx *c, int inst)
{
    if (!c) return MMC_EINVAL;
    if (inst < 0 || inst >= c->count) return fail(const_cast<mmc_ctx *>(c), MMC_EINVAL, "instance %d out of range", inst);
    return MMC_OK;
}

template <typename T>
int energy_impl(mmc_ctx *c, std::vector<double> &sums)
{
    if (c->e_valid) {
        sums = c->e_sums;
        return MMC_OK;
    }
    int rc = demag_chain<T>(c);
    if (rc) return rc;
    hipStream_t st = c->stream;
    HIPCHK(c, timed(c, K_ENERGY, [&] {
               return launch_energy<T>(st, c->dInst, c->cl.dev, c->cl.n(), (const T *)c->d_m,
                                       (const T *)c->d_h, c->d_part, c->any_demag ? 1 : 0);
           }));
    HIPCHK(c, launch_energy_final(st, c->d_efirst, c->count, c->d_part, c->d_eout));
    sums.resize(8 * c->count);
    HIPCHK(c, hipMemcpyAsync(sums.data(), c->d_eout, sizeof(double) * 8 * c->count,
                             hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    c->e_sums = sums;
    c->e_valid = true;
    return MMC_OK;
}

// ---- autotuner --------------------------------------------------------------
// A sweep schedule: fused x + Metropolis passes or separate launches (or the
// tiled two-colour Metropolis kernel), and the tile family.
struct Schedule {
    const char *name;
    bool fuse, tiled;
    int tiles;          // 0 chosen for the launch sizes, 1 minimal, 2 widest
    bool zfold;         // fused tiles of short-z instances fold z in
    bool conc;          // per-length launches of a pass on side streams
    bool rows;          // fused sweeps as the y pass + whole-sweep row-pair tiles
    int any;            // >= 0: y pass and fused passes one launch each over all
                        // instances, tile policy `any` (any_tile); -1: per length
};

// Fused tiles that fold z in where an instance admits them (zfold), else
// the tiles already chosen.
void set_zfold(mmc_ctx *c, bool on)
{
    c->zfold = on;
    for (size_t i = 0; i < c->inst.size(); ++i) {
        Inst &in = c->inst[i];
        if (!c->fuse_ok) { in.zf = 0; continue; }
        const int keep = in.zf ? c->tile_default[i][2] : in.tyf;
        const int r = fused_full(c, in);
        in.tyf = in.zf ? r : keep;
    }
}

int apply_schedule(mmc_ctx *c, const Schedule &s)
{
    for (size_t i = 0; i < c->inst.size(); ++i) {
        Inst &in = c->inst[i];
        const auto &d = s.tiles == 2 ? c->tile_full[i] : c->tile_default[i];
        in.tyx = d[0];
        in.ncol = d[1];
        in.tyf = c->fuse_ok ? d[2] : 0;
        if (s.tiles == 1) {
            in.tyx = 2;
            in.ncol = 1;
            in.tyf = in.tyf && !in.zf ? 1 : in.tyf;
        }
        if (s.any >= 0 && c->any_ok) {
            in.ncol = any_tile(s.any, in.Py);
            in.tyf = any_tile(s.any, in.Px);
        }
    }
    c->any_pol = c->any_ok ? s.any : -1;
    c->fuse = s.fuse;
    c->metro_tiled = s.tiled;
    c->concurrent = s.conc;
    c->rows = s.rows && c->rows_ok;
    for (auto &in : c->inst) in.zf = 0;
    set_zfold(c, s.zfold);
    return retile(c, 0, 0);
}

// Device time of `sweeps` sweeps under the current schedule (one untimed
// run first, which also captures the graph in graph mode).
template <typename T>
int time_sweeps(mmc_ctx *c, int sweeps, double *us)
{
    const uint32_t stage = 0xFFFFFFFFu;           // state is restored afterwards
    int rc = run_sweeps<T>(c, stage, 0, (uint32_t)sweeps);
    if (rc) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    HIPCHK(c, hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, c->stream);
    if (e == hipSuccess) {
        rc = run_sweeps<T>(c, stage, (uint32_t)sweeps, (uint32_t)sweeps);
        if (!rc) e = hipEventRecord(b, c->stream);
    }
    if (e == hipSuccess && !rc) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess && !rc) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(c, MMC_EHIP, "autotune timing: %s", hipGetErrorString(e));
    *us = 1e3 * ms / sweeps;
    return MMC_OK;
}

const Schedule kSchedules[] = {
        {"fused, launch-sized tiles", true, false, 0, false, false, false, -1},
        {"fused, minimal tiles", true, false, 1, false, false, false, -1},
        {"fused, widest tiles", true, false, 2, false, false, false, -1},
        {"fused, z folded", true, false, 0, true, false, false, -1},
        {"separate, launch-sized tiles", false, false, 0, false, false, false, -1},
        {"separate, minimal tiles", false, false, 1, false, false, false, -1},
        {"separate, widest tiles", false, false, 2, false, false, false, -1},
        {"tiled two-colour Metropolis", false, true, 0, false, false, false, -1},
        {"fused, launch-sized, concurrent", true, false, 0, false, true, false, -1},
        {"fused, minimal, concurrent", true, false, 1, false, true, false, -1},
        {"separate, launch-sized, concurrent", false, false, 0, false, true, false, -1},
        {"separate, minimal, concurrent", false, false, 1, false, true, false, -1},
        {"row sweeps, launch-sized tiles", true, false, 0, false, false, true, -1},
        {"row sweeps, minimal tiles", true, false, 1, false, false, true, -1},
        {"row sweeps, minimal, concurrent", true, false, 1, false, true, true, -1},
        {"one launch per pass, minimal tiles", true, false, 1, false, false, false, 0},
        {"one launch per pass, paired short rows", true, false, 1, false, false, false, 1},
        {"one launch per pass, paired rows to 256", true, false, 1, false, false, false, 2},
        {"one launch per pass, four short rows", true, false, 1, false, false, false, 3},
};
constexpr int kNumSchedules = (int)(sizeof(kSchedules) / sizeof(kSchedules[0]));

// Can schedule `s` run on this context, and does it differ from the others
// here (concurrency needs several kernel variants per pass, z folding needs
// short-z instances with z transforms)?
bool schedule_useful(mmc_ctx *c, const Schedule &s)
{
    // concurrent launches only matter where a pass has several kernel variants
    bool multi = false;
    {
        std::map<int, int> px, py;
        for (auto &in : c->inst)
            if (in.cal_demag) { px[in.Px]++; py[in.Py]++; }
        multi = px.size() > 1 || py.size() > 1;
    }
    bool any_zf = false;
    {
        const bool z = c->zfold;
        c->zfold = true;
        for (auto &in : c->inst) {
            Inst t = in;
            fused_full(c, t);
            any_zf = any_zf || t.zf;
        }
        c->zfold = z;
    }
    return !((s.fuse && !c->fuse_ok) || (s.tiled && !c->metro_tiled_ok) || (s.zfold && !any_zf) ||
             (s.conc && !multi) || (s.rows && !c->rows_ok) || (s.any >= 0 && (!c->any_ok || !multi)));
}

template <typename T>
int autotune_impl(mmc_ctx *c, int sweeps, std::string &report)
{
    const Schedule *all = kSchedules;
    void *save = nullptr;
    const size_t mbytes = std::max<size_t>(1, c->m_elems) * c->tsize;
    HIPCHK(c, hipMalloc(&save, mbytes));
    hipError_t e = hipMemcpyAsync(save, c->d_m, mbytes, hipMemcpyDeviceToDevice, c->stream);
    const bool timing = c->timing;
    c->timing = false;
    int rc = e == hipSuccess ? MMC_OK : fail(c, MMC_EHIP, "autotune: %s", hipGetErrorString(e));
    int best = -1;
    double best_us = 0;
    char line[160];
    for (int k = 0; k < kNumSchedules && !rc; ++k) {
        const Schedule &s = all[k];
        if (!schedule_useful(c, s)) continue;
        rc = apply_schedule(c, s);
        double us = 0;
        if (!rc) rc = time_sweeps<T>(c, sweeps, &us);
        if (rc) break;
        snprintf(line, sizeof(line), "%-32s %10.3f us/sweep\n", s.name, us);
        report += line;
        if (best < 0 || us < best_us) { best = k; best_us = us; }
    }
    if (!rc && best >= 0) {
        rc = apply_schedule(c, all[best]);
        snprintf(line, sizeof(line), "chosen: %s\n", all[best].name);
        report += line;
    }
    e = hipMemcpyAsync(c->d_m, save, mbytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(save);
    c->e_valid = false;
    c->timing = timing;
    if (!rc && e != hipSuccess) rc = fail(c, MMC_EHIP, "autotune restore: %s", hipGetErrorString(e));
    return rc;
}

}  // namespace

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

int mmc_create(const mmc_params *p, int precision, mmc_ctx **out)
{
    return create_common(p, 1, precision, out);
}

int mmc_batch_create(const mmc_params *p, int count, int precision, mmc_ctx **out)
{
    return create_common(p, count, precision, out);
}

void mmc_destroy(mmc_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (void *p : {c->d_m, c->d_m2, c->d_h, c->d_a, c->d_khat, (void *)c->d_K, c->d_tw32, c->d_tw64,
                    (void *)c->d_beta, (void *)c->d_ctl, (void *)c->d_part, (void *)c->d_eout,
                    (void *)c->d_efirst, (void *)c->d_H, (void *)c->dInst})
        if (p) (void)hipFree(p);
    for (List *l : {&c->xl, &c->zl, &c->yl, &c->cl, &c->ml, &c->mtl, &c->tl, &c->txl, &c->tzl, &c->tyl, &c->fl, &c->zls, &c->yzl, &c->rl})
        if (l->dev) (void)hipFree(l->dev);
    for (auto s : c->side) (void)hipStreamDestroy(s);
    for (auto e : c->join_ev) (void)hipEventDestroy(e);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *mmc_last_error(const mmc_ctx *c)
{
    if (!c) return "null context";
    return c->err.c_str();
}

int mmc_count(const mmc_ctx *c) { return c ? c->count : 0; }

int mmc_dims(const mmc_ctx *c, int inst, int out[6])
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    const Inst &in = c->inst[inst];
    out[0] = in.nx; out[1] = in.ny; out[2] = in.nz;
    out[3] = in.Px; out[4] = in.Py; out[5] = in.Pz;
    return MMC_OK;
}

int mmc_build_tensor(mmc_ctx *c, const char *cache_dir)
{
    if (!c) return MMC_EINVAL;
    if (!c->any_demag) { c->tensor_ready = true; return MMC_OK; }
    hipStream_t st = c->stream;
    std::vector<char> loaded(c->count, 0);
    if (cache_dir && *cache_dir) {
        for (int i = 0; i < c->count; ++i) {
            const Inst &in = c->inst[i];
            if (!in.cal_demag) continue;
            std::vector<double> K;
            if (load_cache(std::string(cache_dir) + "/" + cache_name(in, c->params[i]), in, c->params[i], K)) {
                HIPCHK(c, hipMemcpy(c->d_K + in.k_off, K.data(), sizeof(double) * K.size(),
                                    hipMemcpyHostToDevice));
                loaded[i] = 1;
            }
        }
    }
    // image slices: bound each launch to about c->tensor_evals Newell
    // evaluations (a launch covers at least one x image)
    int max_img = 0;
    for (auto &in : c->inst) if (in.cal_demag) max_img = std::max(max_img, in.img[0]);
    double evals = 0;
    for (int i = 0; i < c->count; ++i)
        if (!loaded[i] && c->inst[i].cal_demag)
            evals += (double)c->inst[i].NP * 162.0 * (2 * c->inst[i].img[1] + 1) * (2 * c->inst[i].img[2] + 1);
    // build a work list of the instances still to compute
    std::vector<Work> todo;
    for (auto &w : c->tl.host) if (!loaded[w.inst]) todo.push_back(w);
    if (!todo.empty()) {
        Work *dW = nullptr;
        HIPCHK(c, hipMalloc(&dW, sizeof(Work) * todo.size()));
        HIPCHK(c, hipMemcpy(dW, todo.data(), sizeof(Work) * todo.size(), hipMemcpyHostToDevice));
        const int per = (int)std::max(1.0, std::min(1e9, c->tensor_evals / std::max(1.0, evals)));
        int rc = MMC_OK;
        for (int lo = -max_img; lo <= max_img && rc == MMC_OK; lo += per) {
            const int hi = std::min(max_img, lo + per - 1);
            hipError_t e = timed(c, K_TENSOR, [&] {
                return launch_tensor(st, c->dInst, dW, (int)todo.size(), c->d_K, lo, hi,
                                     lo == -max_img, hi == max_img);
            });
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = fail(c, MMC_EHIP, "tensor build failed: %s", hipGetErrorString(e));
        }
        (void)hipFree(dW);
        if (rc) return rc;
    }
    if (cache_dir && *cache_dir) {
        struct stat sb;
        if (stat(cache_dir, &sb) == 0 && S_ISDIR(sb.st_mode))
            for (int i = 0; i < c->count; ++i) {
                const Inst &in = c->inst[i];
                if (!in.cal_demag || loaded[i]) continue;
                std::vector<double> K(6 * in.NP);
                HIPCHK(c, hipMemcpy(K.data(), c->d_K + in.k_off, sizeof(double) * K.size(),
                                    hipMemcpyDeviceToHost));
                save_cache(std::string(cache_dir) + "/" + cache_name(in, c->params[i]), in, c->params[i], K);
            }
    }
    return finish_tensor(c);
}

int mmc_set_tensor(mmc_ctx *c, int inst, const double *K)
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    const Inst &in = c->inst[inst];
    if (!in.cal_demag) return fail(c, MMC_ESTATE, "instance %d has cal_demag off", inst);
    HIPCHK(c, hipMemcpy(c->d_K + in.k_off, K, sizeof(double) * 6 * in.NP, hipMemcpyHostToDevice));
    return finish_tensor(c);
}

int mmc_get_tensor(const mmc_ctx *cc, int inst, double *K)
{
    mmc_ctx *c = const_cast<mmc_ctx *>(cc);
    int rc = check_inst(c, inst);
    if (rc) return rc;
    const Inst &in = c->inst[inst];
    if (!in.cal_demag) return fail(c, MMC_ESTATE, "instance %d has cal_demag off", inst);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(K, c->d_K + in.k_off, sizeof(double) * 6 * in.NP, hipMemcpyDeviceToHost));
    return MMC_OK;
}

int mmc_mirror_tensor(const mmc_ctx *cc, int inst, double K0[6])
{
    int rc = check_inst(cc, inst);
    if (rc) return rc;
    for (int k = 0; k < 6; ++k) K0[k] = cc->inst[inst].K0[k];
    return MMC_OK;
}

int mmc_set_m(mmc_ctx *c, int inst, const double *mx, const double *my, const double *mz)
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    const Inst &in = c->inst[inst];
    const double *src[3] = {mx, my, mz};
    c->e_valid = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) {
        if (c->tsize == 8) {
            HIPCHK(c, hipMemcpy((double *)c->d_m + in.m_off + k * in.N, src[k], sizeof(double) * in.N,
                                hipMemcpyHostToDevice));
        } else {
            std::vector<float> tmp(src[k], src[k] + in.N);
            HIPCHK(c, hipMemcpy((float *)c->d_m + in.m_off + k * in.N, tmp.data(), sizeof(float) * in.N,
                                hipMemcpyHostToDevice));
        }
    }
    return MMC_OK;
}

int mmc_get_m(const mmc_ctx *cc, int inst, double *mx, double *my, double *mz)
{
    mmc_ctx *c = const_cast<mmc_ctx *>(cc);
    int rc = check_inst(c, inst);
    if (rc) return rc;
    const Inst &in = c->inst[inst];
    double *dst[3] = {mx, my, mz};
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) {
        if (c->tsize == 8) {
            HIPCHK(c, hipMemcpy(dst[k], (double *)c->d_m + in.m_off + k * in.N, sizeof(double) * in.N,
                                hipMemcpyDeviceToHost));
        } else {
            std::vector<float> tmp(in.N);
            HIPCHK(c, hipMemcpy(tmp.data(), (float *)c->d_m + in.m_off + k * in.N, sizeof(float) * in.N,
                                hipMemcpyDeviceToHost));
            for (long i = 0; i < in.N; ++i) dst[k][i] = tmp[i];
        }
    }
    return MMC_OK;
}

int mmc_init_random_m(mmc_ctx *c, int inst)
{
    if (!c) return MMC_EINVAL;
    if (inst >= c->count || inst < -1) return fail(c, MMC_EINVAL, "instance %d out of range", inst);
    const int b0 = inst < 0 ? 0 : c->cl.first[inst];
    const int b1 = inst < 0 ? c->cl.n() : c->cl.first[inst + 1];
    c->e_valid = false;
    hipError_t e = c->tsize == 4
                       ? launch_init<float>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0, (float *)c->d_m)
                       : launch_init<double>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0, (double *)c->d_m);
    HIPCHK(c, e);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MMC_OK;
}

int mmc_field(mmc_ctx *c, int inst, int terms, double *hx, double *hy, double *hz)
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    const Inst &in = c->inst[inst];
    if ((terms & MMC_DEMAG) && in.cal_demag) {
        rc = c->tsize == 4 ? demag_chain<float>(c) : demag_chain<double>(c);
        if (rc) return rc;
    }
    const int b0 = c->cl.first[inst], b1 = c->cl.first[inst + 1];
    hipError_t e = c->tsize == 4
                       ? launch_field<float>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0,
                                             (const float *)c->d_m, (const float *)c->d_h, terms, c->d_H)
                       : launch_field<double>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0,
                                              (const double *)c->d_m, (const double *)c->d_h, terms, c->d_H);
    HIPCHK(c, e);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double *dst[3] = {hx, hy, hz};
    for (int k = 0; k < 3; ++k)
        HIPCHK(c, hipMemcpy(dst[k], c->d_H + k * in.N, sizeof(double) * in.N, hipMemcpyDeviceToHost));
    return MMC_OK;
}

int mmc_energy_density(mmc_ctx *c, int inst, int terms, double *e)
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    if (!e) return fail(c, MMC_EINVAL, "null output");
    const Inst &in = c->inst[inst];
    if ((terms & MMC_DEMAG) && in.cal_demag) {
        rc = c->tsize == 4 ? demag_chain<float>(c) : demag_chain<double>(c);
        if (rc) return rc;
    }
    const int b0 = c->cl.first[inst], b1 = c->cl.first[inst + 1];
    hipError_t err = c->tsize == 4
        ? launch_energy_density<float>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0,
                                       (const float *)c->d_m, (const float *)c->d_h, terms, c->d_H)
        : launch_energy_density<double>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0,
                                        (const double *)c->d_m, (const double *)c->d_h, terms, c->d_H);
    HIPCHK(c, err);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(e, c->d_H, sizeof(double) * in.N, hipMemcpyDeviceToHost));
    return MMC_OK;
}

int mmc_trial_delta_e(mmc_ctx *c, int inst, uint32_t stage, uint32_t sweep, double *dE)
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    if (!dE) return fail(c, MMC_EINVAL, "null output");
    const Inst &in = c->inst[inst];
    if (in.cal_demag) {
        rc = c->tsize == 4 ? demag_chain<float>(c) : demag_chain<double>(c);
        if (rc) return rc;
    }
    const int b0 = c->cl.first[inst], b1 = c->cl.first[inst + 1];
    hipError_t e = c->tsize == 4
                       ? launch_delta_e<float>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0,
                                               (const float *)c->d_m, (const float *)c->d_h,
                                               stage, sweep, c->d_H)
                       : launch_delta_e<double>(c->stream, c->dInst, c->cl.dev + b0, b1 - b0,
                                                (const double *)c->d_m, (const double *)c->d_h,
                                                stage, sweep, c->d_H);
    HIPCHK(c, e);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(dE, c->d_H, sizeof(double) * in.N, hipMemcpyDeviceToHost));
    return MMC_OK;
}

int mmc_energy(mmc_ctx *c, int inst, double *E_total, double e_terms[5], double m_avg[3])
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    std::vector<double> s;
    rc = c->tsize == 4 ? energy_impl<float>(c, s) : energy_impl<double>(c, s);
    if (rc) return rc;
    const Inst &in = c->inst[inst];
    const double *v = s.data() + 8 * inst;
    double tot = 0.0;
    double et[5];
    for (int t = 0; t < 5; ++t) {
        et[t] = v[t] * in.e_unit;
        tot += et[t];
    }
    if (E_total) *E_total = tot;
    if (e_terms) for (int t = 0; t < 5; ++t) e_terms[t] = et[t];
    if (m_avg) for (int k = 0; k < 3; ++k) m_avg[k] = v[5 + k] / (double)in.N;
    return MMC_OK;
}

int mmc_energy_noalloc(mmc_ctx *c, int inst, double *E_total, double e_terms[5], double m_avg[3])
{
    int rc = check_inst(c, inst);
    if (rc) return rc;
    void *sp = c->device >= 0 && c->device < kMaxDevices ? g_scratch_ptr[c->device] : nullptr;
    if (!sp) return fail(c, MMC_ESTATE, "scratch of device %d not set up", c->device);
    if (c->count > kNoAllocMaxInst)
        return fail(c, MMC_EINVAL, "mmc_energy_noalloc: %d instances, at most %d", c->count,
                    kNoAllocMaxInst);
    const size_t need = kScratchSums + (size_t)16 * c->cl.n() * sizeof(double);
    if (need > kScratchBytes)
        return fail(c, MMC_EINVAL,
                    "mmc_energy_noalloc: %d reduction blocks need %zu bytes of scratch, above the "
                    "%zu-byte maximum (at most %zu blocks of %d cells)", c->cl.n(), need,
                    kScratchBytes, (kScratchBytes - kScratchSums) / (16 * sizeof(double)), kThreads);
    double *sums = (double *)sp;
    double *part = (double *)((unsigned char *)sp + kScratchSums);
    // untimed launches on the context's own stream only: no events, no side
    // streams, no device memory beyond what the context already holds
    const bool timing = c->timing, conc = c->concurrent;
    c->timing = false;
    c->concurrent = false;
    rc = c->tsize == 4 ? demag_chain<float>(c) : demag_chain<double>(c);
    hipError_t e = hipSuccess;
    if (!rc) {
        hipStream_t st = c->stream;
        e = c->tsize == 4
                ? launch_energy<float>(st, c->dInst, c->cl.dev, c->cl.n(), (const float *)c->d_m,
                                       (const float *)c->d_h, part, c->any_demag ? 1 : 0)
                : launch_energy<double>(st, c->dInst, c->cl.dev, c->cl.n(), (const double *)c->d_m,
                                        (const double *)c->d_h, part, c->any_demag ? 1 : 0);
        if (e == hipSuccess) e = launch_energy_final(st, c->d_efirst, c->count, part, sums);
    }
    double v[8];
    if (!rc && e == hipSuccess)
        e = hipMemcpyAsync(v, sums + 8 * inst, sizeof(v), hipMemcpyDeviceToHost, c->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(c->stream);
    c->timing = timing;
    c->concurrent = conc;
    if (rc) return rc;
    if (e != hipSuccess) return fail(c, MMC_EHIP, "mmc_energy_noalloc: %s", hipGetErrorString(e));
    const Inst &in = c->inst[inst];
    double tot = 0.0, et[5];
    for (int t = 0; t < 5; ++t) {
        et[t] = v[t] * in.e_unit;
        tot += et[t];
    }
    if (E_total) *E_total = tot;
    if (e_terms) for (int t = 0; t < 5; ++t) e_terms[t] = et[t];
    if (m_avg) for (int k = 0; k < 3; ++k) m_avg[k] = v[5 + k] / (double)in.N;
    return MMC_OK;
}

int mmc_set_temperature(mmc_ctx *c, int inst, double T)
{
    if (!c) return MMC_EINVAL;
    if (inst >= c->count || inst < -1) return fail(c, MMC_EINVAL, "instance %d out of range", inst);
    if (!(T > 0)) return fail(c, MMC_EINVAL, "temperature must be positive");
    for (int i = 0; i < c->count; ++i)
        if (inst < 0 || i == inst) c->T[i] = T;
    std::vector<double> beta(c->count);
    for (int i = 0; i < c->count; ++i) beta[i] = c->inst[i].e_unit / (kBoltzmann * c->T[i]);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_beta, beta.data(), sizeof(double) * c->count, hipMemcpyHostToDevice));
    return MMC_OK;
}

int mmc_sweeps(mmc_ctx *c, uint32_t stage, uint32_t sweep0, uint32_t count)
{
    if (!c) return MMC_EINVAL;
    return c->tsize == 4 ? run_sweeps<float>(c, stage, sweep0, count)
                         : run_sweeps<double>(c, stage, sweep0, count);
}

int mmc_autotune(mmc_ctx *c, int sweeps_per_trial, char *report, size_t report_len)
{
    if (!c) return MMC_EINVAL;
    if (sweeps_per_trial < 1 || sweeps_per_trial > 100000)
        return fail(c, MMC_EINVAL, "sweeps_per_trial must be in [1,100000]");
    if (c->any_demag && !c->tensor_ready)
        return fail(c, MMC_ESTATE, "demag tensor not built (call mmc_build_tensor)");
    std::string rep;
    const int rc = c->tsize == 4 ? autotune_impl<float>(c, sweeps_per_trial, rep)
                                 : autotune_impl<double>(c, sweeps_per_trial, rep);
    if (report && report_len) {
        const size_t n = std::min(report_len - 1, rep.size());
        std::memcpy(report, rep.data(), n);
        report[n] = 0;
    }
    return rc;
}

int mmc_set_schedule(mmc_ctx *c, const char *name)
{
    if (!c || !name) return MMC_EINVAL;
    for (int k = 0; k < kNumSchedules; ++k)
        if (!strcmp(name, kSchedules[k].name)) {
            const Schedule &s = kSchedules[k];
            if ((s.fuse && !c->fuse_ok) || (s.tiled && !c->metro_tiled_ok) || (s.any >= 0 && !c->any_ok))
                return fail(c, MMC_EUNSUP, "schedule '%s' does not apply to these instances", name);
            HIPCHK(c, hipStreamSynchronize(c->stream));
            return apply_schedule(c, s);
        }
    return fail(c, MMC_EINVAL, "unknown schedule '%s'", name);
}

int mmc_sweep(mmc_ctx *c, double T, uint32_t stage, uint32_t sweep)
{
    int rc = mmc_set_temperature(c, -1, T);
    if (rc) return rc;
    return mmc_sweeps(c, stage, sweep, 1);
}

int mmc_sync(mmc_ctx *c)
{
    if (!c) return MMC_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MMC_OK;
}

int mmc_anneal(mmc_ctx *c, const mmc_schedule *s, mmc_energy_cb ecb, mmc_snapshot_cb scb,
               void *user, volatile int *stop, long *stages_out)
{
    if (!c || !s) return MMC_EINVAL;
    if (s->sweeps_per_stage < 0) return fail(c, MMC_EINVAL, "sweeps_per_stage < 0");
    double T = s->T_start;
    long stage = 0;
    const Inst &in = c->inst[0];
    std::vector<double> mx, my, mz;
    int rc = MMC_OK;
    while (T > s->T_end) {                                   // src/main.cu:801
        if (stop && *stop) break;
        if ((rc = mmc_set_temperature(c, -1, T))) return rc;
        long sw = 0;
        while (sw < s->sweeps_per_stage) {
            if (stop && *stop) break;
            // next sweep index at which an output is due
            long next = s->sweeps_per_stage;
            auto due = [&](long period) {
                if (period <= 0) return;
                long k = ((sw + period - 1) / period) * period;
                next = std::min(next, k);
            };
            due(s->energy_period);
            due(s->snapshot_period);
            if (next == sw) {
                if (s->snapshot_period > 0 && sw % s->snapshot_period == 0 && scb) {
                    mx.resize(in.N); my.resize(in.N); mz.resize(in.N);
                    if ((rc = mmc_get_m(c, 0, mx.data(), my.data(), mz.data()))) return rc;
                    if (scb(user, T, stage, sw, mx.data(), my.data(), mz.data())) { if (stages_out) *stages_out = stage; return MMC_OK; }
                }
                if (s->energy_period > 0 && sw % s->energy_period == 0 && ecb) {
                    double E = 0;
                    if ((rc = mmc_energy(c, 0, &E, nullptr, nullptr))) return rc;
                    if (ecb(user, T, stage, sw, E)) { if (stages_out) *stages_out = stage; return MMC_OK; }
                }
                next = sw + 1;
                long k2 = s->sweeps_per_stage;
                auto due2 = [&](long period) {
                    if (period <= 0) return;
                    long k = ((sw + 1 + period - 1) / period) * period;
                    k2 = std::min(k2, k);
                };
                due2(s->energy_period);
                due2(s->snapshot_period);
                next = k2;
            }
            // run up to the next output point, in chunks so *stop is polled
            long chunk = std::min(next - sw, 1000L);
            if ((rc = mmc_sweeps(c, (uint32_t)stage, (uint32_t)sw, (uint32_t)chunk))) return rc;
            if ((rc = mmc_sync(c))) return rc;
            sw += chunk;
        }
        ++stage;
        if (stop && *stop) break;
        T = s->use_exp ? T * s->T_step : T - s->T_step;     // src/main.cu:1129-1136
    }
    if (stages_out) *stages_out = stage;
    return MMC_OK;
}

int mmc_set_option(mmc_ctx *c, const char *key, long value)
{
    if (!c || !key) return MMC_EINVAL;
    std::string k(key);
    if (k == "graph") {
        if (value && !c->stream)
            return fail(c, MMC_EUNSUP, "graph replay needs the context's own stream");
        c->use_graph = value != 0;
        return MMC_OK;
    }
    if (k == "default_stream") {
        // 1: launch on the device's default stream (for tools that only
        // observe it); 0: back to a stream of the context's own
        HIPCHK(c, c->stream ? hipStreamSynchronize(c->stream) : hipDeviceSynchronize());
        if (value && c->stream) {
            if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
            if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
            HIPCHK(c, hipStreamDestroy(c->stream));
            c->stream = nullptr;
            c->use_graph = false;
        } else if (!value && !c->stream) {
            HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        }
        return MMC_OK;
    }
    if (k == "graph_sweeps") {
        if (value < 1 || value > 4096) return fail(c, MMC_EINVAL, "graph_sweeps must be in [1,4096]");
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        c->graph_sweeps = (int)value;
        return MMC_OK;
    }
    if (k == "fuse_rows") {
        // row pairs per fused tile: 1, or (value > 1) the full tile
        if (value < 1) return fail(c, MMC_EINVAL, "fuse_rows must be >= 1");
        for (auto &in : c->inst)
            if (in.tyf) in.tyf = (value > 1 || in.zf) ? fused_full(c, in) : 1;
        return retile(c, 0, 0);
    }
    if (k == "concurrent") {
        c->concurrent = value != 0;
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        return MMC_OK;
    }
    if (k == "zfold") {
        set_zfold(c, value != 0);
        return retile(c, 0, 0);
    }
    if (k == "fuse") {
        if (value && !c->fuse_ok)
            return fail(c, MMC_EUNSUP, "fused x-pass + Metropolis sweeps need demag, a "
                                       "checkerboard and a fixed-plan x length for every instance");
        c->fuse = value != 0;
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        return MMC_OK;
    }
    if (k == "rows") {
        if (value && !c->rows_ok)
            return fail(c, MMC_EUNSUP, "row sweeps need checkerboard films of one or two layers "
                                       "with demag, a fused-kernel x length and room for the "
                                       "tile in LDS for every instance");
        c->rows = value != 0;
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        return MMC_OK;
    }
    if (k == "metro_tiled") {
        if (value && !c->metro_tiled_ok)
            return fail(c, MMC_EUNSUP, "tiled Metropolis needs a checkerboard (even periodic axes) "
                                       "and a tile that fits LDS for every instance");
        c->metro_tiled = value != 0;
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        return MMC_OK;
    }
    if (k == "tyx" || k == "ncol") {
        if (value < 1 || value > 4096) return fail(c, MMC_EINVAL, "%s must be in [1,4096]", key);
        return k == "tyx" ? retile(c, (int)value, 0) : retile(c, 0, (int)value);
    }
    if (k == "tensor_evals") {
        // Newell evaluations per tensor-build launch; the x images are split
        // into slices of at least one image each (1: one launch per image)
        if (value < 1) return fail(c, MMC_EINVAL, "tensor_evals must be >= 1");
        c->tensor_evals = (double)value;
        return MMC_OK;
    }
    if (k == "timing") {
        if (!value) harvest(c);
        c->timing = value != 0;
        return MMC_OK;
    }
    return fail(c, MMC_EINVAL, "unknown option '%s'", key);
}

int mmc_kernel_stats(mmc_ctx *c, const char *kind, long *launches, double *total_ms,
                     double *total_bytes, int reset)
{
    if (!c || !kind) return MMC_EINVAL;
    harvest(c);
    for (int k = 0; k < kNumKinds; ++k)
        if (!strcmp(kind, kKinds[k])) {
            if (launches) *launches = c->launches[k];
            if (total_ms) *total_ms = c->total_ms[k];
            if (total_bytes) *total_bytes = c->total_bytes[k];
            if (reset) { c->launches[k] = 0; c->total_ms[k] = 0; c->total_bytes[k] = 0; }
            return MMC_OK;
        }
    return fail(c, MMC_EINVAL, "unknown kernel kind '%s'", kind);
}

int mmc_kernel_bytes(const mmc_ctx *cc, const char *kind, double *bytes)
{
    mmc_ctx *c = const_cast<mmc_ctx *>(cc);
    if (!c || !kind || !bytes) return MMC_EINVAL;
    for (int k = 0; k < kNumKinds; ++k)
        if (!strcmp(kind, kKinds[k])) {
            *bytes = all_bytes(c, k);
            return MMC_OK;
        }
    return fail(c, MMC_EINVAL, "unknown kernel kind '%s'", kind);
}

}  // extern "C"
