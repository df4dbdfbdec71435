// mmc_internal.h -- device data layout shared by the kernels and the C ABI.
//
// One context holds one or more instances (independent anneals).  Every
// per-instance array lives in a pooled device buffer at an offset recorded
// in Inst; every kernel walks a device-resident work list whose items name
// the instance they belong to, so a single launch covers all instances.
//
// HBM layout of one instance (b = sizeof(T), T = float or double):
//   m, h        [3][nz][ny][nx]            T       magnetisation, demag/Ms
//   A           [3][Hx][nz][ny]            cpx<T>  x-spectrum, y fastest
//   B           [3][Hx][Pz][ny]            cpx<T>  xz-spectrum (A if Pz==1; same pool)
//   Khat        [6][Hx][Pz][Py]            cpx<T>  tensor spectrum / (PxPyPz), or with
//                                                  zd the x-y spectra of the Pz z slices
//                                                  / (PxPy)
//   K           [6][Pz][Py][Px]            double  real-space tensor
// with Hx = Px/2 + 1 (half spectrum of the real x transform).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mmc {

constexpr int kThreads = 256;      // threads per workgroup (4 wave64)
constexpr int kMaxStages = 16;     // radix stages per FFT plan
constexpr int kZDirect = 8;        // z lengths up to this use a direct DFT

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

struct FftPlan {
    int n;
    int nst;
    int radix[kMaxStages];
};

struct Inst {
    int nx, ny, nz;
    int Px, Py, Pz, Hx;
    int pbc[3];           // nonzero: periodic axis
    int img[3];           // image half-range of the tensor sum per axis
    int ncolors;          // 2 (checkerboard) or 3 (odd periodic axis)
    int cal_demag;
    int tyx, tyz, ncol;   // tile shapes of the bulk-precision FFT passes; each
                          // pass ping-pongs two LDS buffers of tile size
    int tyx64, tyz64;     // tile shapes of the FP64 tensor transform
    int mtx, mty;         // Metropolis x-y tile (both colours in one pass)
    int tyf;              // complex rows per component of a fused x-pass +
                          // Metropolis tile (0: none)
    int zf;               // fused tiles span all layers and transform z too
    int zd;               // z convolution summed directly in the y pass
                          // (k_yzconv_n); Khat holds x-y spectra of z slices
    int a2;               // even ny, and zd or Pz == 1 with a compile-time y
                          // plan: in fused sweeps the y pass writes the x
                          // spectrum pair-major at pm_off,
                          // [3][nz][ny/2][Hx][2] (a row pair's spectrum
                          // contiguous), which the inverse x pass reads
    int kmode[6];         // Khat storage of each component: 0 zero, 1 real
                          // (imaginary part below rounding), 2 complex
    long N, NP;
    long m_off, h_off;    // T
    long a_off, b_off;    // cpx<T>
    long pm_off;          // cpx<T>, pair-major spectrum (a2): the B region of
                          // a zd instance, a region of its own for Pz == 1
    long khat_off;        // cpx<T>
    long k_off;           // double
    long ka_off;          // cpx<double>, tensor-build scratch
    long twx, twy, twz;   // twiddle tables (same offsets in the FP32 and FP64 pools)
    FftPlan plx, ply, plz;
    double d[3];          // dx, dy, dz [m]
    double lam[3];        // 2A / (mu0 Ms^2 d_a^2)
    double dmi[2];        // D / (mu0 Ms^2 d_a), a = x, y
    double kap;           // Ku1 / (mu0 Ms^2)
    double u[3];          // unit anisotropy axis
    double b[3];          // B / (mu0 Ms)
    double K0[6];         // self ("mirror") tensor
    // the constants above rounded to FP32 once on the host, so FP32 kernels
    // read them as they are rather than converting an FP64 value at each use
    // (the same rounding either way)
    struct PhysF {
        float lam[3], dmi[2], kap, u[3], b[3], K0[6];
    } pf;
    double ms;            // Ms [A/m]
    double e_unit;        // mu0 Ms^2 V [J]
    uint32_t key[2];      // Philox key = seed
};

struct Work {
    int inst, a, b, c;
};

// ---- launch wrappers (mmc_kernels.hip) ------------------------------------
// Each takes a device work list (items [0, n)), launches one workgroup per
// item on `st`, and returns the launch status.
hipError_t launch_tensor(hipStream_t st, const Inst *I, const Work *W, int n,
                         double *K, int ix_lo, int ix_hi, int first, int last);

// Row pairs of a full fixed-plan x-pass tile (16 FP32 / 8 FP64 real rows,
// so a tile's stores along y fill 128-byte segments).
template <typename T> struct kXRows;
template <> struct kXRows<float> { static constexpr int value = 8; };
template <> struct kXRows<double> { static constexpr int value = 4; };

// fixed_n, rows: x length and row pairs of a fixed-plan tile (magnetisation
// mode only), or fixed_n = 0 for the generic kernel.
template <typename T>
hipError_t launch_xfwd(hipStream_t st, const Inst *I, const Work *W, int n,
                       int mode, int fixed_n, int rows, const T *real,
                       cpx<T> *spec, const cpx<T> *tw, size_t lds);
// fixed_pz: the common z length of all items when it has a compile-time
// plan (16, 32, 64; magnetisation mode, y tiles of z_cols columns), else 0.
// half (planned passes only): every item's instance has 2 nz <= Pz, so the
// transforms skip the zero upper half of their input / output.
template <typename T>
hipError_t launch_zfwd(hipStream_t st, const Inst *I, const Work *W, int n,
                       int mode, int fixed_pz, const cpx<T> *src, cpx<T> *dst,
                       const cpx<T> *tw, size_t lds, bool half);
template <typename T>
hipError_t launch_zinv(hipStream_t st, const Inst *I, const Work *W, int n,
                       int fixed_pz, const cpx<T> *src, cpx<T> *dst, const cpx<T> *tw,
                       size_t lds, bool half);
// y columns per item of the compile-time-planned z passes: 32 (256 bytes of
// each z row in FP32), one radix-4 butterfly per thread in a z stage of
// length 32; 64 columns (half the workgroups) and 16 were both slower
__host__ __device__ constexpr int z_cols(size_t) { return 32; }
// fixed_n: the common y length of all items (compile-time-planned kernel),
// or 0 for the generic kernel; half: every item's instance has 2 ny <= Py,
// so the planned transforms skip the zero upper half of their input and
// the discarded upper half of their result.
// a2: in a fused sweep, instances with Inst::a2 store their field spectrum
// pair-major at pm_off for the inverse x pass instead of back in place.
template <typename T>
hipError_t launch_ymul(hipStream_t st, const Inst *I, const Work *W, int n,
                       int fixed_n, int cols, cpx<T> *buf, const cpx<T> *khat,
                       const cpx<T> *tw, size_t lds, bool half, bool a2);
// Threads of a k_yzconv_n workgroup: its items hold all layers of one x
// frequency (few large items), so each gets twice the usual threads.
constexpr int kYzThreads = 512;
// Layers up to which the y pass sums the z convolution directly.
constexpr int kZConvMax = 4;
// Instances with 1 < nz <= kZConvMax layers, Pz <= kZDirect and Py = fixed_n
// (a fixed plan) whose tensor spectrum holds the x-y spectra of its z
// slices (Inst::zd): y transform, Khat product with the z convolution
// summed directly, y inverse, in one pass over A; items (inst, kx), one
// launch per (fixed_n, nz).
// a2: write the result pair-major into the B region of instances with
// Inst::a2 (fused sweeps, whose inverse x pass reads it there).
template <typename T>
hipError_t launch_yzconv(hipStream_t st, const Inst *I, const Work *W, int n,
                         int fixed_n, int nz, cpx<T> *A, const cpx<T> *khat,
                         const cpx<T> *tw, size_t lds, bool a2);
template <typename TO>
hipError_t launch_yfwd_tensor(hipStream_t st, const Inst *I, const Work *W,
                              int n, const cpx<double> *ka, cpx<TO> *khat,
                              const cpx<double> *tw, size_t lds);
template <typename T>
hipError_t launch_xinv(hipStream_t st, const Inst *I, const Work *W, int n,
                       int fixed_n, int rows, const cpx<T> *spec, T *real,
                       const cpx<T> *tw, size_t lds);
// Fused x pass + Metropolis tiles: `rows` row pairs of all three components
// of x length fixed_n (a fixed-plan length); rows is 1 or
// fused_rows(fused_rows_base<T>(), fixed_n).
template <typename T> constexpr int fused_rows_base() { return sizeof(T) == 4 ? 2048 : 1024; }
constexpr int fused_rows(int base, int n) { return base / n > 1 ? base / n : 1; }
// Inverse x pass + colour 0: reads spec, updates m, stores colour-1 h
// compactly in the h buffer.  zf: some item's instance folds z into its
// tiles (Inst::zf; the kernels without zf have no z-folding code).
template <typename T>
hipError_t launch_xinv_metro(hipStream_t st, const Inst *I, const Work *W, int n,
                             int fixed_n, int rows, bool zf, const cpx<T> *spec, T *m,
                             T *h, const double *beta, const uint32_t *ctl,
                             uint32_t sweep_off, const cpx<T> *tw, size_t lds);
// Colour 1 (compact h) + forward x pass of the updated rows into spec.
template <typename T>
hipError_t launch_metro_xfwd(hipStream_t st, const Inst *I, const Work *W, int n,
                             int fixed_n, int rows, bool zf, T *m, const T *h,
                             cpx<T> *spec, const double *beta,
                             const uint32_t *ctl, uint32_t sweep_off,
                             const cpx<T> *tw, size_t lds);
// One launch per pass over a batch of uneven instances: the "any" kernels
// read each work item's transform length and run that length's
// compile-time plan (launch_ymul / launch_xinv_metro / launch_metro_xfwd
// with fixed_n = kAnyLen; `rows` / `cols` is then the tile policy).
// Tile policy 0: one row pair (x) or one column (y) per item; 1: two for
// lengths below 256, whose single rows leave most of a workgroup idle; 2:
// two up to 256; 3: four below 256, two at 256.
constexpr int kAnyLen = -1;
__host__ __device__ constexpr bool any_x_len(int n)
{
    return n == 128 || n == 192 || n == 256 || n == 384 || n == 512;
}
__host__ __device__ constexpr bool any_y_len(int n) { return n == 64 || n == 96 || any_x_len(n); }
__host__ __device__ constexpr int any_tile(int pol, int n)
{
    return pol == 1 ? (n < 256 ? 2 : 1)
         : pol == 2 ? (n <= 256 ? 2 : 1)
         : pol == 3 ? (n < 256 ? 4 : (n == 256 ? 2 : 1)) : 1;
}
constexpr int kAnyPolicies = 4;
// Threads of a k_row_sweep workgroup: a tile is alone on its CU (small films
// have fewer tiles than CUs; large ones fill its LDS), so its Metropolis
// cells and butterflies are spread over sixteen waves to keep their
// dependent latencies short.
constexpr int kRowThreads = 1024;
// A whole sweep of row pairs, all layers, of checkerboard films of nz = 1 or
// 2 layers (x length fixed_n, a fused-kernel length): items (inst, 0, 0,
// y0), y0 even; reads the pre-sweep state m_in and the x spectrum, writes
// the pairs' final state to m_out and their x spectrum back to spec.
template <typename T>
hipError_t launch_row_sweep(hipStream_t st, const Inst *I, const Work *W, int n,
                            int fixed_n, int nz, cpx<T> *spec, const T *m_in, T *m_out,
                            const double *beta, const uint32_t *ctl,
                            uint32_t sweep_off, const cpx<T> *tw, size_t lds);
template <typename T>
hipError_t launch_metro(hipStream_t st, const Inst *I, const Work *W, int n,
                        T *m, const T *h, const double *beta,
                        const uint32_t *ctl, uint32_t sweep_off, int color,
                        int use_demag);
// Both checkerboard colours in one launch; items (inst, x0, y0).
template <typename T>
hipError_t launch_metro_tile(hipStream_t st, const Inst *I, const Work *W, int n,
                             T *m, const T *h, const double *beta,
                             const uint32_t *ctl, uint32_t sweep_off,
                             int use_demag, size_t lds);
template <typename T>
hipError_t launch_energy(hipStream_t st, const Inst *I, const Work *W, int n,
                         const T *m, const T *h, double *partials,
                         int use_demag);
hipError_t launch_energy_final(hipStream_t st, const int *first, int ninst,
                               const double *partials, double *out);
template <typename T>
hipError_t launch_field(hipStream_t st, const Inst *I, const Work *W, int n,
                        const T *m, const T *h, int terms, double *H);
template <typename T>
hipError_t launch_energy_density(hipStream_t st, const Inst *I, const Work *W,
                                 int n, const T *m, const T *h, int terms,
                                 double *out);
template <typename T>
hipError_t launch_delta_e(hipStream_t st, const Inst *I, const Work *W, int n,
                          const T *m, const T *h, uint32_t stage,
                          uint32_t sweep, double *out);
template <typename T>
hipError_t launch_init(hipStream_t st, const Inst *I, const Work *W, int n,
                       T *m);
hipError_t launch_set_ctl(hipStream_t st, uint32_t *ctl, uint32_t stage,
                          uint32_t sweep_base);
// Allow `bytes` of dynamic LDS on every FFT kernel (above the 64 KB default).
hipError_t set_lds_limit(size_t bytes);
// Events for the next launch of this thread: the dispatch packet itself
// records a (start) and b (end), so the measured span is the kernel alone.
// Reset after every launch wrapper call; (nullptr, nullptr) = untimed.
void set_launch_events(hipEvent_t a, hipEvent_t b);

}  // namespace mmc
