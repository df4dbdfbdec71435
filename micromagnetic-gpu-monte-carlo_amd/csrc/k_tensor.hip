// k_tensor.hip -- FP64 Newell demag tensor with periodic-image sums
// (src/demag_cu.cu:119-236 arithmetic, one launch per slice of x images
// instead of src/main.cu:473-491's one launch + sync per image).
#include "mmc_device.h"

namespace mmc {

// Newell f, g (src/demag_cu.cu:16-44): same expression tree, EPSILON guards.
#define MMC_ABS(X) (((X) > 0) ? (X) : (-(X)))
constexpr double kEps = 1e-200;

__device__ __noinline__ double newell_f(double x, double y, double z)
{
    x = MMC_ABS(x);
    y = MMC_ABS(y);
    z = MMC_ABS(z);
    return +y / 2.0 * (z * z - x * x) * asinh(y / (sqrt(x * x + z * z) + kEps))
           + z / 2.0 * (y * y - x * x) * asinh(z / (sqrt(x * x + y * y) + kEps))
           - x * y * z * atan(y * z / (x * sqrt(x * x + y * y + z * z) + kEps))
           + 1.0 / 6.0 * (2 * x * x - y * y - z * z) * sqrt(x * x + y * y + z * z);
}

__device__ __noinline__ double newell_g(double x, double y, double z)
{
    z = MMC_ABS(z);
    return +x * y * z * asinh(z / (sqrt(x * x + y * y) + kEps))
           + y / 6.0 * (3.0 * z * z - y * y) * asinh(x / (sqrt(y * y + z * z) + kEps))
           + x / 6.0 * (3.0 * z * z - x * x) * asinh(y / (sqrt(x * x + z * z) + kEps))
           - z * z * z / 6.0 * atan(x * y / (z * sqrt(x * x + y * y + z * z) + kEps))
           - z * y * y / 2.0 * atan(x * z / (y * sqrt(x * x + y * y + z * z) + kEps))
           - z * x * x / 2.0 * atan(y * z / (x * sqrt(x * x + y * y + z * z) + kEps))
           - x * y * sqrt(x * x + y * y + z * z) / 3.0;
}

__device__ __forceinline__ int tensor_offset(int i, int n, int P, int periodic, int *off)
{
    if (periodic) {                         // src/demag_cu.cu:149-151
        *off = (i + (P + 1) / 2 - 1) % P - (P + 1) / 2 + 1;
        return 1;
    }
    if (i < n) { *off = i; return 1; }
    if (i > P - n) { *off = i - P; return 1; }
    *off = 0;
    return 0;
}

// Sum of one component over the 2^6 corner offsets in the reference's loop
// order (ikx, iky, ikz, ilx, ily, ilz), from the 27 distinct values F[].
// Identical additions to evaluating f/g once per term.
__device__ __forceinline__ double corner_sum(double acc, const double *F)
{
    #pragma unroll
    for (int ikx = 0; ikx < 2; ++ikx)
    #pragma unroll
    for (int iky = 0; iky < 2; ++iky)
    #pragma unroll
    for (int ikz = 0; ikz < 2; ++ikz)
    #pragma unroll
    for (int ilx = 0; ilx < 2; ++ilx)
    #pragma unroll
    for (int ily = 0; ily < 2; ++ily)
    #pragma unroll
    for (int ilz = 0; ilz < 2; ++ilz) {
        const int a = ikx - ilx + 1, b = iky - ily + 1, c = ikz - ilz + 1;
        const double t = F[a * 9 + b * 3 + c];
        const bool sign = ((ikx + iky + ikz + ilx + ily + ilz) % 2 == 0);
        acc += (sign ? t : -t);
    }
    return acc;
}

// One thread per padded cell; x images ix_lo..ix_hi (inclusive) of this
// launch.  Partial sums are kept unnormalised in K between launches so the
// addition order matches a single loop over all images.
__global__ __launch_bounds__(kThreads) void k_tensor(const Inst *__restrict__ I,
                                                     const Work *__restrict__ W,
                                                     double *__restrict__ Kb, int ix_lo,
                                                     int ix_hi, int first, int last)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const long id = (long)w.a + threadIdx.x;
    if (id >= in.NP) return;
    const int Px = in.Px, Py = in.Py, Pz = in.Pz;
    const int x = (int)(id % Px);
    const int y = (int)((id / Px) % Py);
    const int z = (int)(id / ((long)Px * Py));
    double *K = Kb + in.k_off;
    const long NP = in.NP;
    int ox, oy, oz;
    const int ux = tensor_offset(x, in.nx, Px, in.pbc[0], &ox);
    const int uy = tensor_offset(y, in.ny, Py, in.pbc[1], &oy);
    const int uz = tensor_offset(z, in.nz, Pz, in.pbc[2], &oz);
    if (!(ux && uy && uz)) {
        if (first)
            for (int c = 0; c < 6; ++c) K[c * NP + id] = 0.0;
        return;
    }
    const double dx = in.d[0], dy = in.d[1], dz = in.d[2];
    double acc[6];
    for (int c = 0; c < 6; ++c) acc[c] = first ? 0.0 : K[c * NP + id];
    const int py = in.img[1], pz = in.img[2];
    const int lo = max(ix_lo, -in.img[0]), hi = min(ix_hi, in.img[0]);
    for (int iximg = lo; iximg <= hi; ++iximg)
        for (int iyimg = -py; iyimg <= py; ++iyimg)
            for (int izimg = -pz; izimg <= pz; ++izimg) {
                const int ix = ox + iximg * Px;
                const int iy = oy + iyimg * Py;
                const int iz = oz + izimg * Pz;
                double F[27];
                #pragma unroll
                for (int comp = 0; comp < 6; ++comp) {
                    #pragma unroll
                    for (int a = 0; a < 3; ++a)
                    #pragma unroll
                    for (int b = 0; b < 3; ++b)
                    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double X = (ix + a - 1) * dx;
                        const double Y = (iy + b - 1) * dy;
                        const double Z = (iz + c - 1) * dz;
                        double v;
                        switch (comp) {   // src/demag_cu.cu:190-207
                        case 0: v = newell_f(X, Y, Z); break;
                        case 1: v = newell_g(X, Y, Z); break;
                        case 2: v = newell_g(X, Z, Y); break;
                        case 3: v = newell_f(Y, Z, X); break;
                        case 4: v = newell_g(Y, Z, X); break;
                        default: v = newell_f(Z, X, Y); break;
                        }
                        F[a * 9 + b * 3 + c] = v;
                    }
                    acc[comp] = corner_sum(acc[comp], F);
                }
            }
    if (last) {
        const double tmp = -3.14159265358 * 4.0 * dx * dy * dz;   // src/demag_cu.cu:221
        for (int c = 0; c < 6; ++c) K[c * NP + id] = acc[c] / tmp;
    } else {
        for (int c = 0; c < 6; ++c) K[c * NP + id] = acc[c];
    }
}

hipError_t launch_tensor(hipStream_t st, const Inst *I, const Work *W, int n, double *K,
                         int ix_lo, int ix_hi, int first, int last)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_tensor, dim3(n), dim3(kThreads), 0, st, I, W, K, ix_lo, ix_hi, first,
                       last);
    return post();
}


}  // namespace mmc
