/*
 * mmc_oracle.c -- FP64 CPU oracle for the Monte Carlo annealing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never linked
 * into it.  Loaded by tests/, by __graft_entry__.smoke() and by the CPU
 * baseline leg of bench.py.
 *
 * What each part follows (paths are into the reference repository):
 *   constants ............ src/constant.h:21-23, src/main.cu:244-245
 *   padding rule ......... src/main.cu:254-263 (open axes padded to a 2,3,5-
 *                          smooth length >= 2n-1 instead of exactly 2n-1;
 *                          the linear convolution over real cells is the same)
 *   Newell f, g .......... src/demag_cu.h:43-71 (host copies; device copies
 *                          src/demag_cu.cu:16-44 are identical)
 *   tensor assembly ...... src/demag_cu.cu:119-236 (index map :149-151,
 *                          image loops :157-166, 2^6 corner sum :170-207,
 *                          normalisation :221-228); image range uses the GPU
 *                          path's max(pbc,0) (src/main.cu:473-477)
 *   k-space product ...... src/demag_cu.cu:68-116 (H^ = K^ M^ / Np)
 *   energy sum ........... src/main.cu:945-947 (E = sum(e) * dx*dy*dz)
 *   kT ................... src/main.cu:1105 (k_B * T)
 * Routines whose bodies are absent from the reference upload and are
 * restated from the physics their signatures fix (src/effective_H.h:25-43,
 * src/mc_sampling.h:27-41, src/demag_mirror.h:18-24):
 *   exchange / interfacial DMI / uniaxial anisotropy / Zeeman field and
 *   energy density, the single-cell energy change (and its value for every
 *   cell's trial move, orc_trial_delta_e), the mirror (self) tensor and the
 *   Metropolis update.  Parity for those is pinned in tests/ by closed
 *   forms, invariants and statistics, not by reference output.
 */
#include "mmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI_TRUNC 3.14159265358     /* src/constant.h:21 (truncated) */
#define ORC_KB 1.3806488E-23           /* src/constant.h:22 */
#define ORC_EPSILON 1e-200             /* src/constant.h:23 */

struct orc_ctx {
    orc_params p;
    int n[3], P[3], pbc[3];
    long N, NP;
    double *K;      /* 6*NP real-space tensor */
    double *Khat;   /* 6*NP complex (interleaved), K^ / NP */
    double *m;      /* 3*N */
    double *h;      /* 3*N frozen demag field per unit Ms (sweep scratch) */
    int khat_valid;
};

/* ------------------------------------------------------------------ */
/* padding: src/main.cu:254-263                                         */
/* ------------------------------------------------------------------ */
static int is_smooth235(int v)
{
    while (v % 2 == 0) v /= 2;
    while (v % 3 == 0) v /= 3;
    while (v % 5 == 0) v /= 5;
    return v == 1;
}

int orc_padded_len(int n, int pbc)
{
    if (pbc != 0) return n;            /* periodic axis: no padding */
    int v = 2 * n - 1;                 /* reference pads to 2n-1 */
    if (v < 1) v = 1;
    while (!is_smooth235(v)) ++v;
    return v;
}

/* ------------------------------------------------------------------ */
/* Newell f, g: src/demag_cu.h:43-71.  ABS is the reference's ternary   */
/* macro (src/constant.h:17); EPSILON guards each denominator.          */
/* ------------------------------------------------------------------ */
#define ORC_ABS(X) (((X) > 0) ? (X) : (-(X)))

double orc_newell_f(double x, double y, double z)
{
    x = ORC_ABS(x);
    y = ORC_ABS(y);
    z = ORC_ABS(z);
    return +y / 2.0 * (z * z - x * x)
               * asinh(y / (sqrt(x * x + z * z) + ORC_EPSILON))
           + z / 2.0 * (y * y - x * x)
               * asinh(z / (sqrt(x * x + y * y) + ORC_EPSILON))
           - x * y * z
               * atan(y * z / (x * sqrt(x * x + y * y + z * z) + ORC_EPSILON))
           + 1.0 / 6.0 * (2 * x * x - y * y - z * z)
               * sqrt(x * x + y * y + z * z);
}

double orc_newell_g(double x, double y, double z)
{
    z = ORC_ABS(z);
    return +x * y * z * asinh(z / (sqrt(x * x + y * y) + ORC_EPSILON))
           + y / 6.0 * (3.0 * z * z - y * y)
               * asinh(x / (sqrt(y * y + z * z) + ORC_EPSILON))
           + x / 6.0 * (3.0 * z * z - x * x)
               * asinh(y / (sqrt(x * x + z * z) + ORC_EPSILON))
           - z * z * z / 6.0
               * atan(x * y / (z * sqrt(x * x + y * y + z * z) + ORC_EPSILON))
           - z * y * y / 2.0
               * atan(x * z / (y * sqrt(x * x + y * y + z * z) + ORC_EPSILON))
           - z * x * x / 2.0
               * atan(y * z / (x * sqrt(x * x + y * y + z * z) + ORC_EPSILON))
           - x * y * sqrt(x * x + y * y + z * z) / 3.0;
}

/* Offset (in cells) that padded index i stands for.  Periodic axes use the
 * reference's map idx = (i + (P+1)/2 - 1) % P - (P+1)/2 + 1
 * (src/demag_cu.cu:149-151); open axes hold offsets -(n-1)..(n-1) and the
 * extra slots of the longer padding are unused (returns 0). */
int orc_tensor_offset(int i, int n, int P, int pbc, int *off)
{
    if (pbc != 0) {
        *off = (i + (P + 1) / 2 - 1) % P - (P + 1) / 2 + 1;
        return 1;
    }
    if (i < n) { *off = i; return 1; }
    if (i > P - n) { *off = i - P; return 1; }
    *off = 0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11; Random123 published algorithm). */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 23 random bits -> (2k+1)/2^24 in (0,1); exact in FP32 and FP64. */
double orc_u01(uint32_t r)
{
    return (double)(2u * (r >> 9) + 1u) * (1.0 / 16777216.0);
}

/* ------------------------------------------------------------------ */
/* 1-D mixed-radix complex FFT (recursive decimation in time).          */
/* Forward uses exp(-2 pi i jk/n); the inverse is unscaled.             */
/* ------------------------------------------------------------------ */
static int smallest_factor(int n)
{
    if (n % 2 == 0) return 2;
    for (int f = 3; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

static void fft_rec(const double *in, double *out, int n, long is,
                    const double *tw, int N, int tws, double *tmp)
{
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; ++r)
        fft_rec(in + 2 * r * is, out + 2 * (long)r * m, m, is * p, tw, N,
                tws * p, tmp);
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) {
            double yr = out[2 * ((long)r * m + k)];
            double yi = out[2 * ((long)r * m + k) + 1];
            long e = ((long)r * k * tws) % N;
            double wr = tw[2 * e], wi = tw[2 * e + 1];
            tmp[2 * r] = yr * wr - yi * wi;
            tmp[2 * r + 1] = yr * wi + yi * wr;
        }
        for (int q = 0; q < p; ++q) {
            double sr = 0.0, si = 0.0;
            for (int r = 0; r < p; ++r) {
                long e = ((long)((r * q) % p) * m * tws) % N;
                double wr = tw[2 * e], wi = tw[2 * e + 1];
                sr += tmp[2 * r] * wr - tmp[2 * r + 1] * wi;
                si += tmp[2 * r] * wi + tmp[2 * r + 1] * wr;
            }
            tmp[2 * (p + q)] = sr;
            tmp[2 * (p + q) + 1] = si;
        }
        for (int q = 0; q < p; ++q) {
            out[2 * (k + (long)m * q)] = tmp[2 * (p + q)];
            out[2 * (k + (long)m * q) + 1] = tmp[2 * (p + q) + 1];
        }
    }
}

int orc_fft1d(double *data, int n, int inverse)
{
    if (n <= 0) return -1;
    double *tw = (double *)malloc(sizeof(double) * 2 * n);
    double *in = (double *)malloc(sizeof(double) * 2 * n);
    double *tmp = (double *)malloc(sizeof(double) * 4 * (n + 1));
    if (!tw || !in || !tmp) { free(tw); free(in); free(tmp); return -2; }
    for (int t = 0; t < n; ++t) {
        double a = -2.0 * M_PI * (double)t / (double)n;
        tw[2 * t] = cos(a);
        tw[2 * t + 1] = sin(a);
    }
    for (int t = 0; t < n; ++t) {
        in[2 * t] = data[2 * t];
        in[2 * t + 1] = inverse ? -data[2 * t + 1] : data[2 * t + 1];
    }
    fft_rec(in, data, n, 1, tw, n, 1, tmp);
    if (inverse)
        for (int t = 0; t < n; ++t) data[2 * t + 1] = -data[2 * t + 1];
    free(tw); free(in); free(tmp);
    return 0;
}

/* 3-D transform of an x-fastest complex array, one axis at a time. */
static void fft3d(double *a, const int P[3], int inverse)
{
    long sx = 1, sy = P[0], sz = (long)P[0] * P[1];
    long strides[3] = {sx, sy, sz};
    for (int ax = 0; ax < 3; ++ax) {
        int n = P[ax];
        if (n == 1) continue;
        long lines = (long)P[0] * P[1] * P[2] / n;
        #pragma omp parallel
        {
            double *buf = (double *)malloc(sizeof(double) * 2 * n);
            #pragma omp for schedule(static)
            for (long l = 0; l < lines; ++l) {
                /* base index of line l: enumerate the other two axes */
                long base;
                if (ax == 0) base = l * P[0];
                else if (ax == 1) base = (l % P[0]) + (l / P[0]) * sz;
                else base = l;
                long st = strides[ax];
                for (int t = 0; t < n; ++t) {
                    buf[2 * t] = a[2 * (base + t * st)];
                    buf[2 * t + 1] = a[2 * (base + t * st) + 1];
                }
                orc_fft1d(buf, n, inverse);
                for (int t = 0; t < n; ++t) {
                    a[2 * (base + t * st)] = buf[2 * t];
                    a[2 * (base + t * st) + 1] = buf[2 * t + 1];
                }
            }
            free(buf);
        }
    }
}

/* ------------------------------------------------------------------ */
/* context                                                              */
/* ------------------------------------------------------------------ */
orc_ctx *orc_create(const orc_params *p)
{
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    if (!c) return NULL;
    c->p = *p;
    c->n[0] = p->nx; c->n[1] = p->ny; c->n[2] = p->nz;
    c->pbc[0] = p->pbc_x; c->pbc[1] = p->pbc_y; c->pbc[2] = p->pbc_z;
    for (int a = 0; a < 3; ++a) c->P[a] = orc_padded_len(c->n[a], c->pbc[a]);
    c->N = (long)p->nx * p->ny * p->nz;
    c->NP = (long)c->P[0] * c->P[1] * c->P[2];
    c->K = (double *)calloc(6 * c->NP, sizeof(double));
    c->Khat = (double *)calloc(12 * c->NP, sizeof(double));
    c->m = (double *)calloc(3 * c->N, sizeof(double));
    c->h = (double *)calloc(3 * c->N, sizeof(double));
    if (!c->K || !c->Khat || !c->m || !c->h) { orc_destroy(c); return NULL; }
    for (long i = 0; i < c->N; ++i) c->m[2 * c->N + i] = 1.0;
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    free(c->K); free(c->Khat); free(c->m); free(c->h);
    free(c);
}

void orc_dims(const orc_ctx *c, int out[6])
{
    for (int a = 0; a < 3; ++a) { out[a] = c->n[a]; out[3 + a] = c->P[a]; }
}

/* src/demag_cu.cu:119-236, one padded cell at a time, same loop order. */
void orc_build_tensor(orc_ctx *c)
{
    const int Px = c->P[0], Py = c->P[1], Pz = c->P[2];
    const double dx = c->p.dx, dy = c->p.dy, dz = c->p.dz;
    const int px = c->pbc[0] > 0 ? c->pbc[0] : 0;   /* src/main.cu:475-477 */
    const int py = c->pbc[1] > 0 ? c->pbc[1] : 0;
    const int pz = c->pbc[2] > 0 ? c->pbc[2] : 0;
    const long NP = c->NP;
    double *K = c->K;
    #pragma omp parallel for schedule(dynamic, 16) collapse(2)
    for (int z = 0; z < Pz; ++z)
        for (int y = 0; y < Py; ++y)
            for (int x = 0; x < Px; ++x) {
                long id = x + (long)Px * (y + (long)Py * z);
                int ox, oy, oz;
                int ux = orc_tensor_offset(x, c->n[0], Px, c->pbc[0], &ox);
                int uy = orc_tensor_offset(y, c->n[1], Py, c->pbc[1], &oy);
                int uz = orc_tensor_offset(z, c->n[2], Pz, c->pbc[2], &oz);
                if (!(ux && uy && uz)) {
                    for (int k = 0; k < 6; ++k) K[k * NP + id] = 0.0;
                    continue;
                }
                double kxx = 0, kxy = 0, kxz = 0, kyy = 0, kyz = 0, kzz = 0;
                for (int ix_img = -px; ix_img <= px; ++ix_img)
                for (int iy_img = -py; iy_img <= py; ++iy_img)
                for (int iz_img = -pz; iz_img <= pz; ++iz_img) {
                    int ix = ox + ix_img * Px;
                    int iy = oy + iy_img * Py;
                    int iz = oz + iz_img * Pz;
                    for (int ikx = 0; ikx < 2; ++ikx)
                    for (int iky = 0; iky < 2; ++iky)
                    for (int ikz = 0; ikz < 2; ++ikz)
                    for (int ilx = 0; ilx < 2; ++ilx)
                    for (int ily = 0; ily < 2; ++ily)
                    for (int ilz = 0; ilz < 2; ++ilz) {
                        int iix = ix + ikx - ilx;
                        int iiy = iy + iky - ily;
                        int iiz = iz + ikz - ilz;
                        int sign = ((ikx + iky + ikz + ilx + ily + ilz) % 2 == 0);
                        double tmp;
                        tmp = orc_newell_f(iix * dx, iiy * dy, iiz * dz);
                        kxx += (sign ? tmp : -tmp);
                        tmp = orc_newell_g(iix * dx, iiy * dy, iiz * dz);
                        kxy += (sign ? tmp : -tmp);
                        tmp = orc_newell_g(iix * dx, iiz * dz, iiy * dy);
                        kxz += (sign ? tmp : -tmp);
                        tmp = orc_newell_f(iiy * dy, iiz * dz, iix * dx);
                        kyy += (sign ? tmp : -tmp);
                        tmp = orc_newell_g(iiy * dy, iiz * dz, iix * dx);
                        kyz += (sign ? tmp : -tmp);
                        tmp = orc_newell_f(iiz * dz, iix * dx, iiy * dy);
                        kzz += (sign ? tmp : -tmp);
                    }
                }
                double tmp = -ORC_PI_TRUNC * 4.0 * dx * dy * dz;
                K[0 * NP + id] = kxx / tmp;
                K[1 * NP + id] = kxy / tmp;
                K[2 * NP + id] = kxz / tmp;
                K[3 * NP + id] = kyy / tmp;
                K[4 * NP + id] = kyz / tmp;
                K[5 * NP + id] = kzz / tmp;
            }
    c->khat_valid = 0;
}

void orc_set_tensor(orc_ctx *c, const double *K)
{
    memcpy(c->K, K, sizeof(double) * 6 * c->NP);
    c->khat_valid = 0;
}

void orc_get_tensor(const orc_ctx *c, double *K)
{
    memcpy(K, c->K, sizeof(double) * 6 * c->NP);
}

/* Mirror (self) tensor: the padded tensor's zero-offset entry, which holds
 * the cell's self term plus its own periodic images (the six scalars
 * src/main.cu:339-346 passes to the Metropolis kernel). */
void orc_mirror(const orc_ctx *c, double K0[6])
{
    for (int k = 0; k < 6; ++k) K0[k] = c->K[k * c->NP];
}

void orc_set_m(orc_ctx *c, const double *m)
{
    memcpy(c->m, m, sizeof(double) * 3 * c->N);
}

void orc_get_m(const orc_ctx *c, double *m)
{
    memcpy(m, c->m, sizeof(double) * 3 * c->N);
}

/* Uniform-on-sphere start state from Philox (counter word 3 = 1). */
void orc_init_random_m(orc_ctx *c, unsigned long long seed)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (long i = 0; i < c->N; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, 0u, 0u, 1u}, r[4];
        orc_philox4x32_10(ctr, key, r);
        double z = 1.0 - 2.0 * orc_u01(r[0]);
        double phi = 2.0 * M_PI * orc_u01(r[1]);
        double s = sqrt(fmax(0.0, 1.0 - z * z));
        c->m[i] = s * cos(phi);
        c->m[c->N + i] = s * sin(phi);
        c->m[2 * c->N + i] = z;
    }
}

/* ------------------------------------------------------------------ */
/* demag: h = K * m (per unit Ms); src/demag_cu.cu:68-116 product       */
/* ------------------------------------------------------------------ */
static void ensure_khat(orc_ctx *c)
{
    if (c->khat_valid) return;
    const long NP = c->NP;
    for (int k = 0; k < 6; ++k) {
        double *dst = c->Khat + 2 * k * NP;
        for (long i = 0; i < NP; ++i) {
            dst[2 * i] = c->K[k * NP + i];
            dst[2 * i + 1] = 0.0;
        }
        fft3d(dst, c->P, 0);
        for (long i = 0; i < 2 * NP; ++i) dst[i] /= (double)NP;
    }
    c->khat_valid = 1;
}

void orc_demag_fft(orc_ctx *c, double *h)
{
    ensure_khat(c);
    const double *Kh = c->Khat;
    const long NP = c->NP, N = c->N;
    const int nx = c->n[0], ny = c->n[1], nz = c->n[2];
    const int Px = c->P[0], Py = c->P[1];
    double *Mh = (double *)calloc(6 * NP, sizeof(double));
    double *Hh = (double *)calloc(6 * NP, sizeof(double));
    for (int k = 0; k < 3; ++k) {
        double *dst = Mh + 2 * k * NP;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x)
                    dst[2 * (x + (long)Px * (y + (long)Py * z))] =
                        c->m[k * N + x + (long)nx * (y + (long)ny * z)];
        fft3d(dst, c->P, 0);
    }
    /* symmetric 3x3 complex product, src/demag_cu.cu:98-111 */
    static const int rowk[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < NP; ++i) {
        for (int a = 0; a < 3; ++a) {
            double sr = 0, si = 0;
            for (int b = 0; b < 3; ++b) {
                int kk = rowk[a][b];
                double mr = Mh[2 * (b * NP + i)], mi = Mh[2 * (b * NP + i) + 1];
                double kr = Kh[2 * (kk * NP + i)], ki = Kh[2 * (kk * NP + i) + 1];
                sr += mr * kr - mi * ki;
                si += mr * ki + mi * kr;
            }
            Hh[2 * (a * NP + i)] = sr;
            Hh[2 * (a * NP + i) + 1] = si;
        }
    }
    for (int a = 0; a < 3; ++a) {
        double *src = Hh + 2 * a * NP;
        fft3d(src, c->P, 1);
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x)
                    h[a * N + x + (long)nx * (y + (long)ny * z)] =
                        src[2 * (x + (long)Px * (y + (long)Py * z))];
    }
    free(Mh);
    free(Hh);
}

/* Brute-force circular sum over real cells (tiny grids only). */
void orc_demag_direct(orc_ctx *c, double *h)
{
    const long N = c->N, NP = c->NP;
    const int nx = c->n[0], ny = c->n[1];
    const int Px = c->P[0], Py = c->P[1], Pz = c->P[2];
    static const int rowk[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        int xi = (int)(i % nx), yi = (int)((i / nx) % ny), zi = (int)(i / ((long)nx * ny));
        long double acc[3] = {0, 0, 0};
        for (long j = 0; j < N; ++j) {
            int xj = (int)(j % nx), yj = (int)((j / nx) % ny), zj = (int)(j / ((long)nx * ny));
            int ix = ((xi - xj) % Px + Px) % Px;
            int iy = ((yi - yj) % Py + Py) % Py;
            int iz = ((zi - zj) % Pz + Pz) % Pz;
            long kid = ix + (long)Px * (iy + (long)Py * iz);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    acc[a] += (long double)c->K[rowk[a][b] * NP + kid] * c->m[b * N + j];
        }
        for (int a = 0; a < 3; ++a) h[a * N + i] = (double)acc[a];
    }
}

/* ------------------------------------------------------------------ */
/* local fields (A/m).  Restated physics, see file header.              */
/* ------------------------------------------------------------------ */
static double mu0_of(void) { return 4e-7 * ORC_PI_TRUNC; }  /* src/main.cu:244 */

/* neighbour index along one axis, -1 when the bond does not exist */
static int nbr(int i, int d, int n, int pbc)
{
    int j = i + d;
    if (j < 0 || j >= n) {
        if (pbc == 0) return -1;
        j = (j + n) % n;
    }
    if (j == i) return -1;            /* a cell is never its own neighbour */
    return j;
}

static void unit_axis(const orc_params *p, double u[3])
{
    double l = sqrt(p->anis_ux * p->anis_ux + p->anis_uy * p->anis_uy +
                    p->anis_uz * p->anis_uz);
    if (l > 0) { u[0] = p->anis_ux / l; u[1] = p->anis_uy / l; u[2] = p->anis_uz / l; }
    else { u[0] = u[1] = u[2] = 0.0; }
}

/* Field of one term at cell (x,y,z).  Exchange: Laplacian form
 * (2A/mu0 Ms) sum_j (m_j - m_i)/D^2 over existing bonds; DMI: interfacial
 * (normal z), bond derivative of
 *   E = sum_{x-bonds} D V/dx (m_i,z m_j,x - m_i,x m_j,z)
 *     + sum_{y-bonds} D V/dy (m_i,z m_j,y - m_i,y m_j,z);
 * anisotropy (2Ku/mu0 Ms)(m.u)u; Zeeman B/mu0. */
static void local_term(const orc_ctx *c, int term, int x, int y, int z,
                       double H[3])
{
    const orc_params *p = &c->p;
    const int nx = c->n[0], ny = c->n[1];
    const long N = c->N;
    const double *m = c->m;
    const double mu0 = mu0_of();
    long i = x + (long)nx * (y + (long)ny * z);
    H[0] = H[1] = H[2] = 0.0;
    if (term == ORC_EXCH) {
        double fac = 2.0 * p->aex / mu0 / p->ms;       /* src/main.cu:245 */
        double inv2[3] = {1.0 / (p->dx * p->dx), 1.0 / (p->dy * p->dy),
                          1.0 / (p->dz * p->dz)};
        int co[3] = {x, y, z};
        for (int a = 0; a < 3; ++a)
            for (int d = -1; d <= 1; d += 2) {
                int j = nbr(co[a], d, c->n[a], c->pbc[a]);
                if (j < 0) continue;
                int cc[3] = {x, y, z};
                cc[a] = j;
                long jj = cc[0] + (long)nx * (cc[1] + (long)ny * cc[2]);
                for (int k = 0; k < 3; ++k)
                    H[k] += fac * inv2[a] * (m[k * N + jj] - m[k * N + i]);
            }
    } else if (term == ORC_DMI) {
        double fac = p->dind / mu0 / p->ms;
        /* x bonds */
        int jp = nbr(x, +1, nx, c->pbc[0]), jm = nbr(x, -1, nx, c->pbc[0]);
        double pzv = 0, pxv = 0, mzv = 0, mxv = 0;
        if (jp >= 0) { long jj = jp + (long)nx * (y + (long)ny * z); pzv = m[2 * N + jj]; pxv = m[jj]; }
        if (jm >= 0) { long jj = jm + (long)nx * (y + (long)ny * z); mzv = m[2 * N + jj]; mxv = m[jj]; }
        H[0] += fac / p->dx * (pzv - mzv);
        H[2] -= fac / p->dx * (pxv - mxv);
        /* y bonds */
        jp = nbr(y, +1, ny, c->pbc[1]); jm = nbr(y, -1, ny, c->pbc[1]);
        double pz2 = 0, py2 = 0, mz2 = 0, my2 = 0;
        if (jp >= 0) { long jj = x + (long)nx * (jp + (long)ny * z); pz2 = m[2 * N + jj]; py2 = m[N + jj]; }
        if (jm >= 0) { long jj = x + (long)nx * (jm + (long)ny * z); mz2 = m[2 * N + jj]; my2 = m[N + jj]; }
        H[1] += fac / p->dy * (pz2 - mz2);
        H[2] -= fac / p->dy * (py2 - my2);
    } else if (term == ORC_ANIS) {
        double u[3];
        unit_axis(p, u);
        double mu = m[i] * u[0] + m[N + i] * u[1] + m[2 * N + i] * u[2];
        double fac = 2.0 * p->ku1 / mu0 / p->ms;
        for (int k = 0; k < 3; ++k) H[k] = fac * mu * u[k];
    } else if (term == ORC_ZEEMAN) {
        H[0] = p->bext_x / mu0;
        H[1] = p->bext_y / mu0;
        H[2] = p->bext_z / mu0;
    }
}

void orc_field(orc_ctx *c, int terms, double *H)
{
    const long N = c->N;
    double *h = NULL;
    if ((terms & ORC_DEMAG) && c->p.cal_demag) {
        h = (double *)malloc(sizeof(double) * 3 * N);
        orc_demag_fft(c, h);
    }
    const int nx = c->n[0], ny = c->n[1];
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
        double acc[3] = {0, 0, 0}, t[3];
        for (int term = ORC_EXCH; term <= ORC_ZEEMAN; term <<= 1) {
            if (!(terms & term)) continue;
            local_term(c, term, x, y, z, t);
            for (int k = 0; k < 3; ++k) acc[k] += t[k];
        }
        if (h)
            for (int k = 0; k < 3; ++k) acc[k] += c->p.ms * h[k * N + i];
        for (int k = 0; k < 3; ++k) H[k * N + i] = acc[k];
    }
    free(h);
}

/* Per-cell energy density (J/m^3) of the selected terms:
 * e = -1/2 mu0 Ms m.H for the bilinear terms, -Ms m.B for Zeeman. */
void orc_energy_density(orc_ctx *c, int terms, double *e)
{
    const long N = c->N;
    const double mu0 = mu0_of();
    double *h = NULL;
    if ((terms & ORC_DEMAG) && c->p.cal_demag) {
        h = (double *)malloc(sizeof(double) * 3 * N);
        orc_demag_fft(c, h);
    }
    const int nx = c->n[0], ny = c->n[1];
    for (long i = 0; i < N; ++i) {
        int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
        double mi[3] = {c->m[i], c->m[N + i], c->m[2 * N + i]};
        double acc = 0.0, t[3];
        for (int term = ORC_EXCH; term <= ORC_ZEEMAN; term <<= 1) {
            if (!(terms & term)) continue;
            local_term(c, term, x, y, z, t);
            double w = (term == ORC_ZEEMAN) ? -mu0 * c->p.ms : -0.5 * mu0 * c->p.ms;
            acc += w * (mi[0] * t[0] + mi[1] * t[1] + mi[2] * t[2]);
        }
        if (h) {
            double md = mi[0] * h[i] + mi[1] * h[N + i] + mi[2] * h[2 * N + i];
            acc += -0.5 * mu0 * c->p.ms * c->p.ms * md;
        }
        e[i] = acc;
    }
    free(h);
}

/* out = {E_total, E_exch, E_dmi, E_anis, E_zeeman, E_demag, <mx>, <my>, <mz>}.
 * Each term: (sum of densities) * dx * dy * dz, src/main.cu:947. */
void orc_energy(orc_ctx *c, double out[9])
{
    const long N = c->N;
    double *e = (double *)malloc(sizeof(double) * N);
    double tot = 0.0;
    for (int t = 0; t < 5; ++t) {
        orc_energy_density(c, 1 << t, e);
        long double s = 0.0L;
        for (long i = 0; i < N; ++i) s += e[i];
        double E = (double)s * c->p.dx * c->p.dy * c->p.dz;
        out[1 + t] = E;
        tot += E;
    }
    out[0] = tot;
    for (int k = 0; k < 3; ++k) {
        long double s = 0.0L;
        for (long i = 0; i < N; ++i) s += c->m[k * N + i];
        out[6 + k] = (double)(s / (long double)N);
    }
    free(e);
}

/* Energy change (J) of replacing m_i by mnew with the demag field h (per
 * unit Ms) frozen:  dE = -mu0 Ms V d.G - Ku V[(m'.u)^2-(m.u)^2]
 *                        - 1/2 mu0 Ms^2 V [m'K0m' - mK0m],
 * G = (2A/mu0Ms) sum_j m_j/D^2 + H_dmi + B/mu0 + Ms (h_i - K0 m_i).
 * The inputs are those of mcKernel2 (src/mc_sampling.h:27-41). */
double orc_delta_e(orc_ctx *c, const double *h, int cell, const double mnew[3])
{
    const orc_params *p = &c->p;
    const long N = c->N;
    const int nx = c->n[0], ny = c->n[1];
    const int x = cell % nx, y = (cell / nx) % ny, z = cell / (nx * ny);
    const double mu0 = mu0_of();
    const double V = p->dx * p->dy * p->dz;
    double mi[3] = {c->m[cell], c->m[N + cell], c->m[2 * N + cell]};
    double d[3] = {mnew[0] - mi[0], mnew[1] - mi[1], mnew[2] - mi[2]};
    double G[3] = {0, 0, 0};
    /* exchange: neighbours only */
    {
        double fac = 2.0 * p->aex / mu0 / p->ms;
        double inv2[3] = {1.0 / (p->dx * p->dx), 1.0 / (p->dy * p->dy),
                          1.0 / (p->dz * p->dz)};
        int co[3] = {x, y, z};
        for (int a = 0; a < 3; ++a)
            for (int s = -1; s <= 1; s += 2) {
                int j = nbr(co[a], s, c->n[a], c->pbc[a]);
                if (j < 0) continue;
                int cc[3] = {x, y, z};
                cc[a] = j;
                long jj = cc[0] + (long)nx * (cc[1] + (long)ny * cc[2]);
                for (int k = 0; k < 3; ++k) G[k] += fac * inv2[a] * c->m[k * N + jj];
            }
    }
    double t[3];
    local_term(c, ORC_DMI, x, y, z, t);
    for (int k = 0; k < 3; ++k) G[k] += t[k];
    G[0] += p->bext_x / mu0;
    G[1] += p->bext_y / mu0;
    G[2] += p->bext_z / mu0;
    double K0[6] = {0, 0, 0, 0, 0, 0};
    if (p->cal_demag) {
        orc_mirror(c, K0);
        double k0m[3] = {K0[0] * mi[0] + K0[1] * mi[1] + K0[2] * mi[2],
                         K0[1] * mi[0] + K0[3] * mi[1] + K0[4] * mi[2],
                         K0[2] * mi[0] + K0[4] * mi[1] + K0[5] * mi[2]};
        for (int k = 0; k < 3; ++k) G[k] += p->ms * (h[k * N + cell] - k0m[k]);
    }
    double dE = -mu0 * p->ms * V * (d[0] * G[0] + d[1] * G[1] + d[2] * G[2]);
    double u[3];
    unit_axis(p, u);
    double mu_new = mnew[0] * u[0] + mnew[1] * u[1] + mnew[2] * u[2];
    double mu_old = mi[0] * u[0] + mi[1] * u[1] + mi[2] * u[2];
    dE -= p->ku1 * V * (mu_new * mu_new - mu_old * mu_old);
    if (p->cal_demag) {
        double qn = K0[0] * mnew[0] * mnew[0] + K0[3] * mnew[1] * mnew[1] +
                    K0[5] * mnew[2] * mnew[2] +
                    2.0 * (K0[1] * mnew[0] * mnew[1] + K0[2] * mnew[0] * mnew[2] +
                           K0[4] * mnew[1] * mnew[2]);
        double qo = K0[0] * mi[0] * mi[0] + K0[3] * mi[1] * mi[1] +
                    K0[5] * mi[2] * mi[2] +
                    2.0 * (K0[1] * mi[0] * mi[1] + K0[2] * mi[0] * mi[2] +
                           K0[4] * mi[1] * mi[2]);
        dE -= 0.5 * mu0 * p->ms * p->ms * V * (qn - qo);
    }
    return dE;
}

/* ------------------------------------------------------------------ */
/* Metropolis sweep                                                     */
/* ------------------------------------------------------------------ */
int orc_num_colors(const orc_ctx *c)
{
    for (int a = 0; a < 3; ++a)
        if (c->pbc[a] != 0 && c->n[a] >= 3 && (c->n[a] & 1)) return 3;
    return 2;
}

static int axis_color(int i, int n, int pbc)
{
    if (pbc != 0 && n >= 3 && (n & 1) && i == n - 1) return 2;
    return i & 1;
}

/* Proper colouring of the cell graph: (c_x + c_y + c_z) mod K. */
int orc_cell_color(const orc_ctx *c, int x, int y, int z)
{
    int K = orc_num_colors(c);
    return (axis_color(x, c->n[0], c->pbc[0]) + axis_color(y, c->n[1], c->pbc[1]) +
            axis_color(z, c->n[2], c->pbc[2])) % K;
}

/* One sweep: demag frozen at the pre-sweep state (src/main.cu:868-884 then
 * :1101), then every colour class in order.  Trial direction uniform on the
 * sphere, accept if dE <= 0 or u < exp(-dE / kB T).  Philox key = seed,
 * counter = (cell, sweep, stage, 0).  Returns accepted moves. */
/* Energy change of every cell's trial move for (stage, sweep) against the
 * current state, moves not applied: the dE that orc_sweep tests (the
 * quantity mcKernel2 forms from its inputs, src/mc_sampling.h:27-41). */
void orc_trial_delta_e(orc_ctx *c, uint32_t stage, uint32_t sweep, double *dE)
{
    const long N = c->N;
    uint32_t key[2] = {(uint32_t)c->p.seed, (uint32_t)(c->p.seed >> 32)};
    if (c->p.cal_demag) orc_demag_fft(c, c->h);
    else memset(c->h, 0, sizeof(double) * 3 * N);
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, sweep, stage, 0u}, r[4];
        orc_philox4x32_10(ctr, key, r);
        double zz = 1.0 - 2.0 * orc_u01(r[0]);
        double phi = 2.0 * M_PI * orc_u01(r[1]);
        double s = sqrt(fmax(0.0, 1.0 - zz * zz));
        double mn[3] = {s * cos(phi), s * sin(phi), zz};
        dE[i] = orc_delta_e(c, c->h, (int)i, mn);
    }
}

long orc_sweep(orc_ctx *c, double T, uint32_t stage, uint32_t sweep)
{
    const long N = c->N;
    const int nx = c->n[0], ny = c->n[1], nz = c->n[2];
    const double kT = ORC_KB * T;                  /* src/main.cu:1105 */
    uint32_t key[2] = {(uint32_t)c->p.seed, (uint32_t)(c->p.seed >> 32)};
    if (c->p.cal_demag) orc_demag_fft(c, c->h);
    else memset(c->h, 0, sizeof(double) * 3 * N);
    const int K = orc_num_colors(c);
    long acc = 0;
    for (int col = 0; col < K; ++col) {
        #pragma omp parallel for schedule(static) reduction(+ : acc)
        for (long i = 0; i < N; ++i) {
            int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
            if (orc_cell_color(c, x, y, z) != col) continue;
            uint32_t ctr[4] = {(uint32_t)i, sweep, stage, 0u}, r[4];
            orc_philox4x32_10(ctr, key, r);
            double zz = 1.0 - 2.0 * orc_u01(r[0]);
            double phi = 2.0 * M_PI * orc_u01(r[1]);
            double s = sqrt(fmax(0.0, 1.0 - zz * zz));
            double mn[3] = {s * cos(phi), s * sin(phi), zz};
            double dE = orc_delta_e(c, c->h, (int)i, mn);
            if (dE <= 0.0 || orc_u01(r[2]) < exp(-dE / kT)) {
                c->m[i] = mn[0];
                c->m[N + i] = mn[1];
                c->m[2 * N + i] = mn[2];
                ++acc;
            }
        }
    }
    (void)nz;
    return acc;
}
