/*
 * mmc_oracle.h -- FP64 CPU oracle for the Monte Carlo annealing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by the smoke check and by the CPU-baseline leg of the benchmark, and
 * never by the product library.  It restates, in plain C, the arithmetic of
 * the reference (chg-hou/Micromagnetic-GPU-Monte-Carlo); every function cites
 * the reference file:line it follows.  Where the reference's upload has no
 * body for a routine (field, energy, Metropolis, mirror tensor), the oracle
 * restates the physics the reference's kernel signatures fix; those are
 * pinned by closed forms and statistics in tests/, not by reference output.
 *
 * Conventions shared with the product ABI (include/mmc.h):
 *   - magnetisation arrays are SoA, x fastest: idx = x + nx*(y + ny*z);
 *   - the real-space demag tensor is 6 planes [xx,xy,xz,yy,yz,zz] over the
 *     padded grid (Px,Py,Pz), x fastest;
 *   - fields are in A/m, energies in J.
 */
#ifndef MMC_ORACLE_H
#define MMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order as mmc_params in include/mmc.h. */
typedef struct orc_params {
    int nx, ny, nz;
    double dx, dy, dz;
    int pbc_x, pbc_y, pbc_z;
    double ms, aex, dind;
    double anis_ux, anis_uy, anis_uz, ku1;
    double bext_x, bext_y, bext_z;
    int cal_demag;
    int reserved;
    unsigned long long seed;
} orc_params;

typedef struct orc_ctx orc_ctx;

/* Term bits for field/energy selection. */
enum {
    ORC_EXCH = 1, ORC_DMI = 2, ORC_ANIS = 4, ORC_ZEEMAN = 8, ORC_DEMAG = 16,
    ORC_ALL = 31
};

/* ---- scalar pieces ---------------------------------------------------- */
int    orc_padded_len(int n, int pbc);
double orc_newell_f(double x, double y, double z);
double orc_newell_g(double x, double y, double z);
int    orc_tensor_offset(int i, int n, int P, int pbc, int *off);
void   orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                         uint32_t out[4]);
double orc_u01(uint32_t r);

/* ---- 1-D complex FFT (interleaved re,im), forward sign -1 ------------- */
int    orc_fft1d(double *data, int n, int inverse);

/* ---- context ---------------------------------------------------------- */
orc_ctx *orc_create(const orc_params *p);
void     orc_destroy(orc_ctx *c);
void     orc_dims(const orc_ctx *c, int out[6]);   /* nx,ny,nz,Px,Py,Pz */
void     orc_build_tensor(orc_ctx *c);
void     orc_set_tensor(orc_ctx *c, const double *K);
void     orc_get_tensor(const orc_ctx *c, double *K);
void     orc_mirror(const orc_ctx *c, double K0[6]);
void     orc_set_m(orc_ctx *c, const double *m);        /* 3*N */
void     orc_get_m(const orc_ctx *c, double *m);
void     orc_init_random_m(orc_ctx *c, unsigned long long seed);
void     orc_demag_fft(orc_ctx *c, double *h);           /* 3*N, per unit Ms */
void     orc_demag_direct(orc_ctx *c, double *h);
void     orc_field(orc_ctx *c, int terms, double *H);    /* 3*N, A/m */
void     orc_energy(orc_ctx *c, double out[9]);          /* E, 5 terms, <m> */
void     orc_energy_density(orc_ctx *c, int terms, double *e); /* N, J/m^3 */
double   orc_delta_e(orc_ctx *c, const double *h, int cell,
                     const double mnew[3]);
void     orc_trial_delta_e(orc_ctx *c, uint32_t stage, uint32_t sweep,
                           double *dE);                  /* N, J */
long     orc_sweep(orc_ctx *c, double T, uint32_t stage, uint32_t sweep);
int      orc_num_colors(const orc_ctx *c);
int      orc_cell_color(const orc_ctx *c, int x, int y, int z);

#ifdef __cplusplus
}
#endif
#endif
