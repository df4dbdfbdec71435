/*
 * mmc.h -- C ABI of the MI355X Monte Carlo annealing library (libmmc.so).
 *
 * The reference (chg-hou/Micromagnetic-GPU-Monte-Carlo) has no library API:
 * its main() launches __global__ kernels with raw device pointers and scalar
 * physics arguments on the default stream.  Each entry point below replaces
 * one such launch sequence; the comment names the reference code it stands
 * in for (paths relative to the reference repository).
 *
 * Conventions
 *   - every function returns MMC_OK (0) or an MMC_E* code and never exits;
 *     mmc_last_error() returns the message of the last failure;
 *   - host arrays are caller-owned, structure-of-arrays, x fastest
 *     (OOMMF OVF order): index = x + nx*(y + ny*z);
 *   - the real-space demag tensor is 6 planes [xx,xy,xz,yy,yz,zz] over the
 *     padded grid (Px,Py,Pz), x fastest;
 *   - fields are in A/m, energies in J, temperatures in K;
 *   - a context is not thread-safe; distinct contexts are independent.
 */
#ifndef MMC_H
#define MMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical and numerical parameters of one instance.  Field names follow the
 * reference's ini keys (README.md:44-85; variables src/main.cu:108-133). */
typedef struct mmc_params {
    int nx, ny, nz;                 /* cells */
    double dx, dy, dz;              /* cell size [m] */
    int pbc_x, pbc_y, pbc_z;        /* 0 open; >0 periodic with +-p images;
                                       <0 periodic without image sum */
    double ms;                      /* saturation magnetisation [A/m] */
    double aex;                     /* exchange stiffness [J/m] */
    double dind;                    /* interfacial DMI [J/m^2] */
    double anis_ux, anis_uy, anis_uz, ku1;  /* uniaxial axis, Ku1 [J/m^3] */
    double bext_x, bext_y, bext_z;  /* applied field [T] */
    int cal_demag;                  /* cal_demag_flag */
    int reserved;                   /* must be 0 */
    unsigned long long seed;        /* randomseed (Philox key) */
} mmc_params;

typedef struct mmc_ctx mmc_ctx;

/* precision modes */
enum { MMC_FP64 = 0,   /* plain FP64 everywhere (reference -DUSE_DOULBE_IN_GPU) */
       MMC_FP32 = 1 }; /* FP32 FFT/stencil/dE; FP64 tensor and reductions
                          (reference default build, src/constant.h:41-51) */

/* field / energy term bits */
enum { MMC_EXCH = 1, MMC_DMI = 2, MMC_ANIS = 4, MMC_ZEEMAN = 8,
       MMC_DEMAG = 16, MMC_ALL = 31 };

/* status codes */
enum { MMC_OK = 0, MMC_EINVAL = 1, MMC_EHIP = 2, MMC_EIO = 3,
       MMC_ESTATE = 4, MMC_EUNSUP = 5 };

/* ---- lifetime --------------------------------------------------------- */

/* Allocate one instance on the current HIP device.
 * Replaces the cudaMalloc / cufftPlan3d set-up of src/main.cu:396-409,
 * :593-631, :688-756 and the padding rule :254-263.
 * MMC_EINVAL for grids the kernels' 32-bit cell indexing cannot take:
 * nx >= 2^24, ny*nz >= 2^24 or more than (2^32-1)/3 cells. */
int mmc_create(const mmc_params *p, int precision, mmc_ctx **out);

/* Allocate `count` independent instances (uneven sizes allowed) that are
 * advanced together, one launch per kernel stage for all of them.  No
 * reference counterpart: the reference runs one instance per process. */
int mmc_batch_create(const mmc_params *p, int count, int precision,
                     mmc_ctx **out);

/* Free everything.  Replaces the cudaFree / cufftDestroy block of
 * src/main.cu:1200-1252. */
void mmc_destroy(mmc_ctx *ctx);

/* Text of the last error (never NULL). */
const char *mmc_last_error(const mmc_ctx *ctx);

/* out = {nx, ny, nz, Px, Py, Pz} of instance `inst`. */
int mmc_dims(const mmc_ctx *ctx, int inst, int out[6]);

/* Number of instances in the context. */
int mmc_count(const mmc_ctx *ctx);

/* ---- demag tensor ----------------------------------------------------- */

/* Build the FP64 Newell tensor with periodic-image sums on the GPU, then its
 * FFT (downcast to FP32 in MMC_FP32 mode), for every instance.  If
 * cache_dir is non-NULL and holds the reference's cache file for these
 * dimensions it is loaded instead; otherwise the tensor is written there.
 * Replaces src/main.cu:419-675 (ktensorKernel2_step0/1/2, the image loop,
 * the tensor cache, cufftExecD2Z x6 and copyFFT64to32Kernel). */
int mmc_build_tensor(mmc_ctx *ctx, const char *cache_dir);

/* Replace instance `inst`'s real-space tensor (6*Px*Py*Pz doubles) and
 * recompute its FFT.  Test hook; same role as USE_CPU_TO_CALCULATE_K_TENSOR
 * (src/main.cu:462-464), whose host tensor is copied to the device. */
int mmc_set_tensor(mmc_ctx *ctx, int inst, const double *K);

/* Copy instance `inst`'s real-space tensor to the host (6*Px*Py*Pz). */
int mmc_get_tensor(const mmc_ctx *ctx, int inst, double *K);

/* The six self-tensor ("mirror") scalars of src/main.cu:339-346 /
 * src/demag_mirror.h:18-24: the tensor's zero-offset entry [xx..zz]. */
int mmc_mirror_tensor(const mmc_ctx *ctx, int inst, double K0[6]);

/* ---- magnetisation ---------------------------------------------------- */

/* Upload / download unit magnetisation (N doubles per component).
 * Replaces the H2D copies of src/main.cu:719-730 and the snapshot D2H copy
 * of src/main.cu:842-853. */
int mmc_set_m(mmc_ctx *ctx, int inst, const double *mx, const double *my,
              const double *mz);
int mmc_get_m(const mmc_ctx *ctx, int inst, double *mx, double *my,
              double *mz);

/* Seeded uniform-on-sphere start state (Philox keyed by the instance's
 * seed); inst = -1 initialises every instance.  Replaces randInitMx
 * (src/mx_init.h:21-97). */
int mmc_init_random_m(mmc_ctx *ctx, int inst);

/* ---- observables ------------------------------------------------------ */

/* Effective field (A/m) of the selected terms at every real cell.
 * Replaces effectiveHKernel (src/effective_H.h:25-33) plus the demag
 * convolution of src/main.cu:868-900. */
int mmc_field(mmc_ctx *ctx, int inst, int terms, double *hx, double *hy,
              double *hz);

/* Total energy [J], per-term energies [J] (exch, dmi, anis, zeeman, demag)
 * and the mean magnetisation, with compensated FP64 reductions.  One call
 * evaluates every instance of the context; the sums are kept until the
 * state or the tensor changes, so reading all instances of a batch costs
 * one pass.  Replaces calEnergyKernel + thrust::reduce
 * (src/main.cu:902-956). */
int mmc_energy(mmc_ctx *ctx, int inst, double *E_total, double e_terms[5],
               double m_avg[3]);

/* mmc_energy without device memory of its own: the same energy terms and
 * mean magnetisation, bit for bit (same kernels, same summation order),
 * with the reduction's scratch (per-block partial sums and per-instance
 * sums) in a fixed 2 MiB device array of the library instead of memory the
 * context allocates; it allocates nothing and does not read or fill the
 * energy cache.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid allocation).  Limits: at most 1024 instances and at
 * most 15872 reduction blocks of 256 cells (about 4.06 M cells over all
 * instances); MMC_EINVAL beyond.  Every call on one GPU uses that GPU's
 * array, so calls that overlap in time on one GPU (contexts driven from
 * several host threads) race; a call returns after its stream has
 * finished. */
int mmc_energy_noalloc(mmc_ctx *ctx, int inst, double *E_total,
                       double e_terms[5], double m_avg[3]);

/* Energy density [J/m^3] of every real cell (N doubles), the selected terms
 * summed.  Replaces calEnergyKernel's per-cell output d_energy
 * (src/effective_H.h:35-43), which the reference dumps as E_density.ovf
 * under DEBUG_H_FIELD (src/main.cu:1050-1064). */
int mmc_energy_density(mmc_ctx *ctx, int inst, int terms, double *e);

/* Energy change [J] of every cell's Metropolis trial move for (stage,
 * sweep) — Philox counter (cell, sweep, stage, 0) — against the current
 * state and its demag field, the moves not applied (N doubles).  The dE
 * that mcKernel2 computes internally (src/mc_sampling.h:27-41); exposed to
 * state the FP32 error of dE against FP64. */
int mmc_trial_delta_e(mmc_ctx *ctx, int inst, uint32_t stage, uint32_t sweep,
                      double *dE);

/* ---- Monte Carlo ------------------------------------------------------ */

/* Set the temperature [K] of one instance (inst >= 0) or all (-1). */
int mmc_set_temperature(mmc_ctx *ctx, int inst, double T);

/* Run `count` Metropolis sweeps on every instance: per sweep, the demag
 * field of the pre-sweep state, then every colour class of the cell graph
 * (checkerboard; three classes when a periodic axis has odd length).
 * Sweep s uses Philox counter (cell, s, stage, 0).  Replaces the inner
 * loop body of src/main.cu:868-884 + :1101-1120 (R2C x3, calHexchangeKernel,
 * mcKernel2, M/M_out swap). */
int mmc_sweeps(mmc_ctx *ctx, uint32_t stage, uint32_t sweep0, uint32_t count);

/* Time every admissible sweep schedule — fused x-pass + Metropolis kernels
 * or separate launches or the tiled two-colour Metropolis kernel, each with
 * launch-sized, minimal and widest FFT tiles, z folding, and (for batches
 * with several lengths) concurrent per-length launches — over
 * sweeps_per_trial sweeps
 * (twice: once untimed) and keep the fastest.  Every schedule gives the
 * bit-identical Markov chain; the magnetisation is restored afterwards.  A
 * text table of the timings (and the choice) is written to report when it
 * is non-NULL.  No reference counterpart: the reference has one fixed
 * launch sequence (src/main.cu:868-884, :1101-1120). */
int mmc_autotune(mmc_ctx *ctx, int sweeps_per_trial, char *report,
                 size_t report_len);

/* Apply one of the schedules mmc_autotune times, by the name its report
 * prints (e.g. "fused, minimal tiles"), without timing anything: reuse a
 * tuned choice or pin one for profiling.  MMC_EUNSUP if the schedule needs
 * kernels these instances cannot take, MMC_EINVAL for an unknown name.  No
 * reference counterpart. */
int mmc_set_schedule(mmc_ctx *ctx, const char *name);

/* Single-instance convenience: set T, run one sweep. */
int mmc_sweep(mmc_ctx *ctx, double T, uint32_t stage, uint32_t sweep);

/* Wait for all queued work of the context. */
int mmc_sync(mmc_ctx *ctx);

/* ---- anneal driver ---------------------------------------------------- */

typedef struct mmc_schedule {
    double T_start, T_end, T_step;  /* Temperature_start/end/step */
    int use_exp;                    /* Temperature_use_exp */
    long sweeps_per_stage;          /* circle_per_stage */
    long energy_period;             /* energy_output_period (<=0: never) */
    long snapshot_period;           /* ms_output_period (<=0: never) */
} mmc_schedule;

/* Called before sweep `sweep` of stage `stage` at temperature T when the
 * period matches; return non-zero to stop the anneal. */
typedef int (*mmc_energy_cb)(void *user, double T, long stage, long sweep,
                             double E_total);
typedef int (*mmc_snapshot_cb)(void *user, double T, long stage, long sweep,
                               const double *mx, const double *my,
                               const double *mz);

/* Temperature-stage loop of src/main.cu:801-1139 for instance 0: while
 * T > T_end, sweeps_per_stage sweeps, then T *= step (use_exp) or
 * T -= step.  *stop is polled between sweeps (the SIGINT flag of
 * src/main.cu:82-88).  Returns the number of stages run in *stages_out. */
int mmc_anneal(mmc_ctx *ctx, const mmc_schedule *s, mmc_energy_cb ecb,
               mmc_snapshot_cb scb, void *user, volatile int *stop,
               long *stages_out);

/* ---- instrumentation -------------------------------------------------- */

/* Options: "graph" (1 = replay sweeps from a captured hipGraph),
 * "graph_sweeps" (sweeps per captured graph, default 16),
 * "timing" (1 = time every launch with hipEvents recorded by its dispatch
 * packet on the context's stream; sweeps then launch eagerly),
 * "default_stream" (1 = launch on the device's default stream instead of the
 * context's own, for profilers that observe only that one; disables "graph"),
 * "tyx" (x-pass rows per workgroup), "ncol" (y-pass columns per workgroup),
 * "metro_tiled" (1: both checkerboard colours in one tiled launch where
 * every instance has a checkerboard; 0, the default: one launch per colour),
 * "fuse" (1, the default where every instance has demag, a checkerboard and
 * an x length of 128, 192, 256, 384, 512, 1024 or 2048 after padding: the
 * inverse x pass runs with colour 0 and colour 1 with the next forward x
 * pass; 0: separate launches; "metro_tiled" = 1 takes precedence),
 * "fuse_rows" (1: one pair of real rows per fused workgroup; >1: the full
 * tile, 2048 FP32 / 1024 FP64 complex per component over the x length),
 * "rows" (1, with "fuse": every instance being a checkerboard film of one
 * or two layers with a fused-kernel x length and a tile that fits LDS, a
 * sweep is the y pass and one launch of whole-sweep row-pair tiles, each
 * inverting the x spectrum of its pair and both neighbouring pairs in every
 * layer, updating colour 0 on four rows (the halo rows' updates recomputed
 * as their own tiles compute them) and colour 1 on its pair, and
 * transforming the pair forward; the state alternates between two device
 * buffers, settled into the first at the end of every call; default 0),
 * "concurrent" (1: the per-length launches of a pass over a batch run on
 * side streams forked from and joined back into the context's stream;
 * default 0),
 * "zfold" (1: fused tiles of instances with a short padded z axis, Pz <= 8,
 * that still take z transforms span all layers and do the z transform in
 * registers, so sweeps skip the separate z passes; default 0),
 * "tensor_evals" (Newell evaluations per launch of the tensor build, which
 * splits the x images into slices of at least one image; partial sums stay
 * unnormalised between slices, so any slicing gives the bit-identical
 * tensor; default 2e9).
 * Films of 2-4 layers (Pz <= 8) with a y length of 64, 96, 128, 192, 256,
 * 384, 512, 768, 1024 or 2048 after padding take no z transforms at all:
 * one y pass per x frequency sums the z convolution directly ("yzconv").
 * Tile and fusion options only regroup independent work: results are
 * bit-identical.
 * Unknown keys return MMC_EINVAL. */
int mmc_set_option(mmc_ctx *ctx, const char *key, long value);

/* Launches, summed device time [ms] and summed algorithmic HBM bytes of
 * one kernel kind ("xfwd", "zfwd", "ymul", "zinv", "xinv", "metro",
 * "energy", "tensor", "xinv_metro", "metro_xfwd", "yzconv", "row_sweep") recorded
 * while "timing" was on; total_bytes / total_ms is the kind's achieved
 * bandwidth.  Resets the counters when
 * reset != 0.  Instrumentation only; no reference counterpart (the
 * reference prints a CPU-clock "stages/s", src/main.cu:795-827). */
int mmc_kernel_stats(mmc_ctx *ctx, const char *kind, long *launches,
                     double *total_ms, double *total_bytes, int reset);

/* Algorithmic HBM bytes of one pass of `kind` over all instances (each
 * array the pass touches read or written once; a pass may take several
 * launches, one per kernel variant). */
int mmc_kernel_bytes(const mmc_ctx *ctx, const char *kind, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
