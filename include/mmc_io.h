/*
 * mmc_io.h -- configuration and vector-field file formats of the annealer
 * (host-only; part of libmmc.so, usable without a GPU).
 *
 * The reference declares these routines but ships empty bodies; the
 * formats they name are restated here:
 *   - ini input: the 32 keys read by parse_input_ini
 *     (src/parse_input_ini.h:15-47, README.md:44-85), Boost property_tree
 *     INI syntax (key = value; ';' and '#' comment lines; bools
 *     true/false/1/0);
 *   - OOMMF OVF 2.0 vector fields (text, binary 4, binary 8), as
 *     ovf_io::parseOvfFile / writetoOvfFile (src/ovf_io.h:88-99) intended,
 *     following MuMax3's oommf reader/writer cited at src/ovf_io.h:28;
 *     OVF 1.0 binary (big-endian) is read too.
 * Data order is x fastest, then y, then z (src/ovf_io.h:156-164).
 */
#ifndef MMC_IO_H
#define MMC_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh description, src/ovf_io.h:35-41. */
typedef struct mmc_ovf_info {
    int xnodes, ynodes, znodes;
    double xstepsize, ystepsize, zstepsize;
} mmc_ovf_info;

/* Read an OVF file.  *info receives the file's mesh; at most `capacity`
 * cells are written to mx/my/mz (NULL arrays: header only).  Returns 0 or
 * an MMC_E* code; msg (if non-NULL) receives a description of a failure. */
int mmc_ovf_read(const char *path, mmc_ovf_info *info, double *mx, double *my,
                 double *mz, long capacity, char *msg, int msglen);

/* Write an OVF 2.0 file; format is 0 (text), 4 (binary 4) or 8 (binary 8). */
int mmc_ovf_write(const char *path, const mmc_ovf_info *info,
                  const double *mx, const double *my, const double *mz,
                  int format, const char *title);

/* Mesh equality with the reference's 1e-5 relative step-size tolerance
 * (src/ovf_io.h:26-27, :42-53): 1 if equal. */
int mmc_ovf_mesh_equal(const mmc_ovf_info *a, const mmc_ovf_info *b);

/* All ini keys (names as in the file), src/main.cu:108-133. */
typedef struct mmc_config {
    int nx, ny, nz;
    double dx, dy, dz;
    double Bextx, Bexty, Bextz;
    int pbc_x, pbc_y, pbc_z;
    double ms, Aex, Dind;
    double anisUx, anisUy, anisUz, Ku1;
    int use_random_init;
    long long randomseed;
    char ovf_filename[4096];
    double Temperature_start, Temperature_end, Temperature_step;
    int Temperature_use_exp;
    long circle_per_stage;
    long terminal_output_period;
    long energy_output_period;
    long ms_output_period;
    int rand_block_size;
    int cal_demag_flag;
    /* keys this build adds (optional): precision = fp32 | fp64 (default fp32,
     * the reference's default build), ovf_format = 0 | 4 | 8 (default 8) */
    int precision;
    int ovf_format;
} mmc_config;

/* Parse an ini file.  Missing required keys, unknown values and malformed
 * lines are errors (message in msg).  Returns 0 or an MMC_E* code. */
int mmc_ini_parse(const char *path, mmc_config *cfg, char *msg, int msglen);

#ifdef __cplusplus
}
#endif
#endif
