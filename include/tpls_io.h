/*
 * tpls_io.h -- C ABI of the reference's input formats, on either side of the
 * device-resident timestep (host code; no GPU is needed except by the two
 * tpls_case_* functions at the end).
 *
 *  1. 3-D data files (initial_u.dat ... initial_viscosity.dat and every other
 *     .dat array file): a description line, a dimensions line, the six index
 *     bounds, then one value per record in Fortran (i fastest) order
 *     (io.F90:34-169).  Reading accepts whatever Fortran list-directed output
 *     produces (any compiler's layout, E/D/Q or letterless exponents, blank
 *     records, trailing items); writing produces the shortest decimal that
 *     reads back to the same double, so a file round-trips bit-exactly
 *     through this library and through the reference's own load_data.
 *     Files are parsed and formatted by several host threads at once.
 *
 *  2. NAME VALUE options files (tpls_config.opt, initial_config.dat) and the
 *     typed option accessors of options_utils.F90:56-374, with the
 *     reference's error codes, truncation lengths and value formats.
 *
 *  3. The run configuration main_ns_hybrid.F90:242-311 reads through
 *     configure.F90, and loading a case from a configuration directory
 *     (main_ns_hybrid.F90:380-393).
 *
 * Every function returns TPLS_OK (0) or a negative TPLS_ERR_* code, as the
 * functions of tpls_core.h do; tpls_last_error() describes the failure.  The
 * reference's own status values (which it returns through an ierr argument)
 * are returned through an `int *ierr` argument here too.
 */
#ifndef TPLS_IO_H
#define TPLS_IO_H

#include <stddef.h>

#include "tpls_core.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TPLS_ERR_IO = -6,      /* a file could not be opened, read or written */
    TPLS_ERR_FORMAT = -7,  /* a file or value does not parse (where the
                              reference's Fortran runtime would stop the
                              program with a read error) */
    TPLS_ERR_CONFIG = -8   /* a configuration is missing or invalid; *ierr
                              holds the reference's code for it */
};

/* The reference's status codes (options_utils.F90:41-51, configure.F90:22-26). */
enum {
    TPLS_OPTION_NOT_FOUND_ERR = 100,
    TPLS_OPTION_FORMAT_ERR = 101,
    TPLS_OPTION_INVALID_ERR = 102,
    TPLS_OPTION_FILE_NOT_FOUND_ERR = 103,
    TPLS_CONFIG_FILE_NOT_FOUND_ERR = 200,
    TPLS_CONFIG_DIR_NOT_FOUND_ERR = 201
};

/* ---- 3-D data files ---- */

/* save_data (io.F90:47-73): writes data(lx:ly,mx:my,nx:ny) (column-major,
 * i fastest) under the description. */
int tpls_save_data(const char *filename, const char *description, const double *data, int lx, int ly,
                   int mx, int my, int nx, int ny);
/* The six bounds recorded in a data file (lx, ly, mx, my, nx, ny). */
int tpls_data_bounds(const char *filename, int bounds[6]);
/* load_data (io.F90:83-116): reads the bounds and allocates *data (release it
 * with tpls_free).  *ierr = 1 when the file holds fewer values than its
 * bounds ask for (the missing values are zero), else 0. */
int tpls_load_data(const char *filename, double **data, int bounds[6], int *ierr);
/* load_data_into_array (io.F90:127-169): *ierr = 1, 2 or 3 when the file's
 * l, m or n bounds differ from the given ones (data untouched), else 0.  A
 * file that ends early returns TPLS_ERR_FORMAT. */
int tpls_load_data_into_array(const char *filename, double *data, int lx, int ly, int mx, int my, int nx,
                              int ny, int *ierr);
void tpls_free(void *p);

/* ---- NAME VALUE options (options_utils.F90) ----
 *
 * An ordered name -> value table.  Names compare as Fortran character
 * values do (trailing blanks do not count); values keep their leading blanks
 * (put_integer_option stores the I10 field, for instance). */
typedef struct tpls_options tpls_options;

int tpls_options_create(tpls_options **out);
int tpls_options_destroy(tpls_options *o);
/* num_options, and the i-th (0-based) entry; the strings stay valid until
 * the table changes. */
int tpls_options_count(const tpls_options *o);
int tpls_options_entry(const tpls_options *o, int i, const char **name, const char **value);
/* has_option (options_utils.F90:56-70): 1 or 0. */
int tpls_has_option(const tpls_options *o, const char *name);
/* get_option / get_string_option (options_utils.F90:74-91, 140-148): copies
 * the value into value[0..len-1], truncated and NUL-terminated;
 * *ierr = TPLS_OPTION_NOT_FOUND_ERR when there is no such option. */
int tpls_get_option(const tpls_options *o, const char *name, char *value, size_t len, int *ierr);
/* get_integer_option, get_double_option, get_logical_option
 * (options_utils.F90:95-136): the first 60 characters of the value read as
 * one list-directed item.  A value that is not of the type returns
 * TPLS_ERR_FORMAT with *ierr = TPLS_OPTION_FORMAT_ERR. */
int tpls_get_integer_option(const tpls_options *o, const char *name, int *value, int *ierr);
int tpls_get_double_option(const tpls_options *o, const char *name, double *value, int *ierr);
int tpls_get_logical_option(const tpls_options *o, const char *name, int *value, int *ierr);
/* put_option (options_utils.F90:157-174): replaces the value of an existing
 * name, else appends.  The typed variants store the reference's edit
 * descriptors: I10, F20.10 and L1 (options_utils.F90:179-227). */
int tpls_put_option(tpls_options *o, const char *name, const char *value);
int tpls_put_integer_option(tpls_options *o, const char *name, int value);
int tpls_put_double_option(tpls_options *o, const char *name, double value);
int tpls_put_logical_option(tpls_options *o, const char *name, int value);
/* load_options (options_utils.F90:254-332): '#' lines and blank lines are
 * skipped; every other line (its first 80 characters, leading blanks
 * removed) splits at its first blank into a name (first 20 characters) and
 * a value (first 60); later lines replace earlier ones of the same name.
 * *ierr = TPLS_OPTION_FILE_NOT_FOUND_ERR (no such file, *out = NULL) or
 * TPLS_OPTION_FORMAT_ERR (a line with a single word; *out holds the options
 * read before it), else 0. */
int tpls_load_options(const char *filename, tpls_options **out, int *ierr);
/* save_options (options_utils.F90:338-355): hdr lines as "# line", then the
 * options; append_options (options_utils.F90:360-374) adds the options as
 * "name value" lines to an existing file. */
int tpls_save_options(const char *filename, const tpls_options *o, const char *const *hdr, int num_hdr_lines);
int tpls_append_options(const char *filename, const tpls_options *o);

/* ---- run configuration ---- */

/* The values of tpls_config.opt and initial_config.dat that main_ns_hybrid.F90
 * reads (main_ns_hybrid.F90:242-311) beyond those tpls_params holds. */
typedef struct {
    int maxl, maxm, maxn;
    double dpdl, omega_max, maxu;
    int idm_dim;                 /* 0: idm lsm, 1: idm dim */
    double Pe, epn;              /* idm dim only */
    int max_iteration_dim;       /* idm dim only */
    int auto_decomp, num_procs_x, num_procs_y;
    int n_timesteps;
    int phi_dat_frequency, uvw_dat_frequency, backup_frequency;
    int backup_text_format, backup_hdf5_format;
} tpls_run_config;

/* Reads config_dir/initial_config.dat and the options file as
 * main_ns_hybrid.F90:242-311 does (configure.F90:39-764: the same required
 * keys in the same order, the same checks, codes and messages) and sets the
 * file-backed fields of *prm (dx, dy, dz, dt, Re, mu_plus, mu_minus, height,
 * smooth_width, scap, mom_u, mom_v, mom_w, levelset_its, and idm_dim, epn,
 * Pe, dim_its from idm, epn, pe, max_iteration_dim; the others are left as
 * the caller set them) and *run.  The process grid is checked for itself
 * (num_procs_x, num_procs_y >= 1, dividing maxl-1 and maxm-1) but not against
 * a process count: one GPU holds the whole domain.  A failed check returns
 * TPLS_ERR_CONFIG with the reference's code in *ierr and its message in
 * tpls_last_error(). */
int tpls_read_config(const char *config_dir, const char *options_file, tpls_params *prm, tpls_run_config *run,
                     int *ierr);

/* Loads the initial state of a case from config_dir/initial_{u,v,w,pressure,
 * phi,viscosity}.dat, as main_ns_hybrid.F90:380-393 (configure.F90:94-122)
 * loads it, then tpls_case_load()s it.  The viscosity file is read and checked
 * as the reference does but not used: the timestep computes the viscosity
 * from phi (main_ns_hybrid.F90:646-651) before any use.  A missing file or a
 * bounds mismatch returns TPLS_ERR_CONFIG with *ierr = 200 or 1/2/3. */
int tpls_case_load_config(tpls_case *c, const char *config_dir, int *ierr);

/* save_data of a field of a case (its Fortran bounds, see
 * tpls_case_get_field). */
int tpls_case_save_field(tpls_case *c, int field, const char *filename, const char *description);

/* ---- channel outputs and the text backup (io.F90:171-791) ----
 *
 * The reference's outputs for visual comparison, with its argument lists
 * (arrays with their Fortran bounds, l = maxl, m = maxm, n = maxn), its
 * record order, coordinates (cell centres, x = dx/2 + dx*i) and face
 * averages.  Values are written as the shortest decimals that read back to
 * the same doubles; the header lines follow list-directed output.  The
 * caller passes the arrays main_ns_hybrid.F90 passes (its gathered globals:
 * the run driver of tpls_amd/run.py builds them from a case). */
int tpls_save_backup_info(const char *filename, int lx, int ly, int mx, int my, int nx, int ny, int dx, int dy,
                          int dz);                                   /* io.F90:180-200 (integer spacings, as there) */
int tpls_channel_output_phi(const double *phi, int l, int m, int n, double dx, double dy, double dz,
                            const char *filename);                   /* io.F90:217-246 */
/* io.F90:503-534: *ierr = 1 when the file holds fewer records than the grid. */
int tpls_channel_input_phi(const char *filename, double *phi, int l, int m, int n, int *ierr);
int tpls_channel_output_uvw(const double *u3, const double *v3, const double *w3, const double *p, int l, int m,
                            int n, double dx, double dy, double dz, const char *filename);   /* io.F90:551-587 */
int tpls_channel_output2d(const double *u3, const double *v3, const double *w3, const double *p, const double *phi,
                          const double *visc, int l, int m, int n, double dx, double dy, double dz,
                          const char *filename);                     /* io.F90:604-643 */
/* backup_channel / dataload_channel (io.F90:647-791): u2, v2, w2, phi with
 * their conv0/conv1 histories and pres, pres_old, three (two) values per
 * record. */
int tpls_backup_channel(const double *u2, const double *v2, const double *w2, const double *conv0_u,
                        const double *conv1_u, const double *conv0_v, const double *conv1_v, const double *conv0_w,
                        const double *conv1_w, const double *pres_old, const double *pres, const double *phi,
                        const double *conv0_phi, const double *conv1_phi, int l, int m, int n, const char *filename);
int tpls_dataload_channel(double *u2, double *v2, double *w2, double *conv0_u, double *conv1_u, double *conv0_v,
                          double *conv1_v, double *conv0_w, double *conv1_w, double *pres_old, double *pres,
                          double *phi, double *conv0_phi, double *conv1_phi, int l, int m, int n,
                          const char *filename);

/* ---- checkpoint / resume ---- */

/* A full-state snapshot of a case between steps: every device array that
 * carries state into the next step, in its storage precision (u2, v2, w2;
 * conv0/conv1 and csf0/csf1 of each component; conv0/conv1 of phi; phi2;
 * pres; the FP64 pres_old the pressure solve starts from; u_inlet), with a
 * checksum per array and the step count.  A case restored from it steps
 * bit-identically to the case that wrote it.  The reference's backup_channel
 * (io.F90:647-716) leaves out the CSF histories, u_inlet and the step count,
 * so a run cannot resume exactly from it.  The file is written under a
 * temporary name and renamed into place.  Binary, host byte order. */
int tpls_case_checkpoint(tpls_case *c, const char *filename);
/* Restores a snapshot into a case of the same grid and precision (other
 * parameters, dt for instance, may differ).  A checksum mismatch returns
 * TPLS_ERR_FORMAT and leaves the case state undefined. */
int tpls_case_restore(tpls_case *c, const char *filename);
/* Timesteps the case has advanced (restored with a snapshot). */
int tpls_case_steps(tpls_case *c, long *steps);

#ifdef __cplusplus
}
#endif

#endif
