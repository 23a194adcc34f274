/*
 * C host for the RBF-FD shallow-water RK4 path: input reader, RCM node
 * reordering, cfdl/sfdl layouts and padding, the serial single-patch set-up
 * and the profiling report.  These restate the reference's host modules
 * (cited per function) with the same argument meaning, array layouts and
 * error behaviour.  Values are held as doubles; the *_ex readers with
 * fp32 != 0 reproduce the reference's FP32 model type (RUN_32BIT,
 * swe_code/main/include/swe_config.h:75-93) by rounding what they read.
 *
 * All arrays are malloc'ed by the library and owned by the caller
 * (release with swe_gsmd_free / free()).  Functions return 0 on success and
 * a negative SWE_HOST_E* code on failure.
 */
#ifndef SWE_HOST_H
#define SWE_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_HOST_OK 0
#define SWE_HOST_EOPEN (-1)    /* input file cannot be opened (input.c:8)    */
#define SWE_HOST_EREAD (-2)    /* short read / malformed header              */
#define SWE_HOST_EARG (-3)     /* invalid argument                           */
#define SWE_HOST_ENOMEM (-4)
#define SWE_HOST_EGRAPH (-5)   /* stencil graph not connected (RCM)          */
#define SWE_HOST_EFORMAT (-6)  /* not a NetCDF classic file (CDF-1/2/5), or a
                                  text variable where numbers are read      */
#define SWE_HOST_ENOVAR (-7)   /* a NetCDF variable the reader needs is absent */
#define SWE_HOST_ERANGE (-8)   /* a value does not fit the int it is read into */

/* Global static model data, GSMD_struct (swe_config.h:127-165), FP64. */
typedef struct swe_gsmd {
    int Nnodes, padded_Nnodes;
    int Nnbr, padded_Nnbr;
    double a, gh0, dt, gamma;
    double *x, *y, *z, *f, *ghm;           /* [Nnodes]                      */
    double *p_u, *p_v, *p_w, *gradghm;     /* [Nnodes][3] (VECT_LIN_ID)     */
    int *idx;                              /* [Nnodes][Nnbr] (DM_LIN_ID)    */
    double *Dx, *Dy, *Dz, *L;              /* [Nnodes][Nnbr] (DM_LIN_ID)    */
} swe_gsmd;

/* Host layout switches: USE_CFDL / USE_SFDL / SIMD_LENGTH
 * (swe_config.h:16-54), chosen at run time here. */
typedef struct swe_layout {
    int cfdl;
    int sfdl;
    int simd;
} swe_layout;

/* Local patch of the serial run (PSMD_struct, swe_config.h:187-233):
 * one patch covering every padded node, compute_pid_s == 0. */
typedef struct swe_patch {
    int patch_size, compute_size, compute_pid_s, compute_pid_e;
    int halo_pid_s, halo_pid_e;
    int Nnbr, padded_Nnbr;
    double gh0, dt;
    const double *x, *y, *z, *f, *ghm, *p_u, *p_v, *p_w, *gradghm;
    const int *idx;
    const double *Dx, *Dy, *Dz, *L;
} swe_patch;

/* ---- index macros as functions (swe_config.h:21-54) ---------------------- */
long swe_dm_lin_id(const swe_layout *l, int node, int nbr, int nnbr);
long swe_vect_lin_id(const swe_layout *l, int node, int comp);
long swe_svm_lin_id(const swe_layout *l, int pid, int sv);
int swe_roundup(const swe_layout *l, int n);

/* ---- binary input (swe_code/io/input.c:189-414) ------------------------- */
/* get_GSMD_bin: reads constants, geometry, projections and weights; scales
 * Dx, Dy, Dz by 1/a and L by gamma (input.c:314-319) and converts idx to
 * 0-based (input.c:326-328). */
int swe_get_gsmd_bin(const char *path, swe_gsmd *g);
/* The same for the FP32 model type (fp32 != 0, the reference's RUN_32BIT
 * build): every value read is rounded to float, the weights once more after
 * the double-precision scaling (input.c:211-319); values stay stored as
 * doubles (each exactly a float). */
int swe_get_gsmd_bin_ex(const char *path, int fp32, swe_gsmd *g);
/* get_ICs_bin / get_FCs_bin: H_init and H_100, [Nnodes][4] (u,v,w,h). */
int swe_get_ics_bin(const char *path, double **H_init);
int swe_get_fcs_bin(const char *path, double **H_100);
/* which 0 = H_init, 1 = H_100; fp32 rounds to float as above. */
int swe_get_state_bin_ex(const char *path, int which, int fp32, double **H);
void swe_gsmd_free(swe_gsmd *g);

/* ---- NetCDF input (swe_code/io/input.c:14-186, nc2bin.c:10-84) ---------- */
/* The reference's USE_NCIO readers on NetCDF classic files (CDF-1, CDF-2 64-bit
 * offset, CDF-5 64-bit data; NetCDF-4/HDF5 files are SWE_HOST_EFORMAT),
 * without libnetcdf.  Variables: Nnodes, Nnbr (int), a, g, gh0, timestep,
 * gamma, x, y, z, f, ghm [Nnodes], p_u, p_v, p_w, gradghm [Nnodes][3],
 * DPx, DPy, DPz, L, idx [Nnodes][Nnbr] (unit sphere, 1-based), H and H_100
 * [Nnodes][4], of any numeric type (converted as nc_get_var_double /
 * nc_get_var_int convert).  get_GSMD_nc: the same scaling and 0-based ids as
 * the binary reader; a variable of another size than Nnodes/Nnbr imply is
 * SWE_HOST_EREAD. */
int swe_get_gsmd_nc(const char *path, swe_gsmd *g);
int swe_get_gsmd_nc_ex(const char *path, int fp32, swe_gsmd *g);
/* get_ICs_nc / get_FCs_nc: which 0 = H, 1 = H_100. */
int swe_get_state_nc_ex(const char *path, int which, int fp32, double **H);
/* nc2bin: writes the binary input file of a NetCDF one (the weights read,
 * scaled and unscaled again, the ids made 1-based, as nc2bin.c does). */
int swe_nc2bin(const char *nc_path, const char *bin_path);

/* ---- RCM (swe_code/rcm/rcm.c:9-173, reorder_nodes.c:8-87) ---------------- */
/* rcm_mapping: new position -> old node id, malloc'ed [Nnodes]. */
int swe_rcm_mapping(const swe_gsmd *g, int **mapping);
/* reorder_GSMD_struct: permute every per-node array and remap idx. */
int swe_reorder_gsmd(swe_gsmd *g, const int *mapping);
/* reorder_2D_fp_arr: returns a new array, frees the old one. */
double *swe_reorder_2d(double *var, const int *mapping, int dim1, int dim2);
/* print_max_bandwidth (rcm.c:68-96) without printing. */
void swe_stencil_bandwidth(const swe_gsmd *g, int *max_bw, int *avg_bw);

/* ---- layout (swe_code/layout/layout.c, matrix_transformations.c) --------- */
int swe_pad_gsmd(swe_gsmd *g, const swe_layout *l);          /* layout.c:10  */
int swe_convert_sfdl(swe_gsmd *g, const swe_layout *l);      /* layout.c:48  */
double *swe_pad_2d(double *var, int dim1, int pdim1, int dim2, int pdim2);
double *swe_transpose_2d(double *a, int dim1, int dim2);
double *swe_tile_2d(double *a, int dim1, int dim2, int dim2_tiled);
/* main.c:149-154: state [Np][4] -> host state layout (non-CFDL tiles). */
int swe_state_to_layout(double **H, int padded_Nnodes, const swe_layout *l);

/* ---- serial patch set-up (swe_code/mpi/init_patches.c, mpi_size == 1) --- */
int swe_init_patch(const swe_gsmd *g, const swe_layout *l, swe_patch *p);

/* ---- one call for the whole init pipeline of main.c:94-190 -------------- */
typedef struct swe_problem {
    swe_gsmd g;
    swe_layout layout;
    swe_patch patch;
    double *H_init;     /* [patch_size*4] host state layout */
    double *H_100;      /* [patch_size*4] host state layout */
    int *mapping;       /* RCM mapping or NULL               */
} swe_problem;
int swe_problem_load(const char *path, int use_rcm, const swe_layout *l,
                     swe_problem *p);
/* With fp32 != 0: the FP32 model type's reading (swe_get_gsmd_bin_ex). */
int swe_problem_load_ex(const char *path, int use_rcm, const swe_layout *l, int fp32,
                        swe_problem *p);
/* With netcdf != 0 the input is a NetCDF file (main.c:99-106,
 * SWE_USE_NETCDF == 1). */
int swe_problem_load_fmt(const char *path, int use_rcm, const swe_layout *l, int fp32,
                         int netcdf, swe_problem *p);
void swe_problem_free(swe_problem *p);

/* ---- verify_output (main.c:439-477) --------------------------------------- */
int swe_verify_output(const swe_patch *p, const swe_layout *l,
                      const double *H1, const double *H2, double tol);

/* ---- profiling (swe_code/main/profiling.c) -------------------------------- */
typedef struct swe_timers {
    int nattempts, attempt;
    double t_init, t_dev;
    double *t_main, *t_eval_rhs, *t_eval_K, *t_update_D, *t_update_H;
} swe_timers;
double swe_get_time(void);
int swe_timers_init(swe_timers *t, int nattempts);
void swe_timers_free(swe_timers *t);
/* process_profiling_results (profiling.c:54-172): per-phase
 * average/min/max/"stddev" in seconds per timestep; Eval_Rhs excludes the
 * first two attempts (profiling.c:119-125; all zeros with nattempts <= 2).
 * out[phase*4 + {0,1,2,3}] = avg, min, max, stddev for phases
 * main, eval_rhs, eval_K, update_D, update_H.  Prints the report if
 * print != 0. */
void swe_process_profiling(const swe_timers *t, int nsteps, int print,
                           double out[20]);

#ifdef __cplusplus
}
#endif
#endif
