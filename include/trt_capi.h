/*
 * trt_capi.h -- C ABI of the MI355X-native full-order TRT iteration.
 *
 * Plain C, no C++ or torch types: every entry point takes a context handle,
 * plain pointers and sizes, and returns 0 on success or a negative status;
 * trt_last_error() then gives the message.  State lives on the GPU for the
 * whole run; host arrays cross the boundary only through trt_push/trt_pull
 * (initialisation, restart, outputs), with the reference's own array names
 * and Fortran (column-major) shapes, e.g. I_crn(2N_x,2N_y,N_m,N_g).
 *
 * Each routine replaces the reference Fortran subroutine cited beside it
 * (paths relative to the reference repository).  The Fortran side binds them
 * with ISO_C_BINDING (see the fortran/ directory of the package); the
 * reference's Open_Threads arguments have no counterpart (one context = one
 * HIP stream).
 */
#ifndef TRT_CAPI_H
#define TRT_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct trt_ctx trt_ctx;

/* Problem description: the values MainSrc/INPUTS.f08 parses from input.inp
 * (INPUT :10-163, INPUT_SOLVER_OPTS :307-609, INPUT_PARAMETERS :870-1141,
 * INPUT_ENERGY :1517-1573, INPUT_QUAD :1579-1671, INPUT_TEST_TYPE :1676-1729),
 * in the reference's code units (eV, sh, cm, erg/comp_unit). */
typedef struct {
    int nx, ny, nm, ng, nmat;
    const double *dx, *dy;          /* Delx(N_x), Dely(N_y)                         */
    const double *ox, *oy, *w;      /* Omega_x, Omega_y, quad_weight(N_m), quadrant blocks q1..q4 */
    const double *nu;               /* nu_g(N_g+1) group edges                      */
    const int *mat;                 /* Mat(N_x,N_y), 0-based material ids           */
    const double *kmult;            /* Kappa_Mult(nmat)                             */
    double c, dt, theta, comp_unit, cv, kap0, tini;
    int bc_type[4];                 /* left, bottom, right, top: 0 incoming Planckian, 1 reflective */
    double bcT[4];
    int opacity_model;              /* 0 = kapB_calc group Planck mean, 1 = group-midpoint formula */
    int maxit_rte, maxit_mloqd, maxit_gloqd, conv_type;
    double conv_ho, conv_lo, conv_gr1, conv_gr2, chi, line_src;
    int use_line_search, use_safety_search, kapE_dT_flag;
    double E_bound_low, T_bound_low;
    int threads;                    /* accepted and ignored (reference Open_Threads) */
    /* extensions of this implementation */
    int precision;                  /* 0 = FP64 sweep; 1 = FP32 sweep (storage and arithmetic) with FP64
                                       moment accumulation; 2 = FP32 intensity storage with FP64
                                       arithmetic in the corner solve (tile engine only) */
    int tile_x, tile_y;             /* sweep tile shape; 0 = default                 */
    double lin_rtol;                /* relative residual target of the QD Krylov solves; 0 = default 1e-13 */
    int lin_maxit;                  /* iteration cap of the QD Krylov solves (qd_solver 1); 0 = default
                                       2000.  A system BiCGSTAB does not solve goes to the direct solve */
    int sweep_mode;                 /* 0 = default (2, or 4 on grids large enough that column
                                       segments keep the device busy); 1 = one launch per tile diagonal with a
                                       per-cell moment reduction; 2 = one persistent launch per
                                       sweep; 3 = one launch per tile diagonal with per-row
                                       moment reduction; 4 = one persistent launch over column
                                       segments of tiles                                  */
    int qd_solver;                  /* reduced QD systems: 0 = nested-dissection direct solve whose
                                       factors are reused within a time step (iterative refinement
                                       with the last factors of the same systems to the direct
                                       solve's backward error 1e-14; a new factorization when they no
                                       longer contract); 1 = batched BiCGSTAB with the direct solve as
                                       fallback; 2 = block LU over grid rows (the slow direct solve
                                       that backs up the others); 3 = nested dissection factorizing
                                       every solve */
    int grey_emission;              /* grey emission of EGP_FV_NEWT: 0 = the reference's analytic
                                       Q_bar = 15 kap0 c aR T / pi^4 (GLOQD_SOLVES.f08:439-440,766);
                                       1 = sum_g 4 pi kappa_g(T) B_g(T) with the cell's opacity
                                       multiplier, consistent with the group sources (a labelled
                                       deviation from the reference)                   */
    int no_grey;                    /* 0 = use_grey 1 (the grey Newton solve); 1 = use_grey 0: MEB_SOLVE
                                       in the MLOQD loop (MainSrc/ALGORITHMS.f08:767-770) */
} trt_problem;

/* Context life cycle.  trt_ctx_create allocates the device-resident state and
 * performs MISC_INIT, RT_INIT, MGQD_INIT, GQD_INIT and TEMP_INIT
 * (MainSrc/ALGORITHMS.f08:203-218, MainSrc/INITIALIZERS.f08:15-427). */
int trt_ctx_create(trt_ctx **ctx, const trt_problem *p);
int trt_ctx_destroy(trt_ctx *ctx);
const char *trt_last_error(const trt_ctx *ctx);
/* device bytes a context for problem p allocates at creation (without creating
 * it); full time steps add the QD systems on first use */
size_t trt_ctx_bytes(const trt_problem *p);
/* device bytes a context holds now */
size_t trt_ctx_device_bytes(const trt_ctx *ctx);

/* host <-> device copies by reference array name, Fortran shapes.  n is the
 * number of doubles the caller's buffer holds (must equal the array size).
 * trt_array_size returns that size, or 0 for an unknown name.  I_avg, I_edgV,
 * I_edgH, Ic_edgV and Ic_edgH are pull-only views derived from I_crn and the
 * boundary intensities; pushing KapE or RT_Src also refreshes the sweep's
 * per-(cell, group) coefficients. */
size_t trt_array_size(const trt_ctx *ctx, const char *name);
int trt_push(trt_ctx *ctx, const char *name, const double *host, size_t n);
int trt_pull(trt_ctx *ctx, const char *name, double *host, size_t n);

/* MATERIAL_UPDATE, MainSrc/UPDATES.f08:110-156 (kapB_calc/Bg_planck_calc/sigint,
 * MainSrc/TEMP_FUNCTIONS.f08:11-141).  RT_Src is kept per (cell, group) since the
 * reference writes the same value to all corners and angles (:141-150). */
int trt_material_update(trt_ctx *ctx);
/* TRANSPORT_SCB, MainSrc/TRANSPORT_SOLVES.f08:14-327, with the angular moments
 * of COLLAPSE_INTENSITIES (:381-511) accumulated in the same pass. */
int trt_transport_scb(trt_ctx *ctx);
/* RT_BC_UPDATE, MainSrc/UPDATES.f08:12-105: on each reflective side (BC_Type 1) the
 * incoming directions take the outgoing intensities of their mirror directions from the
 * last sweep (the next sweep's inflow, the incoming faces of the I_edgV/I_edgH views);
 * nothing on a context without reflective sides.  trt_step calls it after every sweep
 * (ALGORITHMS.f08:556-565); the Tgen_QDf steps do not, as the reference's Tgen_QDf
 * does not (its reflective sides keep their initial Planckian inflow). */
int trt_rt_bc_update(trt_ctx *ctx);
/* COLLAPSE_INTENSITIES, MainSrc/TRANSPORT_SOLVES.f08:381-511 (reduces the partial
 * moments written by trt_transport_scb). */
int trt_collapse_intensities(trt_ctx *ctx);
/* fg_Calc MainSrc/MLOQD_SOLVES.f08:619-678; Cg_Calc :367-488; MGQD_In_Calc :493-614 */
int trt_fg_calc(trt_ctx *ctx);
int trt_cg_calc(trt_ctx *ctx);
int trt_mgqd_in_calc(trt_ctx *ctx);
/* OLD_MGQD_COEFS MainSrc/MLOQD_SOLVES.f08:12-73; OLD_GREY_COEFS MainSrc/GLOQD_SOLVES.f08:12-42 */
int trt_old_mgqd_coefs(trt_ctx *ctx);
int trt_old_grey_coefs(trt_ctx *ctx);
/* MLOQD_FV, MainSrc/MLOQD_SOLVES.f08:78-362, solving the N_g reduced systems of
 * QD_FV (MainSrc/QD_SOLVERS.f08:40-538) as one batch.  kits (may be NULL)
 * receives N_g Krylov iteration counts, -1 for a group solved by the block-LU
 * direct fallback (the reference's dense fallback, QD_SOLVERS.f08:492-500). */
int trt_mloqd_fv(trt_ctx *ctx, int *kits);
/* COLLAPSE_MG_EF MainSrc/MLOQD_SOLVES.f08:687-730; Cbar_Calc MainSrc/GLOQD_SOLVES.f08:177-231;
 * GREY_COEFS :47-146; GQD_In_Calc :263-290 */
int trt_collapse_mg_ef(trt_ctx *ctx);
int trt_cbar_calc(trt_ctx *ctx);
int trt_grey_coefs(trt_ctx *ctx);
int trt_gqd_in_calc(trt_ctx *ctx);
/* MEB_SOLVE, MainSrc/UPDATES.f08:189-231: Temp from the group energies named eg ("Eg_avg" in
 * the MLOQD loop, "HO_Eg_avg" for tr_no_qd, ALGORITHMS.f08:666,769), KapE and Temp_Old. */
int trt_meb_solve(trt_ctx *ctx, const char *eg);
/* EGP_FV_NEWT, MainSrc/GLOQD_SOLVES.f08:305-860; its receives the Newton count.  With
 * use_safety_search the positivity search of GL:643-752 runs once per Newton iteration
 * (the reference's Safe_Search loop has no exit); a bound that no step length reaches
 * (a value already below it) is an error. */
int trt_egp_fv_newt(trt_ctx *ctx, int mgqd_it, int *its);
/* CONVERGENCE, MainSrc/CONVERGENCE_CHECKS.f08:7-44, on device arrays x, x_old, x_old2
 * named as in trt_pull. */
int trt_convergence(trt_ctx *ctx, const char *x, const char *x_old, const char *x_old2,
                    int conv_type, double eps, int its, double *norm, double *rho, int *converged);
/* One time step of TRT_MLQD_ALGORITHM (run_type mlqd, use_grey = 1),
 * MainSrc/ALGORITHMS.f08:389-897.  its_out (may be NULL) receives
 * {RT iterations, MLOQD iterations, Newton iterations, sweeps}. */
int trt_step(trt_ctx *ctx, int its_out[4]);
/* The time step of the ROM runs (run_type mg_pod / mg_dmd with fg databases, HO_Form 1
 * and 3, MainSrc/ALGORITHMS.f08:573-631).  Before the call the caller pushes the new time
 * level's QD factors and C_g -- what POD_RECONSTRUCT_fg / DMD_RECONSTRUCT_fg and their BCg
 * counterparts write (MainSrc/POD_ROUTINES.f08:28, MainSrc/DMD_ROUTINES.f08:131) -- to the
 * arrays fg_avg_xx_next, fg_avg_yy_next, fg_edgV_xx_next, fg_edgV_xy_next, fg_edgH_yy_next,
 * fg_edgH_xy_next, Cg_L_next, Cg_B_next, Cg_R_next, Cg_T_next (shapes of the arrays without
 * _next).  The step keeps the current factors as the old level's, as the reference does at
 * its start, copies the _next arrays in where the reconstruction writes, and runs the
 * MLOQD/GLOQD iteration with one RT iteration and no sweep (maxit_RTE = 1,
 * MainSrc/INPUTS.f08:94-104).  its_out as trt_step; its_out[3] (sweeps) is 0.
 * Direc_Diff = 1 runs zero fg_edgV_xy and fg_edgH_xy after the reconstruction
 * (ALGORITHMS.f08:593-596): such a caller pushes zeros to fg_edgV_xy_next and
 * fg_edgH_xy_next after reconstructing (the step copies them in unchanged).  A push of
 * BCg_next (the reference's C_g boundary vector, POD_ROUTINES.f08:279-352) is split into
 * Cg_L/B/R/T_next at once, as a reconstruction into BCg_next is. */
int trt_step_given_qdf(trt_ctx *ctx, int its_out[4]);

/* POD reconstruction of one field of the ROM runs: POD_RECONSTRUCT,
 * MainSrc/POD_ROUTINES.f08:372-432, with LINEAR_INTERPOLATION between two data times,
 * MainSrc/GRID_FUNCTIONS.f08:819-836, as POD_RECONSTRUCT_fg calls them
 * (POD_ROUTINES.f08:28-130) for data on the simulation grid (the data-grid map FMAP is not
 * built).  trt_pod_create uploads a basis in the reference's storage -- C (centering
 * vector, the size of array `name`), S (singular values), U (left vectors: per group
 * rank(g) columns of size/N_g values, or rank(1) columns of the whole size for gsum 1),
 * V (right vectors: per group nt x rank(g), t-major), rank (N_g values, or 1 for gsum 1)
 * -- for reconstruction into the context array `name` (normally a *_next array of
 * trt_step_given_qdf, or a dat_* array of a trt_romgrid for data on another grid).
 * trt_pod_reconstruct writes the field at data time t1 (1-based),
 * or, when t2 != t1, interpolated linearly to time x between data times x1 (t1) and x2
 * (t2).  The basis is device memory of its own; destroy it before its context.  Errors:
 * trt_last_error(ctx). */
typedef struct trt_pod trt_pod;
int trt_pod_create(trt_ctx *ctx, trt_pod **out, const char *name, int gsum, int nt, const int *rank,
                   const double *C, const double *S, const double *U, const double *V);
int trt_pod_reconstruct(trt_pod *pod, int t1, int t2, double x, double x1, double x2);
int trt_pod_destroy(trt_pod *pod);
/* For C_g (POD_RECONSTRUCT_BCg / DMD_RECONSTRUCT_BCg, MainSrc/POD_ROUTINES.f08:253-353) the
 * basis reconstructs into BCg_next, shape (2(N_x+N_y), N_g): per group the left side
 * (N_y values), the bottom (N_x), the right (N_y), the top (N_x); each reconstruction into
 * it is split into Cg_L_next, Cg_B_next, Cg_R_next, Cg_T_next. */
/* timing (no counterpart in the reference): reps reconstructions at data time 1 on the
 * context's stream, ms per reconstruction from HIP events; bytes of U */
int trt_pod_timed(trt_pod *pod, int reps, double *ms);
size_t trt_pod_basis_bytes(const trt_pod *pod);
/* DMD reconstruction of one field of the ROM runs: DMD_RECONSTRUCT,
 * MainSrc/DMD_ROUTINES.f08:28-103, as DMD_RECONSTRUCT_fg calls it (:131): the real part
 * of sum_r B_r W_r exp(L_r t), plus C.  L, B (eigenvalues, amplitudes; sum of ranks) and W
 * (modes, in trt_pod_create's U layout) are complex, stored as (re, im) double pairs as
 * Fortran's COMPLEX*16; rank as for trt_pod_create.  trt_dmd_reconstruct writes the field
 * at time t (the reference passes Time - Start_Time_d, ALGORITHMS.f08:617). */
typedef struct trt_dmd trt_dmd;
int trt_dmd_create(trt_ctx *ctx, trt_dmd **out, const char *name, int gsum, const int *rank, const double *C,
                   const double *L, const double *B, const double *W);
int trt_dmd_reconstruct(trt_dmd *dmd, double time);
int trt_dmd_destroy(trt_dmd *dmd);
double trt_time(const trt_ctx *ctx);

/* The ROM runs' data grid: GENERATE_GRIDS_XYT (MainSrc/GRID_FUNCTIONS.f08:661-817) for the
 * context's simulation grid and a data grid of dnx x dny cells (delx_d, dely_d; xlen_d,
 * ylen_d) with dnt data times (delt_d after start_time_d), and nt simulation times (dt
 * after start_time); xlen, ylen are the simulation domain's sizes as the reference passes
 * them.  The data grid holds N_g groups, as the simulation.  Creation registers the
 * data-grid arrays dat_fg_avg_xx, dat_fg_avg_yy (dnx,dny,N_g), dat_fg_edgV_xx,
 * dat_fg_edgV_xy (dnx+1,dny,N_g), dat_fg_edgH_yy, dat_fg_edgH_xy (dnx,dny+1,N_g) and dat_BCg
 * (2(dnx+dny),N_g) on the context -- POD and DMD bases reconstruct into them -- and computes
 * the point maps (MAP_GRIDS, :222-327) on the GPU; a simulation point outside the data
 * domain is an error.  trt_romgrid_tmap gives, for simulation time t (1-based), the data
 * times the reconstruction uses and the interpolation times (TMap, :754-817, as
 * POD_RECONSTRUCT_fg passes them to POD_RECONSTRUCT and LINEAR_INTERPOLATION,
 * POD_ROUTINES.f08:66-73): call trt_pod_reconstruct(b, t1, t2, x, x1, x2) with them.
 * trt_romgrid_map appends the boundary values the reference appends (EXTND_BT_BND,
 * EXTND_LR_BND, AVG_APND_BND with COLLECT_LR_BND / COLLECT_BT_BND / COLLECT_CRN_BND,
 * GRID_FUNCTIONS.f08:406-659, combined as POD_RECONSTRUCT_fg does, POD_ROUTINES.f08:81-232)
 * and maps the data-grid fields onto the simulation grid (FMAP, :14-43) into fg_*_next and,
 * with with_bcg = 1, Cg_*_next from dat_BCg (POD_RECONSTRUCT_BCg, :253-353; needs
 * N_x = N_y and dnx = dny, as the reference's side maps do).  trt_romgrid_maps copies
 * layout's maps and point grids to the host (layouts 0 cells, 1 V faces, 2 H faces, 3..6
 * left, bottom, right, top sides; 4 ints per simulation point).  The context frees a data
 * grid still open when it is destroyed.  Errors: trt_last_error(ctx). */
typedef struct trt_romgrid trt_romgrid;
int trt_romgrid_create(trt_ctx *ctx, trt_romgrid **out, int dnx, int dny, const double *delx_d, const double *dely_d,
                       double xlen_d, double ylen_d, int dnt, double delt_d, double start_time_d, int nt,
                       double start_time, double xlen, double ylen);
int trt_romgrid_tmap(const trt_romgrid *r, int t, int *t1, int *t2, double *x, double *x1, double *x2);
int trt_romgrid_map(trt_romgrid *r, int with_bcg);
int trt_romgrid_maps(const trt_romgrid *r, int layout, int *gmap, int *vmap, double *sim_grid, double *dat_grid);
int trt_romgrid_destroy(trt_romgrid *r);

/* Checkpoint and resume.  trt_restart_out ≙ RESTART_OUT, MainSrc/OUTPUTS.f08:1624-1674
 * (called every restart_freq-th step, ALGORITHMS.f08:889-894); trt_restart_in ≙
 * RESTART_IN, MainSrc/INPUTS.f08:1734-1783 (before the time loop, ALGORITHMS.f08:234-237).
 * The file is the reference's: 25 Fortran unformatted sequential records in its order
 * (Time, Temp, I_crn, fg_avg_xx, fg_avg_yy, fg_edgV_xx, fg_edgV_xy, fg_edgH_yy,
 * fg_edgH_xy, RT_Src, MGQD_Src, GQD_Src, E_avg, Eg_avg, Eg_edgV, Eg_edgH, Fxg_edgV,
 * Fyg_edgH, KapE_Bar, KapE, KapR, Cg_L, Cg_B, Cg_R, Cg_T), FP64, Fortran shapes;
 * 4-byte length markers, records above 2147483639 bytes split into subrecords
 * (negated markers) as gfortran does; TRT_SUBRECORD_LENGTH lowers the split size.
 * RT_Src is written in the reference's shape (2N_x,2N_y,N_m,N_g), the same value at
 * every corner and direction; trt_restart_in rejects a file whose RT_Src is not of
 * that form.  trt_restart_in reads and checks the whole file (record lengths against
 * the context's grid, quadrature and group count) before it changes the context: on
 * an error the context is unchanged.  Intensities of an FP32 context are written as
 * their exact FP64 values, so an FP32 checkpoint resumes bitwise. */
int trt_restart_out(trt_ctx *ctx, const char *path);
int trt_restart_in(trt_ctx *ctx, const char *path);
/* RESTART_TIME_IN, MainSrc/INPUTS.f08:1785-1803: the first record (Time) of a
 * restart file, without a context.  Errors: trt_frecords_last_error(). */
int trt_restart_time_in(const char *path, double *time);
/* Host-only Fortran unformatted sequential records of FP64 arrays, the framing of
 * the restart file (no device, no context; no counterpart in the reference, which
 * uses its Fortran runtime's WRITE/READ).  write: nrec records, record k holding
 * count[k] doubles from data[k], to a new file (append = 0) or after the end of
 * an existing one.  read: the first nrec records of the file, each of which must
 * hold exactly count[k] doubles.  Both return 0 or a negative status with
 * trt_frecords_last_error() (thread-local). */
int trt_frecords_write(const char *path, int nrec, const double *const *data, const size_t *count, int append);
int trt_frecords_read(const char *path, int nrec, double *const *data, const size_t *count);
const char *trt_frecords_last_error(void);

/* ---- Output files -----------------------------------------------------------
 * The reference's NetCDF output file, written from the context's device state.
 * trt_outfile_create ≙ OUTFILE_INIT + OUTFILE_VARDEFS (MainSrc/OUTPUTS.f08:13,839; kind 0,
 * the run file of run_type mlqd, mg_pod, mg_dmd) or OUTFILE_INIT_Tgen (:412; kind 1, the
 * Tgen_QDf file); trt_timestep_outs ≙ TIMESTEP_OUTS (:1429) at time index t (1-based) --
 * with after_step = 1 also the per-step writes of TRT_MLQD_ALGORITHM after a time step
 * (MainSrc/ALGORITHMS.f08:859-869: E_xWave_Speed, Temp_xWave_Speed, and the iteration
 * counts and convergence norms of the step, ALGORITHMS.f08:688-829); 0 for the initial
 * state (Init_out, :268-270); trt_tgen_outs ≙ the writes of Tgen_QDf's time loop
 * (:1036-1095) after trt_tgen_step; trt_outfile_put ≙ the driver's own nf90_put_var calls
 * (PODgsum, rrank_*, ..., :449-506), Fortran order, 1-based starts; trt_outfile_close ≙
 * NF_CLOSE_FILE (MainSrc/NCDF_IO.f08:120).  Dimension, variable and attribute names,
 * types, units and shapes are the reference's (netCDF readers see the Fortran dimensions
 * reversed, e.g. Temperature(N_t,N_y,N_x)).  The container is the NetCDF classic 64-bit
 * offset format (CDF-5, 64-bit data, when a variable exceeds 4 GiB) instead of
 * NETCDF4/HDF5 (NCDF_IO.f08:71); unwritten values hold the default fill values.  The
 * reference's three unlimited iteration dimensions (RT_Its-dim, MGQD_Its-dim,
 * GQD_Its-dim) become fixed ones as long as the longest iteration sequence the run wrote;
 * they and the iteration variables are added at close.  Res_Calc residual variables are
 * not written (the GPU path has no Res_Calc).  The file keeps a pointer to the context:
 * close it first.  Errors: trt_outfile_last_error (trt_last_error(NULL) after a failed
 * create or close). */
typedef struct trt_outfile trt_outfile;
typedef struct {
    int kind;                          /* 0: run file; 1: Tgen_QDf file */
    const char *run_type, *quadrature, *enrgy_strc;   /* attribute text (trailing blanks dropped) */
    double xlen, ylen;                 /* domain (0: sums of Delx, Dely) */
    double tlen, start_time;
    int n_t;                           /* time points; 0: NINT((tlen - start_time)/dt), +1 with
                                          init_out (MainSrc/INPUTS.f08:120-123) */
    int init_out;
    int I_out, HO_Eg_out, HO_Fg_out, HO_E_out, HO_F_out, Eg_out, Fg_out, MGQD_E_out, MGQD_F_out,
        QDfg_out, E_out, F_out, D_out, old_parms_out, its_out, conv_out, kap_out, Src_out;
    int nofill;                        /* 1: no fill values (nf90_set_fill NOFILL) */
    int format;                        /* 0: CDF-2, or CDF-5 when a variable exceeds 4 GiB; 5: CDF-5 */
    /* ROM runs (run_type mg_pod / mg_dmd, OUTPUTS.f08:98-136) */
    const char *pod_type, *dmd_type;
    int direc_diff, n_dsets;
    const double *dset_times;          /* n_dsets values */
    const char *const *dsets;          /* n_dsets database names */
    const double *tpts;                /* n_t time points (kind 1: the temperature file's);
                                          NULL: from start_time and dt (INPUTS.f08:144-151) */
} trt_outopts;
int trt_outfile_create(trt_ctx *ctx, const char *path, const trt_outopts *opts, trt_outfile **out);
int trt_timestep_outs(trt_outfile *f, int t, int after_step);
int trt_tgen_outs(trt_outfile *f, int t);
int trt_outfile_put(trt_outfile *f, const char *var, int ndims, const size_t *start, const size_t *count,
                    const double *vals);
int trt_outfile_close(trt_outfile *f);
const char *trt_outfile_last_error(const trt_outfile *f);
/* Reading a classic-format file (CDF-1, CDF-2, CDF-5): the reads of NF_INQ_DIM and
 * NF_INQ_VAR_nD (MainSrc/NCDF_IO.f08:339-460), e.g. Tgen_QDf's temperature input
 * (ALGORITHMS.f08:1038).  Fortran order, 1-based starts, values converted to FP64.
 * No device, no context. */
typedef struct trt_ncread trt_ncread;
int trt_nc_open(const char *path, trt_ncread **out);
int trt_nc_dim(trt_ncread *r, const char *name, size_t *len);
int trt_nc_get(trt_ncread *r, const char *var, int ndims, const size_t *start, const size_t *count, double *out);
/* attribute name of variable var (NULL: a global attribute), as nf90_get_att reads it:
 * numbers (up to cap, as FP64) into vals and 0 returned, or text (NUL-terminated, at most
 * tcap - 1 characters) into text and 1 returned; *n receives the attribute's length */
int trt_nc_att(trt_ncread *r, const char *var, const char *name, double *vals, size_t cap, char *text, size_t tcap,
               size_t *n);
const char *trt_nc_last_error(const trt_ncread *r);
int trt_nc_close(trt_ncread *r);

/* Sweep timing: runs nsweeps sweeps on the context's stream and returns the
 * total elapsed milliseconds measured with HIP events on that stream. */
int trt_sweep_timed(trt_ctx *ctx, int nsweeps, double *ms);
/* Sweep timing per kernel launch: an event pair around each of the sweep's
 * tile-diagonal launches.  wall_ms spans the whole sweeps, kernel_ms sums the
 * launches, launches counts them. */
int trt_sweep_kernel_timed(trt_ctx *ctx, int nsweeps, double *wall_ms, double *kernel_ms, int *launches);
/* Device-to-device copy of nbytes by a plain kernel, reps times; *ms = mean
 * duration.  The calibration reference for HBM byte counters. */
int trt_calib_copy(size_t nbytes, int reps, double *ms);
/* Accumulated sweep statistics since creation: sweeps run and total sweep ms. */
int trt_sweep_stats(const trt_ctx *ctx, long long *sweeps, double *ms);
/* Accumulated direct QD-solve statistics since creation (qd_solver 0 and 3): solves that
 * factorized, solves on reused factors, and the refinement steps those took. */
int trt_qd_stats(const trt_ctx *ctx, long long *factorized, long long *reused, long long *refine_steps);
/* The sweep configuration in use: sweep_mode (1, 2, 3 or 4), tile shape, and the
 * persistent launch's grid (workgroups; 0 before the first persistent sweep). */
int trt_sweep_config(const trt_ctx *ctx, int *mode, int *tile_x, int *tile_y, int *grid);

/* One step of the sweep-only entry Tgen_QDf (MainSrc/ALGORITHMS.f08:1031-1095)
 * with the Temp array prescribed: MATERIAL_UPDATE, I_crn_old = I_crn,
 * TRANSPORT_SCB, COLLAPSE_INTENSITIES, fg_Calc, Cg_Calc.  No low-order solve. */
int trt_tgen_step(trt_ctx *ctx);

/* ---- Batches ----------------------------------------------------------------
 * No counterpart in the reference, which sweeps one problem per run.  A batch
 * sweeps many contexts (uneven grids, group counts and quadratures; one sweep
 * precision and tile shape) in one persistent launch.  The contexts stay owned
 * by the caller, must outlive the batch, and keep their own streams: each
 * batched sweep starts after the work already enqueued on every member's
 * stream, and later work on a member's stream waits for it. */
typedef struct trt_batch trt_batch;
int trt_batch_create(trt_batch **out, trt_ctx *const *ctxs, int n);
/* As trt_batch_create with the column-segment length of the work items fixed:
 * seg tiles of one tile column per item, their H faces carried in registers
 * (1 = single tiles; 0 = chosen from the available parallelism, as
 * trt_batch_create does).  Every length gives bitwise the same results. */
int trt_batch_create_seg(trt_batch **out, trt_ctx *const *ctxs, int n, int seg);
/* the column-segment length the batch uses (-1 for a null batch) */
int trt_batch_segment_length(const trt_batch *b);
int trt_batch_destroy(trt_batch *b);
const char *trt_batch_last_error(const trt_batch *b);
/* nsweeps batched TRANSPORT_SCB sweeps (each as trt_transport_scb on every
 * member); ms receives their time on the batch stream */
int trt_batch_sweep_timed(trt_batch *b, int nsweeps, double *ms);
/* one Tgen_QDf step of every member, the sweeps in one launch; sweep_ms (may
 * be NULL) receives the batched sweep's time (with whole-step graphs: the step's) */
int trt_batch_tgen_step(trt_batch *b, double *sweep_ms);
/* on = 1: the batched sweep is captured once in a HIP graph and replayed; on = 2: the
 * whole batched Tgen_QDf step -- every member's MATERIAL_UPDATE on its stream, the
 * persistent sweep, every member's COLLAPSE_INTENSITIES + fg_Calc + Cg_Calc -- is one graph
 * launch, captured once per combination of the members' I_crn/I_crn_old buffers (two
 * for a batch stepping together) and replayed; 0 = eager.  Results are bitwise the same
 * in every mode. */
int trt_batch_use_graph(trt_batch *b, int on);
/* One full time step of TRT_MLQD_ALGORITHM (as trt_step) of every member.  The members
 * iterate in lockstep rounds of RT iterations: each round's sweeps, of the members whose
 * RT iteration starts with one, run as one persistent launch (the others' work items are
 * skipped); then every member still iterating runs the rest of its RT iteration -- QD
 * factors, MLOQD and grey Newton iterations, convergence checks -- on its own stream from
 * a host thread of its own.  its (may be NULL) receives 4 counts per member (RT, MLOQD,
 * grey Newton, sweeps).  A member's results are bitwise those of trt_step on it alone. */
int trt_batch_step(trt_batch *b, int *its);
int trt_batch_info(const trt_batch *b, long long *updates_per_sweep, int *grid, int *nitems);

/* Tile-shape autotuner (no counterpart in the reference).  Creates contexts
 * for the n problems with each tile shape the tile engine offers for the
 * precision (trt_autotune_shapes lists them; their count is the return value),
 * times `sweeps` batched sweeps of each and returns the fastest shape.
 * all_ms (may be NULL) receives the ms per sweep of every shape in list order.
 * The problems' own precision, tile and sweep_mode fields are ignored. */
int trt_autotune_tile(const trt_problem *probs, int n, int precision, int sweeps, int *tile_x, int *tile_y,
                      double *best_ms, double *all_ms);
int trt_autotune_shapes(int precision, int *tile_x, int *tile_y, int cap);
/* As trt_autotune_tile, jointly over tile shape and the column-segment length
 * of trt_batch_create_seg (trt_autotune_segments lists the lengths tried; their
 * count is the return value).  all_ms (may be NULL) receives the ms per sweep
 * of shape s with length q at [s * nsegments + q]. */
int trt_autotune_batch(const trt_problem *probs, int n, int precision, int sweeps, int *tile_x, int *tile_y,
                       int *seg, double *best_ms, double *all_ms);
int trt_autotune_segments(int *seg, int cap);
const char *trt_autotune_last_error(void);

/* ---- Request loop -----------------------------------------------------------
 * No counterpart in the reference.  Requests (a problem, an optional
 * prescribed Temp(N_x,N_y) field, a number of Tgen_QDf steps, an arrival time in
 * seconds from the start of trt_server_run) are served with continuous
 * batching: at every tick the loop admits arrived requests in arrival order
 * while the active set's device bytes stay within hbm_budget and its size
 * within max_active, runs one Tgen_QDf step of every active request with all
 * sweeps in one persistent launch, and retires finished requests (outputs
 * copied to host, device state freed).  A request larger than the whole budget
 * fails (status -1).  The problem's vectors are copied at submission.
 * In solve mode (trt_server_set_mode 1) a request's steps are full time steps of
 * TRT_MLQD_ALGORITHM from the initial state (as trt_step; no temperature field):
 * admission allocates its low-order state too, and each tick is trt_batch_step over
 * the active requests (each round of RT iterations' sweeps in one launch, every
 * request's low-order work on its own stream).  A request's results are the same
 * alone and in any mix, in both modes. */
typedef struct trt_server trt_server;
typedef struct {
    double arrival_s, admit_s, done_s;  /* seconds from the start of the run; -1 if never */
    long long updates;                  /* sweep cell-angle-group updates performed */
    size_t device_bytes;                /* device bytes held while active */
    int steps;                          /* steps completed */
    int status;                         /* 0 done, 1 waiting, 2 active, -1 failed */
} trt_request_stats;
typedef struct {
    double wall_s;                      /* duration of the last run */
    long long ticks, updates;
    double sweep_ms;                    /* summed batched-sweep time */
    size_t peak_bytes;                  /* largest device bytes of the active set */
    int peak_active;                    /* largest active set */
    double admit_s, rebuild_s, step_s, retire_s;   /* host seconds in admission (context creation),
                                                      batch rebuilds, steps and retirement */
    double setup_s;                     /* seconds to reserve the arena of hbm_budget bytes (done
                                           once, before the first run's clock starts; held until
                                           trt_server_destroy) */
} trt_server_stats;
int trt_server_create(trt_server **out, size_t hbm_budget, int max_active);
/* 0 (default): requests are Tgen_QDf steps; 1: requests are solves (full time steps) */
int trt_server_set_mode(trt_server *s, int mode);
int trt_server_destroy(trt_server *s);
const char *trt_server_last_error(const trt_server *s);
int trt_server_submit(trt_server *s, const trt_problem *p, const double *temp, int steps, double arrival_s,
                      int *id);
int trt_server_run(trt_server *s);
int trt_server_request(const trt_server *s, int id, trt_request_stats *out);
const char *trt_server_request_error(const trt_server *s, int id);
int trt_server_summary(const trt_server *s, trt_server_stats *out);
/* outputs of a finished request: Temp, HO_E_avg, HO_Eg_avg, HO_Eg_edgV,
 * HO_Eg_edgH, the seven fg_* factors and Cg_L/B/R/T, Fortran shapes */
size_t trt_server_output_size(const trt_server *s, int id, const char *name);
int trt_server_output(const trt_server *s, int id, const char *name, double *host, size_t n);

/* No counterpart in the reference: CONVERGENCE (MainSrc/CONVERGENCE_CHECKS.f08:7-44) for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  x, x_old, x_old2: n values each in the caller's
 * device memory; out: 3 doubles of the caller's device memory receiving {norm, rho,
 * converged (1.0 or 0.0)}, the values trt_convergence returns; stream: a hipStream_t (NULL:
 * the default stream).  The call only enqueues two kernels on the stream: it allocates
 * nothing, copies nothing to the host and does not synchronise.  The reduction's partials
 * sit in a fixed device array of 4 KiB per GPU held until the process ends; n is limited
 * to 2147483647.  Calls that overlap on one GPU share that array and race.  Errors
 * (negative return): the message is trt_last_error(NULL). */
int trt_convergence_noalloc(const double *x, const double *x_old, const double *x_old2, size_t n, int conv_type,
                            double eps, int its, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
