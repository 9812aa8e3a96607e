/*
 * trt_oracle.h -- FP64 CPU restatement of the full-order MLQD iteration of the
 * 2D TRT reference solver (the MainSrc Fortran files).  TEST INFRASTRUCTURE ONLY: this
 * library is the checker for the HIP implementation; it is loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * product path.
 *
 * Array layouts follow the reference exactly (Fortran column-major, first
 * index fastest) so that any array can be compared element-for-element with
 * the reference's own shapes, e.g. I_crn(2N_x,2N_y,N_m,N_g).  All indices in
 * this C code are 0-based.
 */
#ifndef TRT_ORACLE_H
#define TRT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Problem description (inputs of MainSrc/INPUTS.f08 after parsing). */
typedef struct {
    int nx, ny, nm, ng, nmat;
    const double *dx, *dy;          /* Delx(N_x), Dely(N_y) [cm]                      */
    const double *ox, *oy, *w;      /* Omega_x, Omega_y, quad_weight (N_m), sum w = 4pi */
    const double *nu;               /* nu_g(N_g+1) group edges [eV]                   */
    const int *mat;                 /* Mat(N_x,N_y), 0-based material ids             */
    const double *kmult;            /* Kappa_Mult(nmat)                               */
    double c, dt, theta, comp_unit, cv, kap0, tini;
    int bc_type[4];                 /* (left, bottom, right, top), 0 = incoming Planckian */
    double bcT[4];                  /* boundary temperatures [eV]                     */
    int opacity_model;              /* 0 = kapB_calc group Planck mean, 1 = group-midpoint formula */
    int maxit_rte, maxit_mloqd, maxit_gloqd, conv_type;
    double conv_ho, conv_lo, conv_gr1, conv_gr2, chi, line_src;
    int use_line_search, use_safety_search, kapE_dT_flag;
    double E_bound_low, T_bound_low;
    int threads;
    /* options beyond the reference's defaults */
    int grey_emission;              /* 0 = the reference's analytic grey emission 15 kap0 c aR T / pi^4
                                       (GLOQD_SOLVES.f08:439-440); 1 = sum_g 4 pi kappa_g B_g, consistent
                                       with the group sources (a labelled deviation)                 */
    int no_grey;                    /* 0 = grey Newton (use_grey = 1); 1 = use_grey = 0: MEB_SOLVE in the
                                       MLOQD loop (ALGORITHMS.f08:767-770)                            */
} orc_problem;

typedef struct orc_state orc_state;

/* scalar physics (TEMP_FUNCTIONS.f08) */
double orc_sigint(double z);
double orc_Bg_planck(double T, double nu_low, double nu_high, double comp_unit);
double orc_kapB(double T, double nu_low, double nu_high);
double orc_kap_midpoint(double T, double nu_low, double nu_high);

/* state life cycle: MISC_INIT, RT_INIT, MGQD_INIT, GQD_INIT, TEMP_INIT (ALGORITHMS.f08:203-218) */
orc_state *orc_create(const orc_problem *p);
void orc_destroy(orc_state *s);

/* array access by reference variable name; returns NULL if unknown. dims[4] gets
 * the Fortran shape (unused trailing dims = 1). */
double *orc_array(orc_state *s, const char *name, int dims[4]);
int *orc_iarray(orc_state *s, const char *name, int dims[4]);

/* one reference routine per call, operating on the state's arrays */
void orc_material_update(orc_state *s);
/* MEB_SOLVE, UPDATES.f08:189-231: closed-form grey temperature from Eg_avg */
void orc_meb_solve(orc_state *s);
void orc_transport_scb(orc_state *s);
void orc_collapse_intensities(orc_state *s);
void orc_fg_calc(orc_state *s);
void orc_cg_calc(orc_state *s);
void orc_mgqd_in_calc(orc_state *s);
void orc_gqd_in_calc(orc_state *s);
/* RT_BC_UPDATE, UPDATES.f08:12-105 (BC_Type 0: Planckian inflow, 1: reflective) */
void orc_rt_bc_update(orc_state *s);
void orc_old_mgqd_coefs(orc_state *s);
void orc_old_grey_coefs(orc_state *s);
void orc_mloqd_fv(orc_state *s);
void orc_collapse_mg_ef(orc_state *s);
void orc_cbar_calc(orc_state *s);
void orc_grey_coefs(orc_state *s);
int  orc_egp_fv_newt(orc_state *s, int mgqd_it);
int  orc_convergence(int conv_type, double eps, const double *x, const double *x_old,
                     const double *x_old2, int n, int its, double *norm, double *rho);

/* one full time step of TRT_MLQD_ALGORITHM (mlqd, use_grey=1); returns the
 * number of outer (RT) iterations.  it_out (may be NULL) receives
 * {RT_Its, total MLOQD its, total Newton its, total sweeps}. */
int orc_step(orc_state *s, int it_out[4]);
int orc_step_ex(orc_state *s, int it_out[4], int ho_form);
double orc_time(orc_state *s);

/* the reduced QD system (QD_SOLVERS.f08:40-538): build the CSR matrix and RHS in
 * the reference unknown ordering.  Coefficient arrays use the reference shapes:
 * EB_*, MBx_C/B/T, MBy_C/L/R, EB_RHS (N_x,N_y); MBx_R, MBx_RHS (N_x-1,N_y) stored in
 * (N_x,N_y) arrays; MBy_T, MBy_RHS (N_x,N_y-1) stored in (N_x,N_y) arrays; Cp_*, BC_*
 * boundary vectors.  coef[] order: EB_L,EB_B,EB_C,EB_R,EB_T,MBx_C,MBx_R,MBx_B,MBx_T,
 * MBy_C,MBy_T,MBy_L,MBy_R,EB_RHS,MBx_RHS,MBy_RHS (16 arrays of N_x*N_y),
 * bnd[] order: Cp_L,Cp_B,Cp_R,Cp_T,BC_L,BC_B,BC_R,BC_T.
 * Returns nnz; ia (nrows+1), ja, a, rhs must have room for 19NxNy+Nx+Ny entries. */
int orc_qd_build_csr(int nx, int ny, const double *const coef[16], const double *const bnd[8],
                     int *ia, int *ja, double *a, double *rhs);
/* solve the reduced QD system by banded LU with partial pivoting and scatter the
 * solution to E_avg(N_x,N_y), E_edgV(N_x+1,N_y), E_edgH(N_x,N_y+1). Returns 0 on success. */
/* test hook: capture one group's MLOQD system coefficients, optionally skipping the solves */
void orc_set_qd_capture(int group, double *buf, int skip_solve);
/* sensitivity studies: relative perturbation of every QD solution (0 = off) */
void orc_set_lin_noise(double eps);
int orc_qd_fv(int nx, int ny, const double *const coef[16], const double *const bnd[8],
              double *E_avg, double *E_edgV, double *E_edgH);

#ifdef __cplusplus
}
#endif
#endif
