!==================================================================================================================================!
! TRT_CAPI: ISO_C_BINDING interfaces to the C ABI of include/trt_capi.h, plus the context the drop-in
! modules share.  Errors follow the reference convention: a failing call STOPs with the library's message.
!==================================================================================================================================!
MODULE TRT_CAPI

  USE ISO_C_BINDING
  IMPLICIT NONE

  TYPE, BIND(C) :: trt_problem
    INTEGER(C_INT) :: nx, ny, nm, ng, nmat
    TYPE(C_PTR) :: dx, dy, ox, oy, w, nu, mat, kmult
    REAL(C_DOUBLE) :: c, dt, theta, comp_unit, cv, kap0, tini
    INTEGER(C_INT) :: bc_type(4)
    REAL(C_DOUBLE) :: bcT(4)
    INTEGER(C_INT) :: opacity_model
    INTEGER(C_INT) :: maxit_rte, maxit_mloqd, maxit_gloqd, conv_type
    REAL(C_DOUBLE) :: conv_ho, conv_lo, conv_gr1, conv_gr2, chi, line_src
    INTEGER(C_INT) :: use_line_search, use_safety_search, kapE_dT_flag
    REAL(C_DOUBLE) :: E_bound_low, T_bound_low
    INTEGER(C_INT) :: threads, precision, tile_x, tile_y
    REAL(C_DOUBLE) :: lin_rtol
    INTEGER(C_INT) :: lin_maxit
    INTEGER(C_INT) :: sweep_mode
    INTEGER(C_INT) :: qd_solver      ! 0 nested dissection, 1 BiCGSTAB, 2 block LU
    INTEGER(C_INT) :: grey_emission  ! 0 analytic (reference), 1 sum over groups
    INTEGER(C_INT) :: no_grey        ! 1: use_grey = 0 (MEB_SOLVE)
  END TYPE trt_problem

  ! the context of the running problem, created by TRT_GPU_INIT
  TYPE(C_PTR), SAVE :: TRT_CTX = C_NULL_PTR

  ! options of the output file (trt_outopts); flags in the order I_out, HO_Eg_out, HO_Fg_out, HO_E_out,
  ! HO_F_out, Eg_out, Fg_out, MGQD_E_out, MGQD_F_out, fg_out, E_out, F_out, D_out, old_parms_out, its_out,
  ! conv_out, kap_out, Src_out
  TYPE, BIND(C) :: trt_outopts
    INTEGER(C_INT) :: kind
    TYPE(C_PTR) :: run_type, quadrature, enrgy_strc
    REAL(C_DOUBLE) :: xlen, ylen, tlen, start_time
    INTEGER(C_INT) :: n_t, init_out
    INTEGER(C_INT) :: flags(18)
    INTEGER(C_INT) :: nofill, format
    TYPE(C_PTR) :: pod_type, dmd_type
    INTEGER(C_INT) :: direc_diff, n_dsets
    TYPE(C_PTR) :: dset_times, dsets, tpts
  END TYPE trt_outopts

  ! the output file of the running problem, created by TRT_GPU_OUTFILE_INIT
  TYPE(C_PTR), SAVE :: TRT_OUT = C_NULL_PTR

  INTERFACE
    INTEGER(C_INT) FUNCTION trt_ctx_create(ctx, p) BIND(C, name="trt_ctx_create")
      IMPORT :: C_INT, C_PTR, trt_problem
      TYPE(C_PTR) :: ctx
      TYPE(trt_problem) :: p
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_ctx_destroy(ctx) BIND(C, name="trt_ctx_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    TYPE(C_PTR) FUNCTION trt_last_error(ctx) BIND(C, name="trt_last_error")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_SIZE_T) FUNCTION trt_array_size(ctx, name) BIND(C, name="trt_array_size")
      IMPORT :: C_SIZE_T, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: name(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_push(ctx, name, host, n) BIND(C, name="trt_push")
      IMPORT :: C_INT, C_PTR, C_CHAR, C_DOUBLE, C_SIZE_T
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: name(*)
      REAL(C_DOUBLE) :: host(*)
      INTEGER(C_SIZE_T), VALUE :: n
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_pull(ctx, name, host, n) BIND(C, name="trt_pull")
      IMPORT :: C_INT, C_PTR, C_CHAR, C_DOUBLE, C_SIZE_T
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: name(*)
      REAL(C_DOUBLE) :: host(*)
      INTEGER(C_SIZE_T), VALUE :: n
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_material_update(ctx) BIND(C, name="trt_material_update")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_transport_scb(ctx) BIND(C, name="trt_transport_scb")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_rt_bc_update(ctx) BIND(C, name="trt_rt_bc_update")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_collapse_intensities(ctx) BIND(C, name="trt_collapse_intensities")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_fg_calc(ctx) BIND(C, name="trt_fg_calc")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_cg_calc(ctx) BIND(C, name="trt_cg_calc")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_mgqd_in_calc(ctx) BIND(C, name="trt_mgqd_in_calc")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_old_mgqd_coefs(ctx) BIND(C, name="trt_old_mgqd_coefs")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_old_grey_coefs(ctx) BIND(C, name="trt_old_grey_coefs")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_mloqd_fv(ctx, kits) BIND(C, name="trt_mloqd_fv")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT) :: kits(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_collapse_mg_ef(ctx) BIND(C, name="trt_collapse_mg_ef")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_cbar_calc(ctx) BIND(C, name="trt_cbar_calc")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_grey_coefs(ctx) BIND(C, name="trt_grey_coefs")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_gqd_in_calc(ctx) BIND(C, name="trt_gqd_in_calc")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_egp_fv_newt(ctx, mgqd_it, its) BIND(C, name="trt_egp_fv_newt")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: mgqd_it
      INTEGER(C_INT) :: its
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_convergence(ctx, x, xo, xo2, conv_type, eps, its, norm, rho, conv) &
        BIND(C, name="trt_convergence")
      IMPORT :: C_INT, C_PTR, C_CHAR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: x(*), xo(*), xo2(*)
      INTEGER(C_INT), VALUE :: conv_type, its
      REAL(C_DOUBLE), VALUE :: eps
      REAL(C_DOUBLE) :: norm, rho
      INTEGER(C_INT) :: conv
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_meb_solve(ctx, eg) BIND(C, name="trt_meb_solve")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: eg(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_step(ctx, its) BIND(C, name="trt_step")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT) :: its(4)
    END FUNCTION
    REAL(C_DOUBLE) FUNCTION trt_time(ctx) BIND(C, name="trt_time")
      IMPORT :: C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    ! one step of the sweep-only entry Tgen_QDf (ALGORITHMS.f08:1031-1095), Temp prescribed
    INTEGER(C_INT) FUNCTION trt_tgen_step(ctx) BIND(C, name="trt_tgen_step")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    ! the device state as the reference's restart file (RESTART_OUT, OUTPUTS.f08:1624-1674) and back
    ! (RESTART_IN, INPUTS.f08:1734-1783); path is NUL-terminated
    INTEGER(C_INT) FUNCTION trt_restart_out(ctx, path) BIND(C, name="trt_restart_out")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: path(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_restart_in(ctx, path) BIND(C, name="trt_restart_in")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: path(*)
    END FUNCTION
    ! the reference's NetCDF output file (OUTFILE_INIT + OUTFILE_VARDEFS, OUTPUTS.f08:13,839; TIMESTEP_OUTS :1429)
    INTEGER(C_INT) FUNCTION trt_outfile_create(ctx, path, opts, f) BIND(C, name="trt_outfile_create")
      IMPORT :: C_INT, C_PTR, C_CHAR, trt_outopts
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR) :: path(*)
      TYPE(trt_outopts) :: opts
      TYPE(C_PTR) :: f
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_timestep_outs(f, t, after_step) BIND(C, name="trt_timestep_outs")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: f
      INTEGER(C_INT), VALUE :: t, after_step
    END FUNCTION
    INTEGER(C_INT) FUNCTION trt_outfile_close(f) BIND(C, name="trt_outfile_close")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: f
    END FUNCTION
    TYPE(C_PTR) FUNCTION trt_outfile_last_error(f) BIND(C, name="trt_outfile_last_error")
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: f
    END FUNCTION
  END INTERFACE

CONTAINS

!----------------------------------------------------------------------------------------------------------------------------------!
! message of the last failure of the running context (or of trt_ctx_create)
!----------------------------------------------------------------------------------------------------------------------------------!
FUNCTION TRT_ERRMSG() RESULT(msg)
  CHARACTER(LEN=512) :: msg
  TYPE(C_PTR) :: p
  CHARACTER(KIND=C_CHAR), POINTER :: s(:)
  INTEGER :: k
  msg = ''
  p = trt_last_error(TRT_CTX)
  IF (.NOT. C_ASSOCIATED(p)) RETURN
  CALL C_F_POINTER(p, s, [512])
  DO k = 1, 512
    IF (s(k) .EQ. C_NULL_CHAR) EXIT
    msg(k:k) = s(k)
  END DO
END FUNCTION TRT_ERRMSG

!----------------------------------------------------------------------------------------------------------------------------------!
! checkpoint / resume of the running context in the reference's restart file format; the device-resident
! counterpart of RESTART_OUT and RESTART_IN, called where the reference calls them (ALGORITHMS.f08:889-894
! every restart_freq-th step, :234-237 before the time loop)
!----------------------------------------------------------------------------------------------------------------------------------!
SUBROUTINE TRT_GPU_RESTART_OUT(restart_outfile)
  CHARACTER(*), INTENT(IN) :: restart_outfile
  CALL TRT_CHECK(trt_restart_out(TRT_CTX, TRIM(restart_outfile)//C_NULL_CHAR), 'RESTART_OUT')
END SUBROUTINE TRT_GPU_RESTART_OUT

SUBROUTINE TRT_GPU_RESTART_IN(restart_infile)
  CHARACTER(*), INTENT(IN) :: restart_infile
  CALL TRT_CHECK(trt_restart_in(TRT_CTX, TRIM(restart_infile)//C_NULL_CHAR), 'RESTART_IN')
END SUBROUTINE TRT_GPU_RESTART_IN

!----------------------------------------------------------------------------------------------------------------------------------!
! the output file of the running context: created where the reference calls OUTFILE_INIT (main.f08:96) and
! OUTFILE_VARDEFS (ALGORITHMS.f08:251), written where it calls TIMESTEP_OUTS (ALGORITHMS.f08:269 for the
! initial state with Init_out = 1, after_step = 0; :872 after each time step t, after_step = 1) and closed
! where it calls NF_CLOSE_FILE (main.f08:121)
!----------------------------------------------------------------------------------------------------------------------------------!
SUBROUTINE TRT_GPU_OUTFILE_INIT(outfile, run_type, quadrature, enrgy_strc, xlen, ylen, tlen, start_time, init_out, flags)
  CHARACTER(*), INTENT(IN) :: outfile, run_type, quadrature, enrgy_strc
  REAL(C_DOUBLE), INTENT(IN) :: xlen, ylen, tlen, start_time
  INTEGER, INTENT(IN) :: init_out, flags(18)
  TYPE(trt_outopts) :: o
  CHARACTER(KIND=C_CHAR, LEN=256), TARGET :: rt, qd, es
  rt = TRIM(run_type)//C_NULL_CHAR
  qd = TRIM(quadrature)//C_NULL_CHAR
  es = TRIM(enrgy_strc)//C_NULL_CHAR
  CALL TRT_REQUIRE_CTX('TRT_GPU_OUTFILE_INIT')
  o%kind = 0
  o%run_type = C_LOC(rt); o%quadrature = C_LOC(qd); o%enrgy_strc = C_LOC(es)
  o%xlen = xlen; o%ylen = ylen; o%tlen = tlen; o%start_time = start_time
  o%n_t = 0; o%init_out = init_out; o%flags = flags
  o%nofill = 0; o%format = 0
  o%pod_type = C_NULL_PTR; o%dmd_type = C_NULL_PTR; o%direc_diff = 0; o%n_dsets = 0
  o%dset_times = C_NULL_PTR; o%dsets = C_NULL_PTR; o%tpts = C_NULL_PTR
  CALL TRT_CHECK(trt_outfile_create(TRT_CTX, TRIM(outfile)//C_NULL_CHAR, o, TRT_OUT), 'OUTFILE_INIT')
END SUBROUTINE TRT_GPU_OUTFILE_INIT

SUBROUTINE TRT_GPU_TIMESTEP_OUTS(t, after_step)
  INTEGER, INTENT(IN) :: t, after_step
  TYPE(C_PTR) :: p
  CHARACTER(KIND=C_CHAR), POINTER :: s(:)
  INTEGER :: k
  IF (trt_timestep_outs(TRT_OUT, INT(t, C_INT), INT(after_step, C_INT)) .NE. 0) THEN
    p = trt_outfile_last_error(TRT_OUT)
    CALL C_F_POINTER(p, s, [512])
    WRITE(*,'(A)', ADVANCE='NO') ' TRT GPU: TIMESTEP_OUTS failed: '
    DO k = 1, 512
      IF (s(k) .EQ. C_NULL_CHAR) EXIT
      WRITE(*,'(A)', ADVANCE='NO') s(k)
    END DO
    WRITE(*,*)
    ERROR STOP 'TRT GPU call failed'
  END IF
END SUBROUTINE TRT_GPU_TIMESTEP_OUTS

SUBROUTINE TRT_GPU_OUTFILE_CLOSE()
  IF (.NOT. C_ASSOCIATED(TRT_OUT)) RETURN
  CALL TRT_CHECK(trt_outfile_close(TRT_OUT), 'NF_CLOSE_FILE')
  TRT_OUT = C_NULL_PTR
END SUBROUTINE TRT_GPU_OUTFILE_CLOSE

SUBROUTINE TRT_CHECK(rc, what)
  INTEGER(C_INT), INTENT(IN) :: rc
  CHARACTER(*), INTENT(IN) :: what
  IF (rc .NE. 0) THEN
    WRITE(*,*) 'TRT GPU: ', what, ' failed: ', TRIM(TRT_ERRMSG())
    ERROR STOP 'TRT GPU call failed'
  END IF
END SUBROUTINE TRT_CHECK

SUBROUTINE TRT_REQUIRE_CTX(what)
  CHARACTER(*), INTENT(IN) :: what
  IF (.NOT. C_ASSOCIATED(TRT_CTX)) THEN
    WRITE(*,*) 'TRT GPU: ', what, ' called before TRT_GPU_INIT'
    ERROR STOP 'TRT GPU context missing'
  END IF
END SUBROUTINE TRT_REQUIRE_CTX

!----------------------------------------------------------------------------------------------------------------------------------!
! host <-> device copy of a reference-named array (any rank; sequence association)
!----------------------------------------------------------------------------------------------------------------------------------!
SUBROUTINE TRT_PUSH_ARRAY(name, a, n)
  CHARACTER(*), INTENT(IN) :: name
  INTEGER, INTENT(IN) :: n
  REAL(C_DOUBLE), INTENT(IN) :: a(n)
  CALL TRT_REQUIRE_CTX(name)
  CALL TRT_CHECK(trt_push(TRT_CTX, TRIM(name)//C_NULL_CHAR, a, INT(n, C_SIZE_T)), 'push '//name)
END SUBROUTINE TRT_PUSH_ARRAY

SUBROUTINE TRT_PULL_ARRAY(name, a, n)
  CHARACTER(*), INTENT(IN) :: name
  INTEGER, INTENT(IN) :: n
  REAL(C_DOUBLE), INTENT(INOUT) :: a(n)
  CALL TRT_REQUIRE_CTX(name)
  CALL TRT_CHECK(trt_pull(TRT_CTX, TRIM(name)//C_NULL_CHAR, a, INT(n, C_SIZE_T)), 'pull '//name)
END SUBROUTINE TRT_PULL_ARRAY

!----------------------------------------------------------------------------------------------------------------------------------!
! TRT_GPU_INIT: creates the device-resident problem from the values the reference reads in INPUT
! (MainSrc/INPUTS.f08) and initialises it as MISC_INIT/RT_INIT/MGQD_INIT/GQD_INIT/TEMP_INIT do
! (MainSrc/ALGORITHMS.f08:203-218).  Mat holds 1-based material ids as in the reference.
!----------------------------------------------------------------------------------------------------------------------------------!
SUBROUTINE TRT_GPU_INIT(Omega_x, Omega_y, quad_weight, Nu_g, Delx, Dely, Mat, Kappa_Mult, c, Delt, Theta, Comp_Unit, &
  Cv, Kap0, Tini, BC_Type, bcT_left, bcT_bottom, bcT_right, bcT_top, opacity_model, maxit_RTE, maxit_MLOQD, maxit_GLOQD, &
  conv_type, conv_ho, conv_lo, conv_gr1, conv_gr2, chi, line_src, Use_Line_Search, Use_Safety_Search, kapE_dT_flag, &
  E_Bound_Low, T_Bound_Low, precision)

  REAL(C_DOUBLE), TARGET, INTENT(IN) :: Omega_x(:), Omega_y(:), quad_weight(:), Nu_g(:), Delx(:), Dely(:), Kappa_Mult(:)
  INTEGER, INTENT(IN) :: Mat(:,:), BC_Type(4), opacity_model, maxit_RTE, maxit_MLOQD, maxit_GLOQD, conv_type, precision
  REAL(C_DOUBLE), INTENT(IN) :: c, Delt, Theta, Comp_Unit, Cv, Kap0, Tini, bcT_left, bcT_bottom, bcT_right, bcT_top
  REAL(C_DOUBLE), INTENT(IN) :: conv_ho, conv_lo, conv_gr1, conv_gr2, chi, line_src, E_Bound_Low, T_Bound_Low
  LOGICAL, INTENT(IN) :: Use_Line_Search, Use_Safety_Search, kapE_dT_flag

  TYPE(trt_problem) :: p
  INTEGER(C_INT), ALLOCATABLE, TARGET :: mat0(:,:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: ox(:), oy(:), w(:), nu(:), dx(:), dy(:), km(:)

  IF (C_ASSOCIATED(TRT_CTX)) CALL TRT_GPU_FINALIZE()
  ALLOCATE(mat0(SIZE(Mat,1), SIZE(Mat,2)))
  mat0 = Mat - 1
  ox = Omega_x; oy = Omega_y; w = quad_weight; nu = Nu_g; dx = Delx; dy = Dely; km = Kappa_Mult

  p%nx = SIZE(Delx); p%ny = SIZE(Dely); p%nm = SIZE(Omega_x); p%ng = SIZE(Nu_g) - 1; p%nmat = SIZE(Kappa_Mult)
  p%dx = C_LOC(dx); p%dy = C_LOC(dy); p%ox = C_LOC(ox); p%oy = C_LOC(oy); p%w = C_LOC(w); p%nu = C_LOC(nu)
  p%mat = C_LOC(mat0); p%kmult = C_LOC(km)
  p%c = c; p%dt = Delt; p%theta = Theta; p%comp_unit = Comp_Unit; p%cv = Cv; p%kap0 = Kap0; p%tini = Tini
  p%bc_type = BC_Type
  p%bcT = [bcT_left, bcT_bottom, bcT_right, bcT_top]
  p%opacity_model = opacity_model
  p%maxit_rte = maxit_RTE; p%maxit_mloqd = maxit_MLOQD; p%maxit_gloqd = maxit_GLOQD; p%conv_type = conv_type
  p%conv_ho = conv_ho; p%conv_lo = conv_lo; p%conv_gr1 = conv_gr1; p%conv_gr2 = conv_gr2
  p%chi = chi; p%line_src = line_src
  p%use_line_search = MERGE(1, 0, Use_Line_Search)
  p%use_safety_search = MERGE(1, 0, Use_Safety_Search)
  p%kapE_dT_flag = MERGE(1, 0, kapE_dT_flag)
  p%E_bound_low = E_Bound_Low; p%T_bound_low = T_Bound_Low
  p%threads = 1; p%precision = precision; p%tile_x = 0; p%tile_y = 0; p%lin_rtol = 0d0; p%lin_maxit = 0; p%sweep_mode = 0
  p%qd_solver = 0; p%grey_emission = 0; p%no_grey = 0
  CALL TRT_CHECK(trt_ctx_create(TRT_CTX, p), 'trt_ctx_create')
  DEALLOCATE(mat0, ox, oy, w, nu, dx, dy, km)

END SUBROUTINE TRT_GPU_INIT

SUBROUTINE TRT_GPU_FINALIZE()
  INTEGER(C_INT) :: rc
  IF (C_ASSOCIATED(TRT_CTX)) rc = trt_ctx_destroy(TRT_CTX)
  TRT_CTX = C_NULL_PTR
END SUBROUTINE TRT_GPU_FINALIZE

END MODULE TRT_CAPI
