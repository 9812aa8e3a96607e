!==================================================================================================================================!
! trt_mlqd: Fortran host program for the full-order MLQD iteration on the GPU.
!
! Mode 1 (drop-in): the time loop of TRT_MLQD_ALGORITHM, run_type mlqd (MainSrc/ALGORITHMS.f08:389-897), written against the
! reference call surface -- TRANSPORT_SCB, COLLAPSE_INTENSITIES, fg_Calc, Cg_Calc, OLD_MGQD_COEFS, OLD_GREY_COEFS, MLOQD_FV,
! COLLAPSE_MG_EF, Cbar_Calc, GREY_COEFS, EGP_FV_NEWT, MATERIAL_UPDATE and CONVERGENCE from the drop-in modules -- with the
! state in host arrays as the reference keeps it.
! Mode 2 (device-resident): the same step on the GPU through trt_step, state kept on the device between steps.
!
! Usage: trt_mlqd <problem.bin> <results.bin> [<restart_infile> [<restart_outfile> <restart_freq> [<outfile> [<flags>]]]]
!   mode 2 only: the state is read from restart_infile before the time loop ('-' for none; RESTART_IN, ALGORITHMS.f08:234-237)
!   and written to restart_outfile every restart_freq-th step (RESTART_OUT, ALGORITHMS.f08:889-894), in the reference's
!   restart file format ('-' and 0 for none); outfile: the reference's NetCDF output file, written after every step
!   (TIMESTEP_OUTS, ALGORITHMS.f08:872) with the initial state first (Init_out = 1); flags: 18 digits 0/1 in the order
!   I_out HO_Eg_out HO_Fg_out HO_E_out HO_F_out Eg_out Fg_out MGQD_E_out MGQD_F_out fg_out E_out F_out D_out old_parms_out
!   its_out conv_out kap_out Src_out (default: all 1, the reference's defaults, INPUTS.f08:1174-1191).
!   problem.bin (stream, native endianness): int32 nx ny nm ng nmat nsteps opacity_model maxit_rte maxit_mloqd maxit_gloqd
!   conv_type use_line_search kapE_dT_flag bc_type(4) precision mode; float64 c dt theta comp_unit cv kap0 tini bcT(4)
!   conv_ho conv_lo conv_gr1 conv_gr2 chi line_src; float64 Delx(nx) Dely(ny) Omega_x(nm) Omega_y(nm) quad_weight(nm)
!   nu_g(ng+1) Kappa_Mult(nmat); int32 Mat(nx,ny) (1-based).
!   results.bin: per step int32 RT_Its MGQD_Its(total) then float64 Temp, E_avg, HO_E_avg, MGQD_E_avg, fg_avg_xx,
!   fg_avg_yy, fg_avg_xy (Fortran order).
!==================================================================================================================================!
PROGRAM trt_mlqd

  USE ISO_C_BINDING
  USE TRT_CAPI
  USE TRANSPORT_SOLVES
  USE MLOQD_SOLVES
  USE GLOQD_SOLVES
  USE UPDATES
  USE CONVERGENCE_CHECKS
  IMPLICIT NONE

  INTEGER:: N_x, N_y, N_m, N_g, N_mat, N_steps, opacity_model, Maxit_RTE, Maxit_MLOQD, Maxit_GLOQD, Conv_Type
  INTEGER:: line_flag, dT_flag, BC_Type(4), precision, mode, Threads
  REAL*8:: c, Delt, Theta, Comp_Unit, cv, Kap0, Tini, bcT(4), conv_ho, conv_lo, conv_gr1, conv_gr2, chi, line_src
  REAL*8,ALLOCATABLE:: Delx(:), Dely(:), Omega_x(:), Omega_y(:), quad_weight(:), Nu_g(:), Kappa_Mult(:), A(:,:)
  INTEGER,ALLOCATABLE:: Mat(:,:)
  CHARACTER(LEN=1024):: fin, fout, restart_infile, restart_outfile, freq_arg, ncfile, flag_arg
  INTEGER:: restart_freq, out_flags(18), k
  INTEGER:: t, u, v, i, j
  INTEGER(C_INT):: its4(4)

  ! the time step's arrays, named as in TRT_MLQD_ALGORITHM (ALGORITHMS.f08:59-127)
  REAL*8,ALLOCATABLE:: I_avg(:,:,:,:), I_edgV(:,:,:,:), I_edgH(:,:,:,:), I_crn(:,:,:,:), I_crn_old(:,:,:,:)
  REAL*8,ALLOCATABLE:: Ic_edgV(:,:,:,:), Ic_edgH(:,:,:,:), RT_Src(:,:,:,:), RT_Residual(:,:,:)
  INTEGER,ALLOCATABLE:: RT_ResLoc_x(:,:,:), RT_ResLoc_y(:,:,:), MGQD_Kits(:), GQD_Kits(:)
  REAL*8,ALLOCATABLE:: Hg_avg_xx(:,:,:), Hg_avg_yy(:,:,:), Hg_avg_xy(:,:,:), Hg_edgV_xx(:,:,:), Hg_edgV_xy(:,:,:)
  REAL*8,ALLOCATABLE:: Hg_edgH_yy(:,:,:), Hg_edgH_xy(:,:,:), HO_Eg_avg(:,:,:), HO_Eg_edgV(:,:,:), HO_Eg_edgH(:,:,:)
  REAL*8,ALLOCATABLE:: HO_Fxg_edgV(:,:,:), HO_Fyg_edgH(:,:,:), HO_E_avg(:,:), HO_E_edgV(:,:), HO_E_edgH(:,:)
  REAL*8,ALLOCATABLE:: HO_Fx_edgV(:,:), HO_Fy_edgH(:,:)
  REAL*8,ALLOCATABLE:: Eg_avg(:,:,:), Eg_edgV(:,:,:), Eg_edgH(:,:,:), Fxg_edgV(:,:,:), Fyg_edgH(:,:,:)
  REAL*8,ALLOCATABLE:: Eg_avg_old(:,:,:), Eg_edgV_old(:,:,:), Eg_edgH_old(:,:,:), Fxg_edgV_old(:,:,:), Fyg_edgH_old(:,:,:)
  REAL*8,ALLOCATABLE:: fg_avg_xx(:,:,:), fg_avg_yy(:,:,:), fg_avg_xy(:,:,:), fg_edgV_xx(:,:,:), fg_edgV_xy(:,:,:)
  REAL*8,ALLOCATABLE:: fg_edgH_yy(:,:,:), fg_edgH_xy(:,:,:), fg_avg_xx_old(:,:,:), fg_avg_yy_old(:,:,:)
  REAL*8,ALLOCATABLE:: fg_edgV_xx_old(:,:,:), fg_edgV_xy_old(:,:,:), fg_edgH_yy_old(:,:,:), fg_edgH_xy_old(:,:,:)
  REAL*8,ALLOCATABLE:: Cg_L(:,:), Cg_B(:,:), Cg_R(:,:), Cg_T(:,:)
  REAL*8,ALLOCATABLE:: Eg_in_L(:,:), Eg_in_B(:,:), Eg_in_R(:,:), Eg_in_T(:,:), Fg_in_L(:,:), Fg_in_B(:,:), Fg_in_R(:,:)
  REAL*8,ALLOCATABLE:: Fg_in_T(:,:), MGQD_E_avg(:,:), MGQD_E_edgV(:,:), MGQD_E_edgH(:,:), MGQD_Fx_edgV(:,:), MGQD_Fy_edgH(:,:)
  REAL*8,ALLOCATABLE:: G_old(:,:,:), Pold_L(:,:,:), Pold_B(:,:,:), Pold_R(:,:,:), Pold_T(:,:,:)
  REAL*8,ALLOCATABLE:: MGQD_Residual(:,:,:), MGQD_BC_Residual(:,:)
  REAL*8,ALLOCATABLE:: Temp(:,:), Temp_Old(:,:), MGQD_Src(:,:,:), MGQD_Src_old(:,:,:), KapE(:,:,:), KapB(:,:,:)
  REAL*8,ALLOCATABLE:: KapR(:,:,:), KapE_old(:,:,:), KapR_old(:,:,:), Bg(:,:,:)
  REAL*8,ALLOCATABLE:: E_avg(:,:), E_edgV(:,:), E_edgH(:,:), Fx_edgV(:,:), Fy_edgH(:,:), E_avg_old(:,:)
  REAL*8,ALLOCATABLE:: KapE_Bar(:,:), KapE_Bar_old(:,:), KapE_Bar_dT(:,:), KapE_bar_MGQDold(:,:), GQD_Src(:,:)
  REAL*8,ALLOCATABLE:: GQD_Src_old(:,:), Gold_Hat(:,:), Rhat_old(:,:)
  REAL*8,ALLOCATABLE:: DC_xx(:,:), DL_xx(:,:), DR_xx(:,:), DC_yy(:,:), DB_yy(:,:), DT_yy(:,:)
  REAL*8,ALLOCATABLE:: DL_xy(:,:), DR_xy(:,:), DB_xy(:,:), DT_xy(:,:), PL(:,:), PR(:,:), PB(:,:), PT(:,:)
  REAL*8,ALLOCATABLE:: Cb_L(:), Cb_B(:), Cb_R(:), Cb_T(:), E_in_L(:), E_in_B(:), E_in_R(:), E_in_T(:)
  REAL*8,ALLOCATABLE:: F_in_L(:), F_in_B(:), F_in_R(:), F_in_T(:)
  REAL*8,ALLOCATABLE:: Temp_RTold(:,:), Temp_RTold2(:,:), Temp_MGQDold(:,:), Temp_MGQDold2(:,:)
  REAL*8,ALLOCATABLE:: HO_E_avg_RTold(:,:), HO_E_avg_RTold2(:,:), MGQD_E_avg_MGQDold(:,:), MGQD_E_avg_MGQDold2(:,:)
  REAL*8,ALLOCATABLE:: Deltas(:,:), dres(:,:)
  REAL*8:: TR_Tnorm, TR_Enorm, TR_Trho, TR_Erho, MGQD_Tnorm, MGQD_Enorm, MGQD_Trho, MGQD_Erho
  LOGICAL:: RT_Conv, MGQD_conv, Tconv, Econv
  INTEGER:: RT_Its, MGQD_Its, EGP_Its, MGQD_total
  REAL*8,PARAMETER:: Fdt_Weight = 1d0, Fin_Weight = 1d0, Ein_Weight = 1d0

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  restart_infile = ''
  restart_outfile = ''
  restart_freq = 0
  IF (COMMAND_ARGUMENT_COUNT() .GE. 3) CALL GET_COMMAND_ARGUMENT(3, restart_infile)
  IF (TRIM(restart_infile) .EQ. '-') restart_infile = ''
  IF (COMMAND_ARGUMENT_COUNT() .GE. 5) THEN
    CALL GET_COMMAND_ARGUMENT(4, restart_outfile)
    CALL GET_COMMAND_ARGUMENT(5, freq_arg)
    READ(freq_arg, *) restart_freq
  END IF
  IF (TRIM(restart_outfile) .EQ. '-') restart_outfile = ''
  ncfile = ''
  out_flags = 1
  IF (COMMAND_ARGUMENT_COUNT() .GE. 6) CALL GET_COMMAND_ARGUMENT(6, ncfile)
  IF (COMMAND_ARGUMENT_COUNT() .GE. 7) THEN
    CALL GET_COMMAND_ARGUMENT(7, flag_arg)
    DO k = 1, 18
      out_flags(k) = MERGE(1, 0, flag_arg(k:k) .EQ. '1')
    END DO
  END IF
  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) N_x, N_y, N_m, N_g, N_mat, N_steps, opacity_model, Maxit_RTE, Maxit_MLOQD, Maxit_GLOQD, Conv_Type, &
    line_flag, dT_flag, BC_Type, precision, mode
  READ(u) c, Delt, Theta, Comp_Unit, cv, Kap0, Tini, bcT, conv_ho, conv_lo, conv_gr1, conv_gr2, chi, line_src
  ALLOCATE(Delx(N_x), Dely(N_y), Omega_x(N_m), Omega_y(N_m), quad_weight(N_m), Nu_g(N_g+1), Kappa_Mult(N_mat))
  ALLOCATE(Mat(N_x,N_y), A(N_x,N_y))
  READ(u) Delx, Dely, Omega_x, Omega_y, quad_weight, Nu_g, Kappa_Mult, Mat
  CLOSE(u)
  DO j = 1, N_y
    DO i = 1, N_x
      A(i,j) = Delx(i)*Dely(j)
    END DO
  END DO
  Threads = 1

  CALL TRT_GPU_INIT(Omega_x, Omega_y, quad_weight, Nu_g, Delx, Dely, Mat, Kappa_Mult, c, Delt, Theta, Comp_Unit, cv, Kap0, &
    Tini, BC_Type, bcT(1), bcT(2), bcT(3), bcT(4), opacity_model, Maxit_RTE, Maxit_MLOQD, Maxit_GLOQD, Conv_Type, conv_ho, &
    conv_lo, conv_gr1, conv_gr2, chi, line_src, line_flag .NE. 0, .FALSE., dT_flag .NE. 0, -HUGE(1d0), -HUGE(1d0), precision)

  ALLOCATE(Temp(N_x,N_y), E_avg(N_x,N_y), HO_E_avg(N_x,N_y), MGQD_E_avg(N_x,N_y))
  ALLOCATE(fg_avg_xx(N_x,N_y,N_g), fg_avg_yy(N_x,N_y,N_g), fg_avg_xy(N_x,N_y,N_g))
  OPEN(NEWUNIT=v, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')

  IF (mode .EQ. 2) THEN
    !------------------------------------------------------------------!
    !  device-resident: one trt_step per time step                     !
    !------------------------------------------------------------------!
    IF (LEN_TRIM(restart_infile) .GT. 0) CALL TRT_GPU_RESTART_IN(restart_infile)
    IF (LEN_TRIM(ncfile) .GT. 0) THEN
      CALL TRT_GPU_OUTFILE_INIT(ncfile, 'mlqd', 'custom', 'custom', SUM(Delx), SUM(Dely), trt_time(TRT_CTX) + N_steps*Delt, &
        trt_time(TRT_CTX), 1, out_flags)
      CALL TRT_GPU_TIMESTEP_OUTS(1, 0)
    END IF
    DO t = 1, N_steps
      CALL TRT_CHECK(trt_step(TRT_CTX, its4), 'trt_step')
      IF (LEN_TRIM(ncfile) .GT. 0) CALL TRT_GPU_TIMESTEP_OUTS(t + 1, 1)
      CALL TRT_PULL_ARRAY('Temp', Temp, SIZE(Temp))
      CALL TRT_PULL_ARRAY('E_avg', E_avg, SIZE(E_avg))
      CALL TRT_PULL_ARRAY('HO_E_avg', HO_E_avg, SIZE(HO_E_avg))
      CALL TRT_PULL_ARRAY('MGQD_E_avg', MGQD_E_avg, SIZE(MGQD_E_avg))
      CALL TRT_PULL_ARRAY('fg_avg_xx', fg_avg_xx, SIZE(fg_avg_xx))
      CALL TRT_PULL_ARRAY('fg_avg_yy', fg_avg_yy, SIZE(fg_avg_yy))
      CALL TRT_PULL_ARRAY('fg_avg_xy', fg_avg_xy, SIZE(fg_avg_xy))
      WRITE(v) INT(its4(1)), INT(its4(2)), Temp, E_avg, HO_E_avg, MGQD_E_avg, fg_avg_xx, fg_avg_yy, fg_avg_xy
      IF (restart_freq .GT. 0) THEN
        IF (MOD(t, restart_freq) .EQ. 0) CALL TRT_GPU_RESTART_OUT(restart_outfile)
      END IF
    END DO
    CLOSE(v)
    CALL TRT_GPU_OUTFILE_CLOSE()
    CALL TRT_GPU_FINALIZE()
    STOP
  END IF
  IF (LEN_TRIM(restart_infile) .GT. 0 .OR. restart_freq .GT. 0) ERROR STOP 'restart files need mode 2 (device-resident state)'

  !--------------------------------------------------------------------!
  !  drop-in: host arrays and the reference call sequence              !
  !--------------------------------------------------------------------!
  ALLOCATE(I_avg(N_x,N_y,N_m,N_g), I_edgV(N_x+1,N_y,N_m,N_g), I_edgH(N_x,N_y+1,N_m,N_g))
  ALLOCATE(I_crn(2*N_x,2*N_y,N_m,N_g), I_crn_old(2*N_x,2*N_y,N_m,N_g), RT_Src(2*N_x,2*N_y,N_m,N_g))
  ALLOCATE(Ic_edgV(N_x+1,2*N_y,N_m,N_g), Ic_edgH(2*N_x,N_y+1,N_m,N_g))
  ALLOCATE(RT_Residual(4,N_g,5), RT_ResLoc_x(4,N_g,5), RT_ResLoc_y(4,N_g,5), MGQD_Kits(N_g), GQD_Kits(Maxit_GLOQD))
  ALLOCATE(Hg_avg_xx(N_x,N_y,N_g), Hg_avg_yy(N_x,N_y,N_g), Hg_avg_xy(N_x,N_y,N_g), Hg_edgV_xx(N_x+1,N_y,N_g))
  ALLOCATE(Hg_edgV_xy(N_x+1,N_y,N_g), Hg_edgH_yy(N_x,N_y+1,N_g), Hg_edgH_xy(N_x,N_y+1,N_g), HO_Eg_avg(N_x,N_y,N_g))
  ALLOCATE(HO_Eg_edgV(N_x+1,N_y,N_g), HO_Eg_edgH(N_x,N_y+1,N_g), HO_Fxg_edgV(N_x+1,N_y,N_g), HO_Fyg_edgH(N_x,N_y+1,N_g))
  ALLOCATE(HO_E_edgV(N_x+1,N_y), HO_E_edgH(N_x,N_y+1), HO_Fx_edgV(N_x+1,N_y), HO_Fy_edgH(N_x,N_y+1))
  ALLOCATE(Eg_avg(N_x,N_y,N_g), Eg_edgV(N_x+1,N_y,N_g), Eg_edgH(N_x,N_y+1,N_g), Fxg_edgV(N_x+1,N_y,N_g))
  ALLOCATE(Fyg_edgH(N_x,N_y+1,N_g), Eg_avg_old(N_x,N_y,N_g), Eg_edgV_old(N_x+1,N_y,N_g), Eg_edgH_old(N_x,N_y+1,N_g))
  ALLOCATE(Fxg_edgV_old(N_x+1,N_y,N_g), Fyg_edgH_old(N_x,N_y+1,N_g))
  ALLOCATE(fg_edgV_xx(N_x+1,N_y,N_g), fg_edgV_xy(N_x+1,N_y,N_g), fg_edgH_yy(N_x,N_y+1,N_g), fg_edgH_xy(N_x,N_y+1,N_g))
  ALLOCATE(fg_avg_xx_old(N_x,N_y,N_g), fg_avg_yy_old(N_x,N_y,N_g), fg_edgV_xx_old(N_x+1,N_y,N_g))
  ALLOCATE(fg_edgV_xy_old(N_x+1,N_y,N_g), fg_edgH_yy_old(N_x,N_y+1,N_g), fg_edgH_xy_old(N_x,N_y+1,N_g))
  ALLOCATE(Cg_L(N_y,N_g), Cg_B(N_x,N_g), Cg_R(N_y,N_g), Cg_T(N_x,N_g))
  ALLOCATE(Eg_in_L(N_y,N_g), Eg_in_B(N_x,N_g), Eg_in_R(N_y,N_g), Eg_in_T(N_x,N_g))
  ALLOCATE(Fg_in_L(N_y,N_g), Fg_in_B(N_x,N_g), Fg_in_R(N_y,N_g), Fg_in_T(N_x,N_g))
  ALLOCATE(MGQD_E_edgV(N_x+1,N_y), MGQD_E_edgH(N_x,N_y+1), MGQD_Fx_edgV(N_x+1,N_y), MGQD_Fy_edgH(N_x,N_y+1))
  ALLOCATE(G_old(N_x,N_y,N_g), Pold_L(N_x,N_y,N_g), Pold_B(N_x,N_y,N_g), Pold_R(N_x,N_y,N_g), Pold_T(N_x,N_y,N_g))
  ALLOCATE(MGQD_Residual(N_g,5,5), MGQD_BC_Residual(N_g,4))
  ALLOCATE(Temp_Old(N_x,N_y), MGQD_Src(N_x,N_y,N_g), MGQD_Src_old(N_x,N_y,N_g), KapE(N_x,N_y,N_g), KapB(N_x,N_y,N_g))
  ALLOCATE(KapR(N_x,N_y,N_g), KapE_old(N_x,N_y,N_g), KapR_old(N_x,N_y,N_g), Bg(N_x,N_y,N_g))
  ALLOCATE(E_edgV(N_x+1,N_y), E_edgH(N_x,N_y+1), Fx_edgV(N_x+1,N_y), Fy_edgH(N_x,N_y+1), E_avg_old(N_x,N_y))
  ALLOCATE(KapE_Bar(N_x,N_y), KapE_Bar_old(N_x,N_y), KapE_Bar_dT(N_x,N_y), KapE_bar_MGQDold(N_x,N_y), GQD_Src(N_x,N_y))
  ALLOCATE(GQD_Src_old(N_x,N_y), Gold_Hat(N_x,N_y), Rhat_old(N_x,N_y))
  ALLOCATE(DC_xx(N_x,N_y), DL_xx(N_x,N_y), DR_xx(N_x,N_y), DC_yy(N_x,N_y), DB_yy(N_x,N_y), DT_yy(N_x,N_y))
  ALLOCATE(DL_xy(N_x,N_y), DR_xy(N_x,N_y), DB_xy(N_x,N_y), DT_xy(N_x,N_y), PL(N_x,N_y), PR(N_x,N_y), PB(N_x,N_y), PT(N_x,N_y))
  ALLOCATE(Cb_L(N_y), Cb_B(N_x), Cb_R(N_y), Cb_T(N_x), E_in_L(N_y), E_in_B(N_x), E_in_R(N_y), E_in_T(N_x))
  ALLOCATE(F_in_L(N_y), F_in_B(N_x), F_in_R(N_y), F_in_T(N_x))
  ALLOCATE(Temp_RTold(N_x,N_y), Temp_RTold2(N_x,N_y), Temp_MGQDold(N_x,N_y), Temp_MGQDold2(N_x,N_y))
  ALLOCATE(HO_E_avg_RTold(N_x,N_y), HO_E_avg_RTold2(N_x,N_y), MGQD_E_avg_MGQDold(N_x,N_y), MGQD_E_avg_MGQDold2(N_x,N_y))

  ! the initial state that RT_INIT, MGQD_INIT, GQD_INIT and TEMP_INIT set (computed on the device by TRT_GPU_INIT)
  CALL TRT_PULL_ARRAY('I_crn', I_crn, SIZE(I_crn))
  CALL TRT_PULL_ARRAY('I_avg', I_avg, SIZE(I_avg))
  CALL TRT_PULL_ARRAY('I_edgV', I_edgV, SIZE(I_edgV))
  CALL TRT_PULL_ARRAY('I_edgH', I_edgH, SIZE(I_edgH))
  CALL TRT_PULL_ARRAY('Ic_edgV', Ic_edgV, SIZE(Ic_edgV))
  CALL TRT_PULL_ARRAY('Ic_edgH', Ic_edgH, SIZE(Ic_edgH))
  CALL TRT_PULL_ARRAY('HO_Eg_avg', HO_Eg_avg, SIZE(HO_Eg_avg))
  CALL TRT_PULL_ARRAY('HO_Eg_edgV', HO_Eg_edgV, SIZE(HO_Eg_edgV))
  CALL TRT_PULL_ARRAY('HO_Eg_edgH', HO_Eg_edgH, SIZE(HO_Eg_edgH))
  CALL TRT_PULL_ARRAY('HO_E_avg', HO_E_avg, SIZE(HO_E_avg))
  CALL TRT_PULL_ARRAY('Eg_avg', Eg_avg, SIZE(Eg_avg))
  CALL TRT_PULL_ARRAY('Eg_edgV', Eg_edgV, SIZE(Eg_edgV))
  CALL TRT_PULL_ARRAY('Eg_edgH', Eg_edgH, SIZE(Eg_edgH))
  Fxg_edgV = 0d0; Fyg_edgH = 0d0
  CALL TRT_PULL_ARRAY('fg_avg_xx', fg_avg_xx, SIZE(fg_avg_xx))
  CALL TRT_PULL_ARRAY('fg_avg_yy', fg_avg_yy, SIZE(fg_avg_yy))
  CALL TRT_PULL_ARRAY('fg_avg_xy', fg_avg_xy, SIZE(fg_avg_xy))
  CALL TRT_PULL_ARRAY('fg_edgV_xx', fg_edgV_xx, SIZE(fg_edgV_xx))
  CALL TRT_PULL_ARRAY('fg_edgV_xy', fg_edgV_xy, SIZE(fg_edgV_xy))
  CALL TRT_PULL_ARRAY('fg_edgH_yy', fg_edgH_yy, SIZE(fg_edgH_yy))
  CALL TRT_PULL_ARRAY('fg_edgH_xy', fg_edgH_xy, SIZE(fg_edgH_xy))
  CALL TRT_PULL_ARRAY('Cg_L', Cg_L, SIZE(Cg_L)); CALL TRT_PULL_ARRAY('Cg_B', Cg_B, SIZE(Cg_B))
  CALL TRT_PULL_ARRAY('Cg_R', Cg_R, SIZE(Cg_R)); CALL TRT_PULL_ARRAY('Cg_T', Cg_T, SIZE(Cg_T))
  CALL MGQD_In_Calc(Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T,Fg_in_L,Fg_in_B,Fg_in_R,Fg_in_T,I_edgV,I_edgH,Omega_x,Omega_y,&
    quad_weight,c,Comp_Unit,BC_Type,Threads)
  CALL TRT_PULL_ARRAY('E_avg', E_avg, SIZE(E_avg))
  CALL TRT_PULL_ARRAY('E_edgV', E_edgV, SIZE(E_edgV))
  CALL TRT_PULL_ARRAY('E_edgH', E_edgH, SIZE(E_edgH))
  Fx_edgV = 0d0; Fy_edgH = 0d0
  CALL TRT_PULL_ARRAY('Cb_L', Cb_L, SIZE(Cb_L)); CALL TRT_PULL_ARRAY('Cb_B', Cb_B, SIZE(Cb_B))
  CALL TRT_PULL_ARRAY('Cb_R', Cb_R, SIZE(Cb_R)); CALL TRT_PULL_ARRAY('Cb_T', Cb_T, SIZE(Cb_T))
  CALL GQD_In_Calc(E_in_L,E_in_B,E_in_R,E_in_T,F_in_L,F_in_B,F_in_R,F_in_T,Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T,&
    Fg_in_L,Fg_in_B,Fg_in_R,Fg_in_T)
  Temp = Tini
  CALL MATERIAL_UPDATE(RT_Src,MGQD_Src,KapE,KapB,KapR,Bg,Temp,Comp_Unit,Nu_g,Threads,Mat,Kappa_Mult)
  KapE_Bar = 0d0; GQD_Src = 0d0

  DO t = 1, N_steps
    ! AL:396-414
    Temp_Old = Temp
    I_crn_old = I_crn
    fg_avg_xx_old = fg_avg_xx; fg_avg_yy_old = fg_avg_yy
    fg_edgV_xx_old = fg_edgV_xx; fg_edgV_xy_old = fg_edgV_xy
    fg_edgH_yy_old = fg_edgH_yy; fg_edgH_xy_old = fg_edgH_xy
    MGQD_Src_old = MGQD_Src; GQD_Src_old = GQD_Src
    E_avg_old = E_avg; Eg_avg_old = Eg_avg; Eg_edgV_old = Eg_edgV; Eg_edgH_old = Eg_edgH
    Fxg_edgV_old = Fxg_edgV; Fyg_edgH_old = Fyg_edgH
    KapE_Bar_old = KapE_Bar; KapE_old = KapE; KapR_old = KapR
    ! AL:417-420
    CALL OLD_MGQD_COEFS(Eg_avg_old,Eg_edgV_old,Eg_edgH_old,Fxg_edgV_old,Fyg_edgH_old,fg_avg_xx_old,fg_avg_yy_old,&
      fg_edgV_xx_old,fg_edgV_xy_old,fg_edgH_yy_old,fg_edgH_xy_old,KapE_old,KapR_old,MGQD_Src_old,Delx,Dely,A,c,Delt,&
      Theta,G_old,Pold_L,Pold_B,Pold_R,Pold_T,Threads)
    CALL OLD_GREY_COEFS(c,Delt,Theta,cv,A,G_old,Temp_old,KapE_bar_old,E_avg_old,GQD_Src_old,Gold_Hat,Rhat_old)
    ! AL:518-524
    HO_E_avg_RTold2 = 0d0; HO_E_avg_RTold = 0d0; Temp_RTold2 = 0d0; Temp_RTold = 0d0
    RT_Its = 0; MGQD_total = 0; RT_Conv = .FALSE.
    DO WHILE ((.NOT. RT_Conv).AND.(RT_Its .LT. Maxit_RTE))
      RT_Its = RT_Its + 1
      IF (RT_Its .GE. 2) THEN          ! RT_start_Its = 2 for mlqd (AL:302,538)
        CALL TRANSPORT_SCB(I_avg,I_edgV,I_edgH,I_crn,Ic_edgV,Ic_edgH,Omega_x,Omega_y,Delx,Dely,A,&
          KapE,RT_Src,I_crn_old,c,Delt,Threads,RT_Residual,RT_ResLoc_x,RT_ResLoc_y,.FALSE.)
        CALL COLLAPSE_INTENSITIES(Threads,I_avg,I_edgV,I_edgH,Omega_x,Omega_y,quad_weight,Comp_Unit,Hg_avg_xx,Hg_avg_yy,&
          Hg_avg_xy,Hg_edgV_xx,Hg_edgV_xy,Hg_edgH_yy,Hg_edgH_xy,HO_Eg_edgV,HO_Eg_edgH,HO_Eg_avg,HO_Fxg_edgV,HO_Fyg_edgH,&
          HO_E_edgV,HO_E_edgH,HO_E_avg,HO_Fx_edgV,HO_Fy_edgH)
        CALL fg_Calc(fg_avg_xx,fg_avg_yy,fg_avg_xy,fg_edgV_xx,fg_edgV_xy,fg_edgH_yy,fg_edgH_xy,Hg_avg_xx,Hg_avg_yy,Hg_avg_xy,&
          Hg_edgV_xx,Hg_edgV_xy,Hg_edgH_yy,Hg_edgH_xy,HO_Eg_avg,HO_Eg_edgV,HO_Eg_edgH,c,Threads)
        CALL Cg_Calc(Cg_L,Cg_B,Cg_R,Cg_T,I_edgV,I_edgH,Omega_x,Omega_y,quad_weight,Comp_Unit,BC_Type,Threads)
      END IF
      ! AL:676-681
      MGQD_E_avg_MGQDold2 = 0d0; MGQD_E_avg_MGQDold = HO_E_avg
      Temp_MGQDold2 = 0d0; Temp_MGQDold = Temp
      MGQD_Its = 0; MGQD_conv = .FALSE.
      DO WHILE ((.NOT. MGQD_conv).AND.(MGQD_Its .LT. Maxit_MLOQD))
        MGQD_Its = MGQD_Its + 1
        CALL MLOQD_FV(Eg_avg,Eg_edgV,Eg_edgH,Fxg_edgV,Fyg_edgH,fg_avg_xx,fg_avg_yy,fg_edgV_xx,fg_edgV_xy,fg_edgH_yy,&
          fg_edgH_xy,Cg_L,Cg_B,Cg_R,Cg_T,Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T,Fg_in_L,Fg_in_B,Fg_in_R,Fg_in_T,MGQD_Src,KapE,&
          KapR,Delx,Dely,A,c,Delt,Theta,Threads,.FALSE.,MGQD_Residual,MGQD_BC_Residual,G_old,Pold_L,Pold_B,Pold_R,&
          Pold_T,Eg_avg_old,Fxg_edgV_old,Fyg_edgH_old,MGQD_Kits,Fdt_Weight,Fin_Weight,Ein_Weight)
        CALL COLLAPSE_MG_EF(Eg_avg,Eg_edgV,Eg_edgH,Fxg_edgV,Fyg_edgH,Threads,MGQD_E_avg,MGQD_E_edgV,MGQD_E_edgH,&
          MGQD_Fx_edgV,MGQD_Fy_edgH)
        CALL Cbar_Calc(Cb_L,Cb_B,Cb_R,Cb_T,Cg_L,Cg_B,Cg_R,Cg_T,Eg_avg,Eg_edgV,Eg_edgH,Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T)
        KapE_bar_MGQDold = KapE_bar
        CALL GREY_COEFS(Eg_avg,Eg_edgV,Eg_edgH,Fxg_edgV_old,Fyg_edgH_old,fg_avg_xx,fg_avg_yy,fg_edgV_xx,fg_edgV_xy,&
          fg_edgH_yy,fg_edgH_xy,KapE,KapR,A,c,Delt,Theta,Pold_L,Pold_R,Pold_B,Pold_T,KapE_Bar,DC_xx,DL_xx,DR_xx,DC_yy,&
          DB_yy,DT_yy,DL_xy,DR_xy,DB_xy,DT_xy,PL,PR,PB,PT,Fdt_Weight)
        E_avg = MGQD_E_avg; E_edgV = MGQD_E_edgV; E_edgH = MGQD_E_edgH
        Fx_edgV = MGQD_Fx_edgV; Fy_edgH = MGQD_Fy_edgH
        CALL EGP_FV_NEWT(E_avg,E_edgV,E_edgH,Temp,KapE_Bar,Fx_edgV,Fy_edgH,GQD_Src,KapE_Bar_dT,EGP_Its,Deltas,dres,&
          Temp_MGQDold2,KapE_bar_MGQDold,Theta,Delt,Delx,Dely,A,Cb_L,Cb_B,Cb_R,Cb_T,E_in_L,E_in_B,E_in_R,E_in_T,F_in_L,&
          F_in_B,F_in_R,F_in_T,DC_xx,DL_xx,DR_xx,DC_yy,DB_yy,DT_yy,DL_xy,DR_xy,DB_xy,DT_xy,PL,PR,PB,PT,Gold_Hat,&
          Rhat_old,Kap0,cv,Comp_Unit,Chi,line_src,-HUGE(1d0),-HUGE(1d0),conv_gr1,conv_gr2,Maxit_GLOQD,MGQD_Its,&
          line_flag .NE. 0,.FALSE.,.FALSE.,dT_flag .NE. 0,GQD_Kits,Fin_Weight,Ein_Weight)
        CALL MATERIAL_UPDATE(RT_Src,MGQD_Src,KapE,KapB,KapR,Bg,Temp,Comp_Unit,Nu_g,Threads,Mat,Kappa_Mult)
        Tconv = CONVERGENCE(Conv_Type,conv_lo,Temp,Temp_MGQDold,Temp_MGQDold2,MGQD_Its,MGQD_Tnorm,MGQD_Trho)
        Econv = CONVERGENCE(Conv_Type,conv_lo,MGQD_E_avg,MGQD_E_avg_MGQDold,MGQD_E_avg_MGQDold2,MGQD_Its,&
          MGQD_Enorm,MGQD_Erho)
        MGQD_conv = Tconv.AND.Econv
        MGQD_E_avg_MGQDold2 = MGQD_E_avg_MGQDold; MGQD_E_avg_MGQDold = MGQD_E_avg
        Temp_MGQDold2 = Temp_MGQDold; Temp_MGQDold = Temp
      END DO
      MGQD_total = MGQD_total + MGQD_Its
      Tconv = CONVERGENCE(Conv_Type,conv_ho,Temp,Temp_RTold,Temp_RTold2,RT_Its,TR_Tnorm,TR_Trho)
      Econv = CONVERGENCE(Conv_Type,conv_ho,HO_E_avg,HO_E_avg_RTold,HO_E_avg_RTold2,RT_Its,TR_Enorm,TR_Erho)
      RT_conv = Tconv.AND.Econv
      HO_E_avg_RTold2 = HO_E_avg_RTold; HO_E_avg_RTold = HO_E_avg
      Temp_RTold2 = Temp_RTold; Temp_RTold = Temp
    END DO
    WRITE(*,*) t*Delt, RT_Its, TR_Tnorm, TR_Enorm
    WRITE(v) RT_Its, MGQD_total, Temp, E_avg, HO_E_avg, MGQD_E_avg, fg_avg_xx, fg_avg_yy, fg_avg_xy
  END DO
  CLOSE(v)
  CALL TRT_GPU_FINALIZE()

END PROGRAM trt_mlqd
