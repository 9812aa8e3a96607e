!==================================================================================================================================!
! Drop-in modules for the hot path of the reference Fortran solver.  Each module and subroutine keeps the reference's name and
! dummy-argument list, so a driver written against MainSrc/*.f08 compiles and links against these unchanged; the body copies
! the inputs the GPU routine reads to the device-resident state, runs the HIP implementation through the C ABI and copies the
! outputs back into the caller's arrays.  TRT_GPU_INIT (module TRT_CAPI) must have created the problem first.
!
! Open_Threads arguments are accepted and ignored.  Residual diagnostics (Res_Calc = .TRUE.) are not produced by the GPU path and
! stop the run, as the reference stops on an unsupported request.
!==================================================================================================================================!
MODULE TRANSPORT_SOLVES

  USE TRT_CAPI
  IMPLICIT NONE

CONTAINS

! TRANSPORT_SCB, MainSrc/TRANSPORT_SOLVES.f08:14-327
SUBROUTINE TRANSPORT_SCB(I_avg,I_edgV,I_edgH,I_crn,Ic_edgV,Ic_edgH,Omega_x,Omega_y,Delx,Dely,A,KapE_in,Src_in,&
  I_crn_old,c,Delt,Open_Threads,RT_Residual,RT_ResLoc_x,RT_ResLoc_y,Res_Calc)
  REAL*8,INTENT(IN):: Omega_x(:), Omega_y(:)
  REAL*8,INTENT(IN):: Delx(:), Dely(:), A(:,:)
  REAL*8,INTENT(IN):: KapE_in(:,:,:), Src_in(:,:,:,:)
  REAL*8,INTENT(IN):: I_crn_old(:,:,:,:)
  REAL*8,INTENT(IN):: c, Delt
  INTEGER,INTENT(IN):: Open_Threads
  LOGICAL,INTENT(IN):: Res_Calc
  REAL*8,INTENT(OUT):: I_avg(:,:,:,:), I_edgV(:,:,:,:), I_edgH(:,:,:,:)
  REAL*8,INTENT(OUT):: I_crn(:,:,:,:), Ic_edgV(:,:,:,:), Ic_edgH(:,:,:,:)
  REAL*8,INTENT(OUT):: RT_Residual(:,:,:)
  INTEGER,INTENT(OUT):: RT_ResLoc_x(:,:,:), RT_ResLoc_y(:,:,:)
  REAL*8,ALLOCATABLE:: src(:,:,:)
  INTEGER:: i, j, g

  IF (Res_Calc) STOP 'TRANSPORT_SCB: Res_Calc is not available on the GPU path'
  ! the reference source is corner- and angle-uniform (UPDATES.f08:141-150); the device keeps one value per cell and group
  ALLOCATE(src(SIZE(KapE_in,1), SIZE(KapE_in,2), SIZE(KapE_in,3)))
  DO g = 1, SIZE(src,3)
    DO j = 1, SIZE(src,2)
      DO i = 1, SIZE(src,1)
        src(i,j,g) = Src_in(2*i-1, 2*j-1, 1, g)
      END DO
    END DO
  END DO
  CALL TRT_PUSH_ARRAY('KapE', KapE_in, SIZE(KapE_in))
  CALL TRT_PUSH_ARRAY('RT_Src', src, SIZE(src))
  CALL TRT_PUSH_ARRAY('I_crn_old', I_crn_old, SIZE(I_crn_old))
  CALL TRT_CHECK(trt_transport_scb(TRT_CTX), 'TRANSPORT_SCB')
  CALL TRT_PULL_ARRAY('I_crn', I_crn, SIZE(I_crn))
  CALL TRT_PULL_ARRAY('I_avg', I_avg, SIZE(I_avg))
  CALL TRT_PULL_ARRAY('I_edgV', I_edgV, SIZE(I_edgV))
  CALL TRT_PULL_ARRAY('I_edgH', I_edgH, SIZE(I_edgH))
  CALL TRT_PULL_ARRAY('Ic_edgV', Ic_edgV, SIZE(Ic_edgV))
  CALL TRT_PULL_ARRAY('Ic_edgH', Ic_edgH, SIZE(Ic_edgH))
  DEALLOCATE(src)
END SUBROUTINE TRANSPORT_SCB

! COLLAPSE_INTENSITIES, MainSrc/TRANSPORT_SOLVES.f08:381-511.  The angular sums were accumulated by the preceding
! TRANSPORT_SCB on the device; the intensity arguments are those the preceding call returned.
SUBROUTINE COLLAPSE_INTENSITIES(Open_Threads,I_avg,I_edgV,I_edgH,Omega_x,Omega_y,quad_weight,Comp_Unit,Hg_avg_xx,&
  Hg_avg_yy,Hg_avg_xy,Hg_edgV_xx,Hg_edgV_xy,Hg_edgH_yy,Hg_edgH_xy,Eg_edgV,Eg_edgH,Eg_avg,Fxg_edgV,Fyg_edgH,E_edgV,&
  E_edgH,E_avg,Fx_edgV,Fy_edgH)
  REAL*8,INTENT(IN):: I_avg(:,:,:,:)
  REAL*8,INTENT(IN):: I_edgV(:,:,:,:), I_edgH(:,:,:,:)
  REAL*8,INTENT(IN):: Omega_x(:), Omega_y(:), quad_weight(:)
  REAL*8,INTENT(IN):: Comp_Unit
  INTEGER,INTENT(IN):: Open_Threads
  REAL*8,INTENT(OUT):: Hg_avg_xx(:,:,:), Hg_avg_yy(:,:,:), Hg_avg_xy(:,:,:)
  REAL*8,INTENT(OUT):: Hg_edgV_xx(:,:,:), Hg_edgV_xy(:,:,:)
  REAL*8,INTENT(OUT):: Hg_edgH_yy(:,:,:), Hg_edgH_xy(:,:,:)
  REAL*8,INTENT(OUT):: Eg_avg(:,:,:), Eg_edgV(:,:,:), Eg_edgH(:,:,:)
  REAL*8,INTENT(OUT):: Fxg_edgV(:,:,:), Fyg_edgH(:,:,:)
  REAL*8,INTENT(OUT):: E_avg(:,:), E_edgV(:,:), E_edgH(:,:)
  REAL*8,INTENT(OUT):: Fx_edgV(:,:), Fy_edgH(:,:)

  CALL TRT_CHECK(trt_collapse_intensities(TRT_CTX), 'COLLAPSE_INTENSITIES')
  CALL TRT_PULL_ARRAY('Hg_avg_xx', Hg_avg_xx, SIZE(Hg_avg_xx))
  CALL TRT_PULL_ARRAY('Hg_avg_yy', Hg_avg_yy, SIZE(Hg_avg_yy))
  CALL TRT_PULL_ARRAY('Hg_avg_xy', Hg_avg_xy, SIZE(Hg_avg_xy))
  CALL TRT_PULL_ARRAY('Hg_edgV_xx', Hg_edgV_xx, SIZE(Hg_edgV_xx))
  CALL TRT_PULL_ARRAY('Hg_edgV_xy', Hg_edgV_xy, SIZE(Hg_edgV_xy))
  CALL TRT_PULL_ARRAY('Hg_edgH_yy', Hg_edgH_yy, SIZE(Hg_edgH_yy))
  CALL TRT_PULL_ARRAY('Hg_edgH_xy', Hg_edgH_xy, SIZE(Hg_edgH_xy))
  CALL TRT_PULL_ARRAY('HO_Eg_avg', Eg_avg, SIZE(Eg_avg))
  CALL TRT_PULL_ARRAY('HO_Eg_edgV', Eg_edgV, SIZE(Eg_edgV))
  CALL TRT_PULL_ARRAY('HO_Eg_edgH', Eg_edgH, SIZE(Eg_edgH))
  CALL TRT_PULL_ARRAY('HO_Fxg_edgV', Fxg_edgV, SIZE(Fxg_edgV))
  CALL TRT_PULL_ARRAY('HO_Fyg_edgH', Fyg_edgH, SIZE(Fyg_edgH))
  CALL TRT_PULL_ARRAY('HO_E_avg', E_avg, SIZE(E_avg))
  CALL TRT_PULL_ARRAY('HO_E_edgV', E_edgV, SIZE(E_edgV))
  CALL TRT_PULL_ARRAY('HO_E_edgH', E_edgH, SIZE(E_edgH))
  CALL TRT_PULL_ARRAY('HO_Fx_edgV', Fx_edgV, SIZE(Fx_edgV))
  CALL TRT_PULL_ARRAY('HO_Fy_edgH', Fy_edgH, SIZE(Fy_edgH))
END SUBROUTINE COLLAPSE_INTENSITIES

END MODULE TRANSPORT_SOLVES

!==================================================================================================================================!
MODULE MLOQD_SOLVES

  USE TRT_CAPI
  IMPLICIT NONE

CONTAINS

! OLD_MGQD_COEFS, MainSrc/MLOQD_SOLVES.f08:12-73
SUBROUTINE OLD_MGQD_COEFS(Eg_avg_old,Eg_edgV_old,Eg_edgH_old,Fxg_edgV_old,Fyg_edgH_old,fg_avg_xx_old,fg_avg_yy_old,&
  fg_edgV_xx_old,fg_edgV_xy_old,fg_edgH_yy_old,fg_edgH_xy_old,KapE_old,KapR_old,Src_old,Delx,Dely,A,c,Delt,Theta,&
  G_old,Pold_L,Pold_B,Pold_R,Pold_T,Open_Threads)
  REAL*8,INTENT(IN):: Eg_avg_old(:,:,:), Eg_edgV_old(:,:,:), Eg_edgH_old(:,:,:)
  REAL*8,INTENT(IN):: Fxg_edgV_old(:,:,:), Fyg_edgH_old(:,:,:)
  REAL*8,INTENT(IN):: fg_avg_xx_old(:,:,:), fg_avg_yy_old(:,:,:)
  REAL*8,INTENT(IN):: fg_edgV_xx_old(:,:,:), fg_edgV_xy_old(:,:,:)
  REAL*8,INTENT(IN):: fg_edgH_yy_old(:,:,:), fg_edgH_xy_old(:,:,:)
  REAL*8,INTENT(IN):: KapE_old(:,:,:), KapR_old(:,:,:), Src_old(:,:,:)
  REAL*8,INTENT(IN):: Delx(:), Dely(:), A(:,:)
  REAL*8,INTENT(IN):: c, Delt, Theta
  INTEGER,INTENT(IN):: Open_Threads
  REAL*8,INTENT(OUT):: G_old(:,:,:), Pold_L(:,:,:), Pold_B(:,:,:), Pold_R(:,:,:), Pold_T(:,:,:)

  CALL TRT_PUSH_ARRAY('Eg_avg_old', Eg_avg_old, SIZE(Eg_avg_old))
  CALL TRT_PUSH_ARRAY('Eg_edgV_old', Eg_edgV_old, SIZE(Eg_edgV_old))
  CALL TRT_PUSH_ARRAY('Eg_edgH_old', Eg_edgH_old, SIZE(Eg_edgH_old))
  CALL TRT_PUSH_ARRAY('Fxg_edgV_old', Fxg_edgV_old, SIZE(Fxg_edgV_old))
  CALL TRT_PUSH_ARRAY('Fyg_edgH_old', Fyg_edgH_old, SIZE(Fyg_edgH_old))
  CALL TRT_PUSH_ARRAY('fg_avg_xx_old', fg_avg_xx_old, SIZE(fg_avg_xx_old))
  CALL TRT_PUSH_ARRAY('fg_avg_yy_old', fg_avg_yy_old, SIZE(fg_avg_yy_old))
  CALL TRT_PUSH_ARRAY('fg_edgV_xx_old', fg_edgV_xx_old, SIZE(fg_edgV_xx_old))
  CALL TRT_PUSH_ARRAY('fg_edgV_xy_old', fg_edgV_xy_old, SIZE(fg_edgV_xy_old))
  CALL TRT_PUSH_ARRAY('fg_edgH_yy_old', fg_edgH_yy_old, SIZE(fg_edgH_yy_old))
  CALL TRT_PUSH_ARRAY('fg_edgH_xy_old', fg_edgH_xy_old, SIZE(fg_edgH_xy_old))
  CALL TRT_PUSH_ARRAY('KapE_old', KapE_old, SIZE(KapE_old))
  CALL TRT_PUSH_ARRAY('KapR_old', KapR_old, SIZE(KapR_old))
  CALL TRT_PUSH_ARRAY('MGQD_Src_old', Src_old, SIZE(Src_old))
  CALL TRT_CHECK(trt_old_mgqd_coefs(TRT_CTX), 'OLD_MGQD_COEFS')
  CALL TRT_PULL_ARRAY('G_old', G_old, SIZE(G_old))
  CALL TRT_PULL_ARRAY('Pold_L', Pold_L, SIZE(Pold_L))
  CALL TRT_PULL_ARRAY('Pold_B', Pold_B, SIZE(Pold_B))
  CALL TRT_PULL_ARRAY('Pold_R', Pold_R, SIZE(Pold_R))
  CALL TRT_PULL_ARRAY('Pold_T', Pold_T, SIZE(Pold_T))
END SUBROUTINE OLD_MGQD_COEFS

! MLOQD_FV, MainSrc/MLOQD_SOLVES.f08:78-362 (run_type mlqd: Fdt/Fin/Ein weights of 1)
SUBROUTINE MLOQD_FV(Eg_avg,Eg_edgV,Eg_edgH,Fxg_edgV,Fyg_edgH,fg_avg_xx,fg_avg_yy,fg_edgV_xx,fg_edgV_xy,fg_edgH_yy,&
  fg_edgH_xy,Cg_L,Cg_B,Cg_R,Cg_T,Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T,Fg_in_L,Fg_in_B,Fg_in_R,Fg_in_T,Src,KapE,KapR,Delx,&
  Dely,A,c,Delt,Theta,Open_Threads,Res_Calc,MGQD_Residual,MGQD_BC_Residual,G_old,Pold_L,Pold_B,Pold_R,Pold_T,&
  Eg_avg_old,Fxg_edgV_old,Fyg_edgH_old,MGQD_Kits,Fdt_Weight,Fin_Weight,Ein_Weight)
  REAL*8,INTENT(INOUT):: Eg_avg(:,:,:), Eg_edgV(:,:,:), Eg_edgH(:,:,:)
  REAL*8,INTENT(OUT):: Fxg_edgV(:,:,:), Fyg_edgH(:,:,:)
  REAL*8,INTENT(OUT):: MGQD_Residual(:,:,:), MGQD_BC_Residual(:,:)
  INTEGER,INTENT(OUT):: MGQD_Kits(*)
  REAL*8,INTENT(IN):: G_old(:,:,:), Pold_L(:,:,:), Pold_B(:,:,:), Pold_R(:,:,:), Pold_T(:,:,:)
  REAL*8,INTENT(IN):: fg_avg_xx(:,:,:), fg_avg_yy(:,:,:)
  REAL*8,INTENT(IN):: fg_edgV_xx(:,:,:), fg_edgV_xy(:,:,:)
  REAL*8,INTENT(IN):: fg_edgH_yy(:,:,:), fg_edgH_xy(:,:,:)
  REAL*8,INTENT(IN):: Cg_L(:,:), Cg_B(:,:), Cg_R(:,:), Cg_T(:,:)
  REAL*8,INTENT(IN):: Eg_in_L(:,:), Eg_in_B(:,:), Eg_in_R(:,:), Eg_in_T(:,:)
  REAL*8,INTENT(IN):: Fg_in_L(:,:), Fg_in_B(:,:), Fg_in_R(:,:), Fg_in_T(:,:)
  REAL*8,INTENT(IN):: Src(:,:,:)
  REAL*8,INTENT(IN):: Eg_avg_old(:,:,:)
  REAL*8,INTENT(IN):: Fxg_edgV_old(:,:,:), Fyg_edgH_old(:,:,:)
  REAL*8,INTENT(IN):: KapE(:,:,:), KapR(:,:,:)
  REAL*8,INTENT(IN):: Delx(:), Dely(:), A(:,:)
  REAL*8,INTENT(IN):: c, Delt
  REAL*8,INTENT(IN):: Theta
  REAL*8,INTENT(IN):: Fdt_Weight, Fin_Weight, Ein_Weight
  INTEGER,INTENT(IN):: Open_Threads
  LOGICAL,INTENT(IN):: Res_Calc
  INTEGER(C_INT),ALLOCATABLE:: kits(:)

  IF (Res_Calc) STOP 'MLOQD_FV: Res_Calc is not available on the GPU path'
  IF (Fdt_Weight .NE. 1d0 .OR. Fin_Weight .NE. 1d0 .OR. Ein_Weight .NE. 1d0) &
    STOP 'MLOQD_FV: the GPU path implements run_type mlqd (unit Fdt/Fin/Ein weights)'
  CALL TRT_PUSH_ARRAY('Eg_avg', Eg_avg, SIZE(Eg_avg))
  CALL TRT_PUSH_ARRAY('Eg_edgV', Eg_edgV, SIZE(Eg_edgV))
  CALL TRT_PUSH_ARRAY('Eg_edgH', Eg_edgH, SIZE(Eg_edgH))
  CALL TRT_PUSH_ARRAY('fg_avg_xx', fg_avg_xx, SIZE(fg_avg_xx))
  CALL TRT_PUSH_ARRAY('fg_avg_yy', fg_avg_yy, SIZE(fg_avg_yy))
  CALL TRT_PUSH_ARRAY('fg_edgV_xx', fg_edgV_xx, SIZE(fg_edgV_xx))
  CALL TRT_PUSH_ARRAY('fg_edgV_xy', fg_edgV_xy, SIZE(fg_edgV_xy))
  CALL TRT_PUSH_ARRAY('fg_edgH_yy', fg_edgH_yy, SIZE(fg_edgH_yy))
  CALL TRT_PUSH_ARRAY('fg_edgH_xy', fg_edgH_xy, SIZE(fg_edgH_xy))
  CALL TRT_PUSH_ARRAY('Cg_L', Cg_L, SIZE(Cg_L))
  CALL TRT_PUSH_ARRAY('Cg_B', Cg_B, SIZE(Cg_B))
  CALL TRT_PUSH_ARRAY('Cg_R', Cg_R, SIZE(Cg_R))
  CALL TRT_PUSH_ARRAY('Cg_T', Cg_T, SIZE(Cg_T))
  CALL TRT_PUSH_ARRAY('Eg_in_L', Eg_in_L, SIZE(Eg_in_L))
  CALL TRT_PUSH_ARRAY('Eg_in_B', Eg_in_B, SIZE(Eg_in_B))
  CALL TRT_PUSH_ARRAY('Eg_in_R', Eg_in_R, SIZE(Eg_in_R))
  CALL TRT_PUSH_ARRAY('Eg_in_T', Eg_in_T, SIZE(Eg_in_T))
  CALL TRT_PUSH_ARRAY('Fg_in_L', Fg_in_L, SIZE(Fg_in_L))
  CALL TRT_PUSH_ARRAY('Fg_in_B', Fg_in_B, SIZE(Fg_in_B))
  CALL TRT_PUSH_ARRAY('Fg_in_R', Fg_in_R, SIZE(Fg_in_R))
  CALL TRT_PUSH_ARRAY('Fg_in_T', Fg_in_T, SIZE(Fg_in_T))
  CALL TRT_PUSH_ARRAY('MGQD_Src', Src, SIZE(Src))
  CALL TRT_PUSH_ARRAY('KapE', KapE, SIZE(KapE))
  CALL TRT_PUSH_ARRAY('KapR', KapR, SIZE(KapR))
  CALL TRT_PUSH_ARRAY('G_old', G_old, SIZE(G_old))
  CALL TRT_PUSH_ARRAY('Pold_L', Pold_L, SIZE(Pold_L))
  CALL TRT_PUSH_ARRAY('Pold_B', Pold_B, SIZE(Pold_B))
  CALL TRT_PUSH_ARRAY('Pold_R', Pold_R, SIZE(Pold_R))
  CALL TRT_PUSH_ARRAY('Pold_T', Pold_T, SIZE(Pold_T))
  CALL TRT_PUSH_ARRAY('Eg_avg_old', Eg_avg_old, SIZE(Eg_avg_old))
  CALL TRT_PUSH_ARRAY('Fxg_edgV_old', Fxg_edgV_old, SIZE(Fxg_edgV_old))
  CALL TRT_PUSH_ARRAY('Fyg_edgH_old', Fyg_edgH_old, SIZE(Fyg_edgH_old))
  ALLOCATE(kits(SIZE(Eg_avg,3)))
  CALL TRT_CHECK(trt_mloqd_fv(TRT_CTX, kits), 'MLOQD_FV')
  MGQD_Kits(1:SIZE(kits)) = kits
  CALL TRT_PULL_ARRAY('Eg_avg', Eg_avg, SIZE(Eg_avg))
  CALL TRT_PULL_ARRAY('Eg_edgV', Eg_edgV, SIZE(Eg_edgV))
  CALL TRT_PULL_ARRAY('Eg_edgH', Eg_edgH, SIZE(Eg_edgH))
  CALL TRT_PULL_ARRAY('Fxg_edgV', Fxg_edgV, SIZE(Fxg_edgV))
  CALL TRT_PULL_ARRAY('Fyg_edgH', Fyg_edgH, SIZE(Fyg_edgH))
  DEALLOCATE(kits)
END SUBROUTINE MLOQD_FV

! Cg_Calc, MainSrc/MLOQD_SOLVES.f08:367-488: from the boundary moments of the preceding TRANSPORT_SCB
SUBROUTINE Cg_Calc(Cg_L,Cg_B,Cg_R,Cg_T,I_edgV,I_edgH,Omega_x,Omega_y,quad_weight,Comp_Unit,BC_Type,Open_Threads)
  REAL*8,INTENT(OUT):: Cg_L(:,:), Cg_B(:,:), Cg_R(:,:), Cg_T(:,:)
  REAL*8,INTENT(IN):: I_edgV(:,:,:,:), I_edgH(:,:,:,:)
  REAL*8,INTENT(IN):: Omega_x(:), Omega_y(:), quad_weight(:)
  REAL*8,INTENT(IN):: Comp_Unit
  INTEGER,INTENT(IN):: BC_Type(:), Open_Threads

  CALL TRT_CHECK(trt_cg_calc(TRT_CTX), 'Cg_Calc')
  CALL TRT_PULL_ARRAY('Cg_L', Cg_L, SIZE(Cg_L))
  CALL TRT_PULL_ARRAY('Cg_B', Cg_B, SIZE(Cg_B))
  CALL TRT_PULL_ARRAY('Cg_R', Cg_R, SIZE(Cg_R))
  CALL TRT_PULL_ARRAY('Cg_T', Cg_T, SIZE(Cg_T))
END SUBROUTINE Cg_Calc

! MGQD_In_Calc, MainSrc/MLOQD_SOLVES.f08:493-614 (zero on reflective sides)
SUBROUTINE MGQD_In_Calc(Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T,Fg_in_L,Fg_in_B,Fg_in_R,Fg_in_T,I_edgV,I_edgH,Omega_x,Omega_y,&
  quad_weight,c,Comp_Unit,BC_Type,Open_Threads)
  REAL*8,INTENT(OUT):: Eg_in_L(:,:), Eg_in_B(:,:), Eg_in_R(:,:), Eg_in_T(:,:)
  REAL*8,INTENT(OUT):: Fg_in_L(:,:), Fg_in_B(:,:), Fg_in_R(:,:), Fg_in_T(:,:)
  REAL*8,INTENT(IN):: I_edgV(:,:,:,:), I_edgH(:,:,:,:)
  REAL*8,INTENT(IN):: Omega_x(:), Omega_y(:), quad_weight(:)
  REAL*8,INTENT(IN):: c, Comp_Unit
  INTEGER,INTENT(IN):: BC_Type(:), Open_Threads

  CALL TRT_CHECK(trt_mgqd_in_calc(TRT_CTX), 'MGQD_In_Calc')
  CALL TRT_PULL_ARRAY('Eg_in_L', Eg_in_L, SIZE(Eg_in_L))
  CALL TRT_PULL_ARRAY('Eg_in_B', Eg_in_B, SIZE(Eg_in_B))
  CALL TRT_PULL_ARRAY('Eg_in_R', Eg_in_R, SIZE(Eg_in_R))
  CALL TRT_PULL_ARRAY('Eg_in_T', Eg_in_T, SIZE(Eg_in_T))
  CALL TRT_PULL_ARRAY('Fg_in_L', Fg_in_L, SIZE(Fg_in_L))
  CALL TRT_PULL_ARRAY('Fg_in_B', Fg_in_B, SIZE(Fg_in_B))
  CALL TRT_PULL_ARRAY('Fg_in_R', Fg_in_R, SIZE(Fg_in_R))
  CALL TRT_PULL_ARRAY('Fg_in_T', Fg_in_T, SIZE(Fg_in_T))
END SUBROUTINE MGQD_In_Calc

! fg_Calc, MainSrc/MLOQD_SOLVES.f08:619-678
SUBROUTINE fg_Calc(fg_avg_xx,fg_avg_yy,fg_avg_xy,fg_edgV_xx,fg_edgV_xy,fg_edgH_yy,fg_edgH_xy,Hg_avg_xx,Hg_avg_yy,Hg_avg_xy,&
  Hg_edgV_xx,Hg_edgV_xy,Hg_edgH_yy,Hg_edgH_xy,HO_Eg_avg,HO_Eg_edgV,HO_Eg_edgH,c,Open_Threads)
  REAL*8,INTENT(OUT):: fg_avg_xx(:,:,:), fg_avg_yy(:,:,:), fg_avg_xy(:,:,:)
  REAL*8,INTENT(OUT):: fg_edgV_xx(:,:,:), fg_edgV_xy(:,:,:)
  REAL*8,INTENT(OUT):: fg_edgH_yy(:,:,:), fg_edgH_xy(:,:,:)
  REAL*8,INTENT(IN):: Hg_avg_xx(:,:,:), Hg_avg_yy(:,:,:), Hg_avg_xy(:,:,:)
  REAL*8,INTENT(IN):: Hg_edgV_xx(:,:,:), Hg_edgV_xy(:,:,:)
  REAL*8,INTENT(IN):: Hg_edgH_yy(:,:,:), Hg_edgH_xy(:,:,:)
  REAL*8,INTENT(IN):: HO_Eg_avg(:,:,:), HO_Eg_edgV(:,:,:), HO_Eg_edgH(:,:,:)
  REAL*8,INTENT(IN):: c
  INTEGER,INTENT(IN):: Open_Threads

  CALL TRT_PUSH_ARRAY('Hg_avg_xx', Hg_avg_xx, SIZE(Hg_avg_xx))
  CALL TRT_PUSH_ARRAY('Hg_avg_yy', Hg_avg_yy, SIZE(Hg_avg_yy))
  CALL TRT_PUSH_ARRAY('Hg_avg_xy', Hg_avg_xy, SIZE(Hg_avg_xy))
  CALL TRT_PUSH_ARRAY('Hg_edgV_xx', Hg_edgV_xx, SIZE(Hg_edgV_xx))
  CALL TRT_PUSH_ARRAY('Hg_edgV_xy', Hg_edgV_xy, SIZE(Hg_edgV_xy))
  CALL TRT_PUSH_ARRAY('Hg_edgH_yy', Hg_edgH_yy, SIZE(Hg_edgH_yy))
  CALL TRT_PUSH_ARRAY('Hg_edgH_xy', Hg_edgH_xy, SIZE(Hg_edgH_xy))
  CALL TRT_PUSH_ARRAY('HO_Eg_avg', HO_Eg_avg, SIZE(HO_Eg_avg))
  CALL TRT_PUSH_ARRAY('HO_Eg_edgV', HO_Eg_edgV, SIZE(HO_Eg_edgV))
  CALL TRT_PUSH_ARRAY('HO_Eg_edgH', HO_Eg_edgH, SIZE(HO_Eg_edgH))
  CALL TRT_CHECK(trt_fg_calc(TRT_CTX), 'fg_Calc')
  CALL TRT_PULL_ARRAY('fg_avg_xx', fg_avg_xx, SIZE(fg_avg_xx))
  CALL TRT_PULL_ARRAY('fg_avg_yy', fg_avg_yy, SIZE(fg_avg_yy))
  CALL TRT_PULL_ARRAY('fg_avg_xy', fg_avg_xy, SIZE(fg_avg_xy))
  CALL TRT_PULL_ARRAY('fg_edgV_xx', fg_edgV_xx, SIZE(fg_edgV_xx))
  CALL TRT_PULL_ARRAY('fg_edgV_xy', fg_edgV_xy, SIZE(fg_edgV_xy))
  CALL TRT_PULL_ARRAY('fg_edgH_yy', fg_edgH_yy, SIZE(fg_edgH_yy))
  CALL TRT_PULL_ARRAY('fg_edgH_xy', fg_edgH_xy, SIZE(fg_edgH_xy))
END SUBROUTINE fg_Calc

! COLLAPSE_MG_EF, MainSrc/MLOQD_SOLVES.f08:687-730
SUBROUTINE COLLAPSE_MG_EF(Eg_avg,Eg_edgV,Eg_edgH,Fxg_edgV,Fyg_edgH,Open_Threads,E_avg,E_edgV,E_edgH,Fx_edgV,Fy_edgH)
  REAL*8,INTENT(IN):: Eg_avg(:,:,:), Eg_edgV(:,:,:), Eg_edgH(:,:,:)
  REAL*8,INTENT(IN):: Fxg_edgV(:,:,:), Fyg_edgH(:,:,:)
  INTEGER,INTENT(IN):: Open_Threads
  REAL*8,INTENT(OUT):: E_avg(:,:), E_edgV(:,:), E_edgH(:,:)
  REAL*8,INTENT(OUT):: Fx_edgV(:,:), Fy_edgH(:,:)

  CALL TRT_PUSH_ARRAY('Eg_avg', Eg_avg, SIZE(Eg_avg))
  CALL TRT_PUSH_ARRAY('Eg_edgV', Eg_edgV, SIZE(Eg_edgV))
  CALL TRT_PUSH_ARRAY('Eg_edgH', Eg_edgH, SIZE(Eg_edgH))
  CALL TRT_PUSH_ARRAY('Fxg_edgV', Fxg_edgV, SIZE(Fxg_edgV))
  CALL TRT_PUSH_ARRAY('Fyg_edgH', Fyg_edgH, SIZE(Fyg_edgH))
  CALL TRT_CHECK(trt_collapse_mg_ef(TRT_CTX), 'COLLAPSE_MG_EF')
  CALL TRT_PULL_ARRAY('MGQD_E_avg', E_avg, SIZE(E_avg))
  CALL TRT_PULL_ARRAY('MGQD_E_edgV', E_edgV, SIZE(E_edgV))
  CALL TRT_PULL_ARRAY('MGQD_E_edgH', E_edgH, SIZE(E_edgH))
  CALL TRT_PULL_ARRAY('MGQD_Fx_edgV', Fx_edgV, SIZE(Fx_edgV))
  CALL TRT_PULL_ARRAY('MGQD_Fy_edgH', Fy_edgH, SIZE(Fy_edgH))
END SUBROUTINE COLLAPSE_MG_EF

END MODULE MLOQD_SOLVES

!==================================================================================================================================!
MODULE GLOQD_SOLVES

  USE TRT_CAPI
  IMPLICIT NONE

CONTAINS

! OLD_GREY_COEFS, MainSrc/GLOQD_SOLVES.f08:12-42
SUBROUTINE OLD_GREY_COEFS(c,Delt,Theta,cv,A,Gold,Temp_old,KapE_bar_old,E_avg_old,GQD_Src_old,Gold_Hat,Rhat_old)
  REAL*8,INTENT(IN):: c, Delt, Theta, cv, A(:,:)
  REAL*8,INTENT(IN):: Gold(:,:,:), Temp_old(:,:), KapE_bar_old(:,:), E_avg_old(:,:), GQD_Src_old(:,:)
  REAL*8,INTENT(OUT):: Gold_Hat(:,:), Rhat_old(:,:)

  CALL TRT_PUSH_ARRAY('G_old', Gold, SIZE(Gold))
  CALL TRT_PUSH_ARRAY('Temp_Old', Temp_old, SIZE(Temp_old))
  CALL TRT_PUSH_ARRAY('KapE_Bar_old', KapE_bar_old, SIZE(KapE_bar_old))
  CALL TRT_PUSH_ARRAY('E_avg_old', E_avg_old, SIZE(E_avg_old))
  CALL TRT_PUSH_ARRAY('GQD_Src_old', GQD_Src_old, SIZE(GQD_Src_old))
  CALL TRT_CHECK(trt_old_grey_coefs(TRT_CTX), 'OLD_GREY_COEFS')
  CALL TRT_PULL_ARRAY('Gold_Hat', Gold_Hat, SIZE(Gold_Hat))
  CALL TRT_PULL_ARRAY('Rhat_old', Rhat_old, SIZE(Rhat_old))
END SUBROUTINE OLD_GREY_COEFS

! GREY_COEFS, MainSrc/GLOQD_SOLVES.f08:47-146 (run_type mlqd: Fdt_Weight = 1)
SUBROUTINE GREY_COEFS(Eg_avg,Eg_edgV,Eg_edgH,Fxg_edgV_old,Fyg_edgH_old,fg_avg_xx,fg_avg_yy,fg_edgV_xx,fg_edgV_xy,fg_edgH_yy,&
  fg_edgH_xy,KapE,KapR,A,c,Delt,Theta,Pold_L,Pold_R,Pold_B,Pold_T,KapE_Bar,DC_xx,DL_xx,DR_xx,DC_yy,DB_yy,DT_yy,DL_xy,&
  DR_xy,DB_xy,DT_xy,PL,PR,PB,PT,Fdt_Weight)
  REAL*8,INTENT(IN):: Eg_avg(:,:,:), Eg_edgV(:,:,:), Eg_edgH(:,:,:)
  REAL*8,INTENT(IN):: Fxg_edgV_old(:,:,:), Fyg_edgH_old(:,:,:)
  REAL*8,INTENT(IN):: fg_avg_xx(:,:,:), fg_avg_yy(:,:,:)
  REAL*8,INTENT(IN):: fg_edgV_xx(:,:,:), fg_edgV_xy(:,:,:)
  REAL*8,INTENT(IN):: fg_edgH_yy(:,:,:), fg_edgH_xy(:,:,:)
  REAL*8,INTENT(IN):: KapE(:,:,:), KapR(:,:,:)
  REAL*8,INTENT(IN):: A(:,:)
  REAL*8,INTENT(IN):: c, Delt, Theta, Fdt_Weight
  REAL*8,INTENT(IN):: Pold_L(:,:,:), Pold_R(:,:,:), Pold_B(:,:,:), Pold_T(:,:,:)
  REAL*8,INTENT(OUT):: KapE_Bar(:,:)
  REAL*8,INTENT(OUT):: DC_xx(:,:), DL_xx(:,:), DR_xx(:,:)
  REAL*8,INTENT(OUT):: DC_yy(:,:), DB_yy(:,:), DT_yy(:,:)
  REAL*8,INTENT(OUT):: DL_xy(:,:), DR_xy(:,:), DB_xy(:,:), DT_xy(:,:)
  REAL*8,INTENT(OUT):: PL(:,:), PR(:,:), PB(:,:), PT(:,:)

  IF (Fdt_Weight .NE. 1d0) STOP 'GREY_COEFS: the GPU path implements run_type mlqd (Fdt_Weight = 1)'
  CALL TRT_PUSH_ARRAY('Eg_avg', Eg_avg, SIZE(Eg_avg))
  CALL TRT_PUSH_ARRAY('Eg_edgV', Eg_edgV, SIZE(Eg_edgV))
  CALL TRT_PUSH_ARRAY('Eg_edgH', Eg_edgH, SIZE(Eg_edgH))
  CALL TRT_PUSH_ARRAY('Fxg_edgV_old', Fxg_edgV_old, SIZE(Fxg_edgV_old))
  CALL TRT_PUSH_ARRAY('Fyg_edgH_old', Fyg_edgH_old, SIZE(Fyg_edgH_old))
  CALL TRT_PUSH_ARRAY('fg_avg_xx', fg_avg_xx, SIZE(fg_avg_xx))
  CALL TRT_PUSH_ARRAY('fg_avg_yy', fg_avg_yy, SIZE(fg_avg_yy))
  CALL TRT_PUSH_ARRAY('fg_edgV_xx', fg_edgV_xx, SIZE(fg_edgV_xx))
  CALL TRT_PUSH_ARRAY('fg_edgV_xy', fg_edgV_xy, SIZE(fg_edgV_xy))
  CALL TRT_PUSH_ARRAY('fg_edgH_yy', fg_edgH_yy, SIZE(fg_edgH_yy))
  CALL TRT_PUSH_ARRAY('fg_edgH_xy', fg_edgH_xy, SIZE(fg_edgH_xy))
  CALL TRT_PUSH_ARRAY('KapE', KapE, SIZE(KapE))
  CALL TRT_PUSH_ARRAY('KapR', KapR, SIZE(KapR))
  CALL TRT_PUSH_ARRAY('Pold_L', Pold_L, SIZE(Pold_L))
  CALL TRT_PUSH_ARRAY('Pold_R', Pold_R, SIZE(Pold_R))
  CALL TRT_PUSH_ARRAY('Pold_B', Pold_B, SIZE(Pold_B))
  CALL TRT_PUSH_ARRAY('Pold_T', Pold_T, SIZE(Pold_T))
  CALL TRT_CHECK(trt_grey_coefs(TRT_CTX), 'GREY_COEFS')
  CALL TRT_PULL_ARRAY('KapE_Bar', KapE_Bar, SIZE(KapE_Bar))
  CALL TRT_PULL_ARRAY('DC_xx', DC_xx, SIZE(DC_xx))
  CALL TRT_PULL_ARRAY('DL_xx', DL_xx, SIZE(DL_xx))
  CALL TRT_PULL_ARRAY('DR_xx', DR_xx, SIZE(DR_xx))
  CALL TRT_PULL_ARRAY('DC_yy', DC_yy, SIZE(DC_yy))
  CALL TRT_PULL_ARRAY('DB_yy', DB_yy, SIZE(DB_yy))
  CALL TRT_PULL_ARRAY('DT_yy', DT_yy, SIZE(DT_yy))
  CALL TRT_PULL_ARRAY('DL_xy', DL_xy, SIZE(DL_xy))
  CALL TRT_PULL_ARRAY('DR_xy', DR_xy, SIZE(DR_xy))
  CALL TRT_PULL_ARRAY('DB_xy', DB_xy, SIZE(DB_xy))
  CALL TRT_PULL_ARRAY('DT_xy', DT_xy, SIZE(DT_xy))
  CALL TRT_PULL_ARRAY('PL', PL, SIZE(PL))
  CALL TRT_PULL_ARRAY('PR', PR, SIZE(PR))
  CALL TRT_PULL_ARRAY('PB', PB, SIZE(PB))
  CALL TRT_PULL_ARRAY('PT', PT, SIZE(PT))
END SUBROUTINE GREY_COEFS

! Cbar_Calc, MainSrc/GLOQD_SOLVES.f08:177-231
SUBROUTINE Cbar_Calc(Cb_L,Cb_B,Cb_R,Cb_T,Cg_L,Cg_B,Cg_R,Cg_T,Eg_avg,Eg_edgV,Eg_edgH,Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T)
  REAL*8,INTENT(OUT):: Cb_L(:), Cb_B(:), Cb_R(:), Cb_T(:)
  REAL*8,INTENT(IN):: Cg_L(:,:), Cg_B(:,:), Cg_R(:,:), Cg_T(:,:)
  REAL*8,INTENT(IN):: Eg_avg(:,:,:), Eg_edgV(:,:,:), Eg_edgH(:,:,:)
  REAL*8,INTENT(IN):: Eg_in_L(:,:), Eg_in_B(:,:), Eg_in_R(:,:), Eg_in_T(:,:)

  CALL TRT_PUSH_ARRAY('Cg_L', Cg_L, SIZE(Cg_L))
  CALL TRT_PUSH_ARRAY('Cg_B', Cg_B, SIZE(Cg_B))
  CALL TRT_PUSH_ARRAY('Cg_R', Cg_R, SIZE(Cg_R))
  CALL TRT_PUSH_ARRAY('Cg_T', Cg_T, SIZE(Cg_T))
  CALL TRT_PUSH_ARRAY('Eg_edgV', Eg_edgV, SIZE(Eg_edgV))
  CALL TRT_PUSH_ARRAY('Eg_edgH', Eg_edgH, SIZE(Eg_edgH))
  CALL TRT_PUSH_ARRAY('Eg_in_L', Eg_in_L, SIZE(Eg_in_L))
  CALL TRT_PUSH_ARRAY('Eg_in_B', Eg_in_B, SIZE(Eg_in_B))
  CALL TRT_PUSH_ARRAY('Eg_in_R', Eg_in_R, SIZE(Eg_in_R))
  CALL TRT_PUSH_ARRAY('Eg_in_T', Eg_in_T, SIZE(Eg_in_T))
  CALL TRT_CHECK(trt_cbar_calc(TRT_CTX), 'Cbar_Calc')
  CALL TRT_PULL_ARRAY('Cb_L', Cb_L, SIZE(Cb_L))
  CALL TRT_PULL_ARRAY('Cb_B', Cb_B, SIZE(Cb_B))
  CALL TRT_PULL_ARRAY('Cb_R', Cb_R, SIZE(Cb_R))
  CALL TRT_PULL_ARRAY('Cb_T', Cb_T, SIZE(Cb_T))
END SUBROUTINE Cbar_Calc

! GQD_In_Calc, MainSrc/GLOQD_SOLVES.f08:263-290
SUBROUTINE GQD_In_Calc(E_in_L,E_in_B,E_in_R,E_in_T,F_in_L,F_in_B,F_in_R,F_in_T,Eg_in_L,Eg_in_B,Eg_in_R,Eg_in_T,&
  Fg_in_L,Fg_in_B,Fg_in_R,Fg_in_T)
  REAL*8,INTENT(OUT):: E_in_L(:), E_in_B(:), E_in_R(:), E_in_T(:)
  REAL*8,INTENT(OUT):: F_in_L(:), F_in_B(:), F_in_R(:), F_in_T(:)
  REAL*8,INTENT(IN):: Eg_in_L(:,:), Eg_in_B(:,:), Eg_in_R(:,:), Eg_in_T(:,:)
  REAL*8,INTENT(IN):: Fg_in_L(:,:), Fg_in_B(:,:), Fg_in_R(:,:), Fg_in_T(:,:)

  CALL TRT_PUSH_ARRAY('Eg_in_L', Eg_in_L, SIZE(Eg_in_L))
  CALL TRT_PUSH_ARRAY('Eg_in_B', Eg_in_B, SIZE(Eg_in_B))
  CALL TRT_PUSH_ARRAY('Eg_in_R', Eg_in_R, SIZE(Eg_in_R))
  CALL TRT_PUSH_ARRAY('Eg_in_T', Eg_in_T, SIZE(Eg_in_T))
  CALL TRT_PUSH_ARRAY('Fg_in_L', Fg_in_L, SIZE(Fg_in_L))
  CALL TRT_PUSH_ARRAY('Fg_in_B', Fg_in_B, SIZE(Fg_in_B))
  CALL TRT_PUSH_ARRAY('Fg_in_R', Fg_in_R, SIZE(Fg_in_R))
  CALL TRT_PUSH_ARRAY('Fg_in_T', Fg_in_T, SIZE(Fg_in_T))
  CALL TRT_CHECK(trt_gqd_in_calc(TRT_CTX), 'GQD_In_Calc')
  CALL TRT_PULL_ARRAY('E_in_L', E_in_L, SIZE(E_in_L))
  CALL TRT_PULL_ARRAY('E_in_B', E_in_B, SIZE(E_in_B))
  CALL TRT_PULL_ARRAY('E_in_R', E_in_R, SIZE(E_in_R))
  CALL TRT_PULL_ARRAY('E_in_T', E_in_T, SIZE(E_in_T))
  CALL TRT_PULL_ARRAY('F_in_L', F_in_L, SIZE(F_in_L))
  CALL TRT_PULL_ARRAY('F_in_B', F_in_B, SIZE(F_in_B))
  CALL TRT_PULL_ARRAY('F_in_R', F_in_R, SIZE(F_in_R))
  CALL TRT_PULL_ARRAY('F_in_T', F_in_T, SIZE(F_in_T))
END SUBROUTINE GQD_In_Calc

! EGP_FV_NEWT, MainSrc/GLOQD_SOLVES.f08:305-860.  Chi, line_src, the bounds, Eps1/Eps2, Maxits and the flags are those
! given to TRT_GPU_INIT; Temp_Mold and KapE_bar_Mold feed the finite-difference dKapE_Bar/dT (GL:402-417).
SUBROUTINE EGP_FV_NEWT(E_avg,E_edgV,E_edgH,Temp,KapE_Bar,Fx_edgV,Fy_edgH,Q_bar,KapE_Bar_dT,Its,Deltas,dresiduals,Temp_Mold,&
  KapE_bar_Mold,Theta,Delt,Delx,Dely,A,Cb_L,Cb_B,Cb_R,Cb_T,E_in_L,E_in_B,E_in_R,E_in_T,F_in_L,F_in_B,F_in_R,F_in_T,&
  DC_xx,DL_xx,DR_xx,DC_yy,DB_yy,DT_yy,DL_xy,DR_xy,DB_xy,DT_xy,PL,PR,PB,PT,Gold_Hat,Rhat_old,Kap0,cv,Comp_Unit,Chi,&
  line_src,E_Bound_Low,T_Bound_Low,Eps1,Eps2,Maxits,MGQD_It,Use_Line_Search,Use_Safety_Search,Res_Calc,kapE_dT_flag,GQD_Kits,&
  Fin_Weight,Ein_Weight)
  REAL*8,INTENT(INOUT):: E_avg(:,:), E_edgV(:,:), E_edgH(:,:)
  REAL*8,INTENT(INOUT):: Fx_edgV(:,:), Fy_edgH(:,:)
  REAL*8,INTENT(INOUT):: Temp(:,:), KapE_Bar(:,:)
  REAL*8,INTENT(OUT):: Q_bar(:,:), KapE_Bar_dT(:,:)
  INTEGER,INTENT(OUT):: Its, GQD_Kits(*)
  REAL*8,ALLOCATABLE,INTENT(OUT):: Deltas(:,:), dresiduals(:,:)
  REAL*8,INTENT(IN):: Temp_Mold(:,:), KapE_bar_Mold(:,:)
  REAL*8,INTENT(IN):: Theta, Delt
  REAL*8,INTENT(IN):: Delx(:), Dely(:), A(:,:)
  REAL*8,INTENT(IN):: Cb_L(:), Cb_B(:), Cb_R(:), Cb_T(:)
  REAL*8,INTENT(IN):: E_in_L(:), E_in_B(:), E_in_R(:), E_in_T(:)
  REAL*8,INTENT(IN):: F_in_L(:), F_in_B(:), F_in_R(:), F_in_T(:)
  REAL*8,INTENT(IN):: DC_xx(:,:), DL_xx(:,:), DR_xx(:,:)
  REAL*8,INTENT(IN):: DC_yy(:,:), DB_yy(:,:), DT_yy(:,:)
  REAL*8,INTENT(IN):: DL_xy(:,:), DR_xy(:,:), DB_xy(:,:), DT_xy(:,:)
  REAL*8,INTENT(IN):: PL(:,:), PR(:,:), PB(:,:), PT(:,:)
  REAL*8,INTENT(IN):: Gold_Hat(:,:), Rhat_old(:,:)
  REAL*8,INTENT(IN):: Kap0, cv, Comp_Unit
  REAL*8,INTENT(IN):: Fin_Weight, Ein_Weight
  REAL*8,INTENT(IN):: Chi, line_src, E_Bound_Low, T_Bound_Low, Eps1, Eps2
  INTEGER,INTENT(IN):: Maxits, MGQD_It
  LOGICAL,INTENT(IN):: Use_Line_Search, Use_Safety_Search, Res_Calc, kapE_dT_flag
  INTEGER(C_INT):: n

  IF (Res_Calc) STOP 'EGP_FV_NEWT: Res_Calc is not available on the GPU path'
  IF (Fin_Weight .NE. 1d0 .OR. Ein_Weight .NE. 1d0) STOP 'EGP_FV_NEWT: the GPU path implements run_type mlqd'
  CALL TRT_PUSH_ARRAY('E_avg', E_avg, SIZE(E_avg))
  CALL TRT_PUSH_ARRAY('E_edgV', E_edgV, SIZE(E_edgV))
  CALL TRT_PUSH_ARRAY('E_edgH', E_edgH, SIZE(E_edgH))
  CALL TRT_PUSH_ARRAY('Fx_edgV', Fx_edgV, SIZE(Fx_edgV))
  CALL TRT_PUSH_ARRAY('Fy_edgH', Fy_edgH, SIZE(Fy_edgH))
  CALL TRT_PUSH_ARRAY('Temp', Temp, SIZE(Temp))
  CALL TRT_PUSH_ARRAY('KapE_Bar', KapE_Bar, SIZE(KapE_Bar))
  CALL TRT_PUSH_ARRAY('Temp_MGQDold2', Temp_Mold, SIZE(Temp_Mold))
  CALL TRT_PUSH_ARRAY('KapE_bar_MGQDold', KapE_bar_Mold, SIZE(KapE_bar_Mold))
  CALL TRT_PUSH_ARRAY('Cb_L', Cb_L, SIZE(Cb_L))
  CALL TRT_PUSH_ARRAY('Cb_B', Cb_B, SIZE(Cb_B))
  CALL TRT_PUSH_ARRAY('Cb_R', Cb_R, SIZE(Cb_R))
  CALL TRT_PUSH_ARRAY('Cb_T', Cb_T, SIZE(Cb_T))
  CALL TRT_PUSH_ARRAY('E_in_L', E_in_L, SIZE(E_in_L))
  CALL TRT_PUSH_ARRAY('E_in_B', E_in_B, SIZE(E_in_B))
  CALL TRT_PUSH_ARRAY('E_in_R', E_in_R, SIZE(E_in_R))
  CALL TRT_PUSH_ARRAY('E_in_T', E_in_T, SIZE(E_in_T))
  CALL TRT_PUSH_ARRAY('F_in_L', F_in_L, SIZE(F_in_L))
  CALL TRT_PUSH_ARRAY('F_in_B', F_in_B, SIZE(F_in_B))
  CALL TRT_PUSH_ARRAY('F_in_R', F_in_R, SIZE(F_in_R))
  CALL TRT_PUSH_ARRAY('F_in_T', F_in_T, SIZE(F_in_T))
  CALL TRT_PUSH_ARRAY('DC_xx', DC_xx, SIZE(DC_xx))
  CALL TRT_PUSH_ARRAY('DL_xx', DL_xx, SIZE(DL_xx))
  CALL TRT_PUSH_ARRAY('DR_xx', DR_xx, SIZE(DR_xx))
  CALL TRT_PUSH_ARRAY('DC_yy', DC_yy, SIZE(DC_yy))
  CALL TRT_PUSH_ARRAY('DB_yy', DB_yy, SIZE(DB_yy))
  CALL TRT_PUSH_ARRAY('DT_yy', DT_yy, SIZE(DT_yy))
  CALL TRT_PUSH_ARRAY('DL_xy', DL_xy, SIZE(DL_xy))
  CALL TRT_PUSH_ARRAY('DR_xy', DR_xy, SIZE(DR_xy))
  CALL TRT_PUSH_ARRAY('DB_xy', DB_xy, SIZE(DB_xy))
  CALL TRT_PUSH_ARRAY('DT_xy', DT_xy, SIZE(DT_xy))
  CALL TRT_PUSH_ARRAY('PL', PL, SIZE(PL))
  CALL TRT_PUSH_ARRAY('PR', PR, SIZE(PR))
  CALL TRT_PUSH_ARRAY('PB', PB, SIZE(PB))
  CALL TRT_PUSH_ARRAY('PT', PT, SIZE(PT))
  CALL TRT_PUSH_ARRAY('Gold_Hat', Gold_Hat, SIZE(Gold_Hat))
  CALL TRT_PUSH_ARRAY('Rhat_old', Rhat_old, SIZE(Rhat_old))
  CALL TRT_CHECK(trt_egp_fv_newt(TRT_CTX, INT(MGQD_It, C_INT), n), 'EGP_FV_NEWT')
  Its = n
  GQD_Kits(1:Its) = 0
  CALL TRT_PULL_ARRAY('E_avg', E_avg, SIZE(E_avg))
  CALL TRT_PULL_ARRAY('E_edgV', E_edgV, SIZE(E_edgV))
  CALL TRT_PULL_ARRAY('E_edgH', E_edgH, SIZE(E_edgH))
  CALL TRT_PULL_ARRAY('Fx_edgV', Fx_edgV, SIZE(Fx_edgV))
  CALL TRT_PULL_ARRAY('Fy_edgH', Fy_edgH, SIZE(Fy_edgH))
  CALL TRT_PULL_ARRAY('Temp', Temp, SIZE(Temp))
  CALL TRT_PULL_ARRAY('KapE_Bar', KapE_Bar, SIZE(KapE_Bar))
  CALL TRT_PULL_ARRAY('GQD_Src', Q_bar, SIZE(Q_bar))
  CALL TRT_PULL_ARRAY('KapE_Bar_dT', KapE_Bar_dT, SIZE(KapE_Bar_dT))
END SUBROUTINE EGP_FV_NEWT

END MODULE GLOQD_SOLVES

!==================================================================================================================================!
MODULE UPDATES

  USE TRT_CAPI
  IMPLICIT NONE

CONTAINS

! MATERIAL_UPDATE, MainSrc/UPDATES.f08:110-156 (Mat and Kappa_Mult are those given to TRT_GPU_INIT)
SUBROUTINE MATERIAL_UPDATE(RT_Src,MGQD_Src,KapE,KapB,KapR,Bg,Temp,Comp_Unit,Nu_g,Open_Threads,Mat,Kappa_Mult)
  REAL*8,INTENT(OUT):: RT_Src(:,:,:,:), MGQD_Src(:,:,:)
  REAL*8,INTENT(OUT):: KapE(:,:,:), KapB(:,:,:), KapR(:,:,:), Bg(:,:,:)
  REAL*8,INTENT(IN):: Temp(:,:), comp_unit, nu_g(:), Kappa_Mult(:)
  INTEGER,INTENT(IN):: Open_Threads, Mat(:,:)
  REAL*8,ALLOCATABLE:: src(:,:,:)
  INTEGER:: i, j, g, m

  CALL TRT_PUSH_ARRAY('Temp', Temp, SIZE(Temp))
  CALL TRT_CHECK(trt_material_update(TRT_CTX), 'MATERIAL_UPDATE')
  CALL TRT_PULL_ARRAY('MGQD_Src', MGQD_Src, SIZE(MGQD_Src))
  CALL TRT_PULL_ARRAY('KapE', KapE, SIZE(KapE))
  CALL TRT_PULL_ARRAY('KapB', KapB, SIZE(KapB))
  CALL TRT_PULL_ARRAY('KapR', KapR, SIZE(KapR))
  CALL TRT_PULL_ARRAY('Bg', Bg, SIZE(Bg))
  ALLOCATE(src(SIZE(KapE,1), SIZE(KapE,2), SIZE(KapE,3)))
  CALL TRT_PULL_ARRAY('RT_Src', src, SIZE(src))
  DO g = 1, SIZE(src,3)
    DO m = 1, SIZE(RT_Src,3)
      DO j = 1, SIZE(src,2)
        DO i = 1, SIZE(src,1)
          RT_Src(2*i-1,2*j-1,m,g) = src(i,j,g)
          RT_Src(2*i,2*j-1,m,g) = src(i,j,g)
          RT_Src(2*i-1,2*j,m,g) = src(i,j,g)
          RT_Src(2*i,2*j,m,g) = src(i,j,g)
        END DO
      END DO
    END DO
  END DO
  DEALLOCATE(src)
END SUBROUTINE MATERIAL_UPDATE

! RT_BC_UPDATE, MainSrc/UPDATES.f08:12-105 (BC_Type, bcT_*, Nu_g and Comp_Unit are those given to TRT_GPU_INIT): on the
! reflective sides the incoming directions take the mirror directions' outgoing intensities of the last TRANSPORT_SCB; the
! four intensity arrays are returned with the updated incoming faces.
SUBROUTINE RT_BC_UPDATE(BC_Type,bcT_left,bcT_lower,bcT_right,bcT_upper,Comp_Unit,Nu_g,I_edgV,I_edgH,Ic_edgV,Ic_edgH)
  REAL*8,INTENT(IN):: bcT_left, bcT_lower, bcT_right, bcT_upper
  REAL*8,INTENT(IN):: Comp_Unit, Nu_g(:)
  INTEGER,INTENT(IN):: BC_Type(:)
  REAL*8,INTENT(OUT):: I_edgV(:,:,:,:), I_edgH(:,:,:,:), Ic_edgV(:,:,:,:), Ic_edgH(:,:,:,:)

  CALL TRT_CHECK(trt_rt_bc_update(TRT_CTX), 'RT_BC_UPDATE')
  CALL TRT_PULL_ARRAY('I_edgV', I_edgV, SIZE(I_edgV))
  CALL TRT_PULL_ARRAY('I_edgH', I_edgH, SIZE(I_edgH))
  CALL TRT_PULL_ARRAY('Ic_edgV', Ic_edgV, SIZE(Ic_edgV))
  CALL TRT_PULL_ARRAY('Ic_edgH', Ic_edgH, SIZE(Ic_edgH))
END SUBROUTINE RT_BC_UPDATE

! MEB_SOLVE, MainSrc/UPDATES.f08:189-231 (Bg, Delt, cV, Kap0, Comp_Unit are the context's; Eg_HO is
! pushed as the context's Eg_avg, the array the MLOQD loop passes, ALGORITHMS.f08:769)
SUBROUTINE MEB_SOLVE(Temp,Eg_HO,Bg,KapE,Temp_old,Delt,cV,Kap0,Comp_Unit)
  REAL*8,INTENT(INOUT):: Temp(:,:)
  REAL*8,INTENT(IN):: Eg_HO(:,:,:), Bg(:,:,:), KapE(:,:,:)
  REAL*8,INTENT(IN):: Temp_old(:,:), Delt, cV, Kap0, Comp_Unit

  CALL TRT_PUSH_ARRAY('Eg_avg', Eg_HO, SIZE(Eg_HO))
  CALL TRT_PUSH_ARRAY('KapE', KapE, SIZE(KapE))
  CALL TRT_PUSH_ARRAY('Temp_Old', Temp_old, SIZE(Temp_old))
  CALL TRT_CHECK(trt_meb_solve(TRT_CTX, 'Eg_avg'//C_NULL_CHAR), 'MEB_SOLVE')
  CALL TRT_PULL_ARRAY('Temp', Temp, SIZE(Temp))
END SUBROUTINE MEB_SOLVE

END MODULE UPDATES

!==================================================================================================================================!
MODULE CONVERGENCE_CHECKS

  USE TRT_CAPI
  IMPLICIT NONE

CONTAINS

! CONVERGENCE, MainSrc/CONVERGENCE_CHECKS.f08:7-44, evaluated by the device reductions on copies of x, x_old, x_old2
! (cell-sized arrays, as every call site of the reference passes).
FUNCTION CONVERGENCE(Conv_Type,Eps,x,x_old,x_old2,Its,Norm,Rho)
  REAL*8,INTENT(IN):: Eps
  REAL*8,INTENT(IN):: x(:,:), x_old(:,:), x_old2(:,:)
  INTEGER,INTENT(IN):: Conv_Type, Its
  REAL*8,INTENT(OUT):: Norm
  REAL*8,INTENT(OUT):: Rho
  LOGICAL:: CONVERGENCE
  INTEGER(C_INT):: conv

  CALL TRT_PUSH_ARRAY('_conv_x', x, SIZE(x))
  CALL TRT_PUSH_ARRAY('_conv_x_old', x_old, SIZE(x_old))
  CALL TRT_PUSH_ARRAY('_conv_x_old2', x_old2, SIZE(x_old2))
  CALL TRT_CHECK(trt_convergence(TRT_CTX, '_conv_x'//C_NULL_CHAR, '_conv_x_old'//C_NULL_CHAR, &
    '_conv_x_old2'//C_NULL_CHAR, INT(Conv_Type, C_INT), Eps, INT(Its, C_INT), Norm, Rho, conv), 'CONVERGENCE')
  CONVERGENCE = conv .NE. 0
END FUNCTION CONVERGENCE

END MODULE CONVERGENCE_CHECKS
