! iso_c_binding interface to libhosnwt (include/hosnwt.h).
!
! A Fortran caller of the reference's solver loop (main/HOS-NWT.f90:446-581)
! replaces CALL runge4_adapt2(...) by hos_rk_step / hos_commit, or the whole
! inner DO WHILE(time_cur < time_next) loop by hos_advance.  Arrays are the
! reference's column-major REAL(RP) (m1,m2) / (m3_add,m2) arrays, passed by
! reference.
MODULE hosnwt_gpu
USE, INTRINSIC :: iso_c_binding
IMPLICIT NONE
!
TYPE, BIND(C) :: hos_params
   INTEGER(C_INT) :: n1, n2, n3, mHOS, p1, p2, icase, igeom, i_wmk, iramp, iabsnb, precision
   REAL(C_DOUBLE) :: xlen, ylen, d_hinge, Tramp, T_stop, T_mono
   TYPE(C_PTR)    :: nu
   INTEGER(C_INT) :: n_freq
   TYPE(C_PTR)    :: freq, pos_ini_re, pos_ini_im
   REAL(C_DOUBLE) :: coeffilt(3) = 1.0_C_DOUBLE
   TYPE(C_PTR)    :: dposdy_ini_re = C_NULL_PTR, dposdy_ini_im = C_NULL_PTR, nu_x = C_NULL_PTR, nu_y = C_NULL_PTR
   ! dipole wavemaker (igeom = 3): per component frequency, mu(:,1), mu(:,2), mu(:,3)
   INTEGER(C_INT) :: n_dip = 0
   TYPE(C_PTR)    :: dip_freq = C_NULL_PTR, dip_mux_re = C_NULL_PTR, dip_mux_im = C_NULL_PTR
   TYPE(C_PTR)    :: dip_muz_re = C_NULL_PTR, dip_muz_im = C_NULL_PTR, dip_z = C_NULL_PTR
   REAL(C_DOUBLE) :: x_dip = 2.5_C_DOUBLE
   INTEGER(C_INT) :: n_side = 6
END TYPE hos_params
!
INTERFACE
   INTEGER(C_INT) FUNCTION hos_create(p, ctx) BIND(C, NAME='hos_create')
      IMPORT :: C_INT, C_PTR, hos_params
      TYPE(hos_params), INTENT(IN) :: p
      TYPE(C_PTR), INTENT(OUT)     :: ctx
   END FUNCTION hos_create
   INTEGER(C_INT) FUNCTION hos_destroy(ctx) BIND(C, NAME='hos_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END FUNCTION hos_destroy
   TYPE(C_PTR) FUNCTION hos_last_error() BIND(C, NAME='hos_last_error')
      IMPORT :: C_PTR
   END FUNCTION hos_last_error
   INTEGER(C_INT) FUNCTION hos_initialize(ctx) BIND(C, NAME='hos_initialize')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END FUNCTION hos_initialize
   INTEGER(C_INT) FUNCTION hos_set_state(ctx, eta, phis, a1st_add, a_add_ramp, a1st, eta1st) &
        BIND(C, NAME='hos_set_state')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: eta(*), phis(*), a1st_add(*), a_add_ramp(*), a1st(*), eta1st(*)
   END FUNCTION hos_set_state
   INTEGER(C_INT) FUNCTION hos_get_state(ctx, eta, phis, a1st_add, a_add_ramp, a1st, eta1st) &
        BIND(C, NAME='hos_get_state')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: eta(*), phis(*), a1st_add(*), a_add_ramp(*), a1st(*), eta1st(*)
   END FUNCTION hos_get_state
   ! second-order wavemaker state (i_wmk = 2): a2nd_add (m3_add,m2), a2nd (modal), eta2nd
   INTEGER(C_INT) FUNCTION hos_set_state2(ctx, a2nd_add, a2nd, eta2nd) BIND(C, NAME='hos_set_state2')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: a2nd_add(*), a2nd(*), eta2nd(*)
   END FUNCTION hos_set_state2
   INTEGER(C_INT) FUNCTION hos_get_state2(ctx, a2nd_add, a2nd, eta2nd) BIND(C, NAME='hos_get_state2')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: a2nd_add(*), a2nd(*), eta2nd(*)
   END FUNCTION hos_get_state2
   ! third-order wavemaker paddle modes (i_wmk = 3): a3rd_add (m3_add,m2)
   INTEGER(C_INT) FUNCTION hos_set_state3(ctx, a3rd_add) BIND(C, NAME='hos_set_state3')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: a3rd_add(*)
   END FUNCTION hos_set_state3
   INTEGER(C_INT) FUNCTION hos_get_state3(ctx, a3rd_add) BIND(C, NAME='hos_get_state3')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: a3rd_add(*)
   END FUNCTION hos_get_state3
   ! one solveHOS2 evaluation (resol_HOS.f90:270-656)
   INTEGER(C_INT) FUNCTION hos_rhs(ctx, t, delta_t, G1, G2, a1st, eta1st, a_add_1, dG1, dG2, da1st, deta1st, &
        da_add_1, da_add_ramp) BIND(C, NAME='hos_rhs')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: t, delta_t
      REAL(C_DOUBLE), INTENT(IN)  :: G1(*), G2(*), a1st(*), eta1st(*), a_add_1(*)
      REAL(C_DOUBLE), INTENT(OUT) :: dG1(*), dG2(*), da1st(*), deta1st(*), da_add_1(*), da_add_ramp(*)
   END FUNCTION hos_rhs
   INTEGER(C_INT) FUNCTION hos_rhs2(ctx, t, delta_t, G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2, &
        dG1, dG2, da1st, deta1st, da_add_1, da_add_ramp, da2nd, deta2nd, da_add_2) BIND(C, NAME='hos_rhs2')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: t, delta_t
      REAL(C_DOUBLE), INTENT(IN)  :: G1(*), G2(*), a1st(*), eta1st(*), a_add_1(*), a2nd(*), eta2nd(*), a_add_2(*)
      REAL(C_DOUBLE), INTENT(OUT) :: dG1(*), dG2(*), da1st(*), deta1st(*), da_add_1(*), da_add_ramp(*)
      REAL(C_DOUBLE), INTENT(OUT) :: da2nd(*), deta2nd(*), da_add_2(*)
   END FUNCTION hos_rhs2
   INTEGER(C_INT) FUNCTION hos_rhs3(ctx, t, delta_t, G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2, &
        a_add_3, dG1, dG2, da1st, deta1st, da_add_1, da_add_ramp, da2nd, deta2nd, da_add_2, da_add_3) &
        BIND(C, NAME='hos_rhs3')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: t, delta_t
      REAL(C_DOUBLE), INTENT(IN)  :: G1(*), G2(*), a1st(*), eta1st(*), a_add_1(*), a2nd(*), eta2nd(*), a_add_2(*)
      REAL(C_DOUBLE), INTENT(IN)  :: a_add_3(*)
      REAL(C_DOUBLE), INTENT(OUT) :: dG1(*), dG2(*), da1st(*), deta1st(*), da_add_1(*), da_add_ramp(*)
      REAL(C_DOUBLE), INTENT(OUT) :: da2nd(*), deta2nd(*), da_add_2(*), da_add_3(*)
   END FUNCTION hos_rhs3
   INTEGER(C_INT) FUNCTION hos_rk_step(ctx, t, h, err) BIND(C, NAME='hos_rk_step')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: t, h
      REAL(C_DOUBLE), INTENT(OUT) :: err
   END FUNCTION hos_rk_step
   INTEGER(C_INT) FUNCTION hos_commit(ctx) BIND(C, NAME='hos_commit')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END FUNCTION hos_commit
   INTEGER(C_INT) FUNCTION hos_end_ramp(ctx) BIND(C, NAME='hos_end_ramp')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END FUNCTION hos_end_ramp
   INTEGER(C_INT) FUNCTION hos_advance(ctx, t_cur, t_next, h_rk, toler, dt_stab, delt_out, n_acc, n_rej) &
        BIND(C, NAME='hos_advance')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(INOUT) :: t_cur, h_rk
      REAL(C_DOUBLE), VALUE :: t_next, toler, dt_stab, delt_out
      INTEGER(C_INT), INTENT(OUT) :: n_acc, n_rej
   END FUNCTION hos_advance
   INTEGER(C_INT) FUNCTION hos_advance_ensemble(n, ctx, t_cur, t_next, h_rk, toler, dt_stab, delt_out, &
        n_acc, n_rej) BIND(C, NAME='hos_advance_ensemble')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      INTEGER(C_INT), VALUE :: n
      TYPE(C_PTR), INTENT(IN) :: ctx(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: t_cur(*), h_rk(*)
      REAL(C_DOUBLE), VALUE :: t_next, toler
      REAL(C_DOUBLE), INTENT(IN) :: dt_stab(*), delt_out(*)
      INTEGER(C_INT), INTENT(OUT) :: n_acc(*), n_rej(*)
   END FUNCTION hos_advance_ensemble
   INTEGER(C_INT) FUNCTION hos_get_physical(ctx, eta, phis, deta, phi_add) BIND(C, NAME='hos_get_physical')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: eta(*), phis(*), deta(*), phi_add(*)
   END FUNCTION hos_get_physical
   INTEGER(C_INT) FUNCTION hos_get_wmk(ctx, alpha, pos_top) BIND(C, NAME='hos_get_wmk')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: alpha, pos_top
   END FUNCTION hos_get_wmk
   INTEGER(C_INT) FUNCTION hos_get_wmk_motion(ctx, alpha, alphat, x_top, v_top) BIND(C, NAME='hos_get_wmk_motion')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: alpha, alphat, x_top(*), v_top(*)
   END FUNCTION hos_get_wmk_motion
   INTEGER(C_INT) FUNCTION hos_get_modes(ctx, eta_modes, phis_modes) BIND(C, NAME='hos_get_modes')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: eta_modes(*), phis_modes(*)
   END FUNCTION hos_get_modes
   ! SWENSE output modes at an output time (IO/velocities.f90:223-784, output.f90:417-467)
   INTEGER(C_INT) FUNCTION hos_swense_modes(ctx, t, meta2, specx, specy, specz, spect, modes_fs, modes_fst, &
        modes_add, modes_addt) BIND(C, NAME='hos_swense_modes')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: t
      INTEGER(C_INT), VALUE :: meta2
      REAL(C_DOUBLE), INTENT(OUT) :: specx(*), specy(*), specz(*), spect(*), modes_fs(*), modes_fst(*)
      REAL(C_DOUBLE), INTENT(OUT) :: modes_add(*), modes_addt(*)
   END FUNCTION hos_swense_modes
   INTEGER(C_INT) FUNCTION hos_probes(ctx, np, xp, yp, eta_p) BIND(C, NAME='hos_probes')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: np
      REAL(C_DOUBLE), INTENT(IN)  :: xp(*), yp(*)
      REAL(C_DOUBLE), INTENT(OUT) :: eta_p(*)
   END FUNCTION hos_probes
   INTEGER(C_INT) FUNCTION hos_get_slope_max(ctx, sx, sy) BIND(C, NAME='hos_get_slope_max')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: sx, sy
   END FUNCTION hos_get_slope_max
   INTEGER(C_INT) FUNCTION hos_check_slope(ctx, threshold, coeff, fx, fy) BIND(C, NAME='hos_check_slope')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: threshold, coeff
      INTEGER(C_INT), INTENT(OUT) :: fx, fy
   END FUNCTION hos_check_slope
END INTERFACE
!
CONTAINS
!
! Map a nonzero status to ERROR STOP with the library's message
! (the reference STOPs on bad input, e.g. initial_condition.f90:41-59).
SUBROUTINE hos_check(rc, what)
   INTEGER(C_INT), INTENT(IN)   :: rc
   CHARACTER(LEN=*), INTENT(IN) :: what
   CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
   CHARACTER(LEN=512) :: text
   INTEGER :: i
   IF (rc == 0) RETURN
   CALL C_F_POINTER(hos_last_error(), msg, [512])
   text = ' '
   DO i = 1, 512
      IF (msg(i) == C_NULL_CHAR) EXIT
      text(i:i) = msg(i)
   END DO
   WRITE(*,'(A,A,A)') TRIM(what), ': ', TRIM(text)
   ERROR STOP 'libhosnwt call failed'
END SUBROUTINE hos_check
!
END MODULE hosnwt_gpu
