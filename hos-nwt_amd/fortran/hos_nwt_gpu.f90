! Fortran driver of the GPU time-stepping core.
!
! Reads the reference's input_HOS-NWT.dat (same 52-column description + value
! format, IO/input.f90:62-133), nondimensionalises (main/HOS-NWT.f90:99-119),
! builds the paddle motion of a monochromatic wave (icase = 2, front wave,
! piston or hinged flap) or a sloshing mode (icase = 1), the front beach, the
! RK4 stability bound (main/HOS-NWT.f90:132-196), and runs the time loop of
! main/HOS-NWT.f90:446-581 with the inner adaptive loop on the GPU
! (hos_advance).  Writes Results/vol_energy.dat (IO/output.f90:60-70,
! :251-262).  The reference's sizes are compile-time parameters
! (common_vars.f90:31-35); here they are runtime arguments:
!     hos_nwt_gpu [n1 n2 n3 mHOS [precision]]      (default 513 1 33 3 0)
PROGRAM hos_nwt_gpu
USE, INTRINSIC :: iso_c_binding
USE hosnwt_gpu
IMPLICIT NONE
INTEGER, PARAMETER :: RP = KIND(1.0D0), N_TOT = 52
REAL(RP), PARAMETER :: PI = 3.141592653589793238462643383279502888419_rp, G = 9.81_rp
REAL(RP), PARAMETER :: TWOPI = 2.0_rp * PI, TINY = EPSILON(1.0_rp)
INTEGER :: n1, n2, n3, mHOS, prec, nargs
CHARACTER(LEN=32) :: arg
! input_HOS-NWT.dat fields
REAL(RP) :: xlen, ylen, h, aslosh, amp_mono, nu_mono, theta_mono, ph_mono, xd_mono, nuc_low, nuc_high
REAL(RP) :: Hs, Tp, gamma, d_hinge, Tramp, xabsf, coeffabsf, T_stop, toler, f_out
INTEGER :: icase, islosh, ibat, i_cut, iseed, i_wmk, igeom, iramp, iabsnb, iprobes, idim, i3d, imodes, iwmk, i_sw
CHARACTER(LEN=100) :: file_name, pro_file
! derived
REAL(RP) :: Tscale, T_mono, w_mono, dt_stab, delt, delt_out, time_cur, time_next, h_rk
REAL(RP) :: k_w, freq1, delx, dely, x0, x2, x3, c8, d1, d2, d3, voltot, enepot, enekin, alpha, pos_top
REAL(RP) :: t_adim
COMPLEX(RP) :: tf, ctmp
REAL(RP), ALLOCATABLE, TARGET :: nu(:,:), eta(:,:), phis(:,:), deta(:,:), phi_add(:,:), zero2(:,:), zadd(:,:)
REAL(RP), ALLOCATABLE, TARGET :: freq(:), pre(:), pim(:), dyre(:), dyim(:), nux(:,:)
TYPE(hos_params) :: p
TYPE(C_PTR) :: ctx
INTEGER(C_INT) :: n_acc, n_rej, fx, fy
INTEGER :: i1, i2, n_rk_tot, n_er_tot, unit_in
!
n1 = 513; n2 = 1; n3 = 33; mHOS = 3; prec = 0
nargs = COMMAND_ARGUMENT_COUNT()
IF (nargs >= 4) THEN
   CALL GET_COMMAND_ARGUMENT(1, arg); READ(arg,*) n1
   CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg,*) n2
   CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg,*) n3
   CALL GET_COMMAND_ARGUMENT(4, arg); READ(arg,*) mHOS
END IF
IF (nargs >= 5) THEN
   CALL GET_COMMAND_ARGUMENT(5, arg); READ(arg,*) prec
END IF
!
unit_in = 100
OPEN(unit_in, FILE='input_HOS-NWT.dat', STATUS='OLD')
CALL blank(); CALL rd_r(xlen); CALL rd_r(ylen); CALL rd_r(h)
CALL blank(); CALL rd_i(icase)
CALL blank(); CALL rd_i(islosh); CALL rd_r(aslosh)
CALL blank(); CALL rd_r(amp_mono); CALL rd_r(nu_mono); CALL rd_r(theta_mono); CALL rd_r(ph_mono)
CALL rd_i(ibat); CALL rd_r(xd_mono)
CALL blank(); CALL rd_c(file_name); CALL rd_i(i_cut); CALL rd_r(nuc_low); CALL rd_r(nuc_high)
CALL blank(); CALL rd_r(Hs); CALL rd_r(Tp); CALL rd_r(gamma); CALL rd_i(iseed)
CALL blank(); CALL rd_i(i_wmk); CALL rd_i(igeom); CALL rd_r(d_hinge); CALL rd_i(iramp); CALL rd_r(Tramp)
CALL blank(); CALL rd_i(iabsnb); CALL rd_r(xabsf); CALL rd_r(coeffabsf)
CALL blank(); CALL rd_i(iprobes); CALL rd_c(pro_file)
CALL blank(); CALL rd_r(T_stop); CALL rd_r(toler); CALL rd_r(f_out)
CALL blank(); CALL rd_i(idim); CALL rd_i(i3d); CALL rd_i(imodes); CALL rd_i(iwmk); CALL rd_i(i_sw)
CLOSE(unit_in)
IF (icase /= 1 .AND. icase /= 2) ERROR STOP 'this driver builds icase = 1 or 2 (files: use the Python host)'
IF (icase == 2 .AND. ibat /= 2) ERROR STOP 'only the front wave (ibat = 2) is built here'
islosh = islosh + 1
IF (n2 == 1) theta_mono = 0.0_rp
IF (iramp == 0) Tramp = 0.0_rp
IF (i_wmk < 1 .OR. i_wmk > 3) ERROR STOP 'i_wmk = 1, 2 or 3 is built here'
!
! nondimensionalisation (HOS-NWT.f90:99-119)
Tscale = 1.0_rp / SQRT(G / h)
xlen = xlen / h; ylen = ylen / h; amp_mono = amp_mono / h; d_hinge = d_hinge / h; aslosh = aslosh / h
Tramp = Tramp / Tscale; T_stop = T_stop / Tscale; f_out = f_out * Tscale; nu_mono = nu_mono * Tscale
w_mono = nu_mono * TWOPI; T_mono = 1.0_rp / nu_mono
IF (n2 == 1) ylen = 1.0_rp
!
! RK4 stability bound (HOS-NWT.f90:132-196)
c8 = SQRT(8.0_rp / PI**2)
d1 = c8 * xlen / REAL(n1 - 1, RP)
d2 = d1 + 1.0_rp
IF (n2 /= 1) d2 = c8 * ylen / REAL(n2 - 1, RP)
d3 = c8 * 4.0_rp / REAL(2 * (n3 - 1), RP)
IF (icase == 1) THEN
   dt_stab = MIN(d1, d2)
ELSE
   dt_stab = MIN(d1, d2, d3)
END IF
delt_out = 1.0_rp / f_out
delt = MIN(0.95_rp * dt_stab, delt_out)
!
ALLOCATE(nu(n1,n2), eta(n1,n2), phis(n1,n2), deta(n1,n2), phi_add(n1,n2), zero2(n1,n2), zadd(2*(n3-1),n2))
nu = 0.0_rp; eta = 0.0_rp; phis = 0.0_rp; zero2 = 0.0_rp; zadd = 0.0_rp
delx = xlen / REAL(n1 - 1, RP)
dely = 0.0_rp
IF (n2 /= 1) dely = ylen / REAL(n2 - 1, RP)
! front beach (initial_condition.f90:724-794)
IF (iabsnb /= 0 .AND. ABS(xabsf - 1.0_rp) > TINY) THEN
   x0 = xlen * xabsf; x2 = xlen - x0
   DO i1 = 1, n1
      IF ((i1 - 1) * delx >= x0) THEN
         x3 = ((i1 - 1) * delx - x0) / x2
         nu(i1,:) = coeffabsf * x3 * x3
      END IF
   END DO
END IF
!
ALLOCATE(freq(1), pre(n2), pim(n2), dyre(n2), dyim(n2), nux(n1,n2))
freq = 0.0_rp; pre = 0.0_rp; pim = 0.0_rp; dyre = 0.0_rp; dyim = 0.0_rp; nux = 0.0_rp
! beach x derivative as initial_condition.f90:768 builds it (2 c x3), read with i_wmk = 2
IF (iabsnb /= 0 .AND. ABS(xabsf - 1.0_rp) > TINY) THEN
   DO i1 = 1, n1
      IF ((i1 - 1) * delx >= x0) nux(i1,:) = coeffabsf * 2.0_rp * ((i1 - 1) * delx - x0) / x2
   END DO
END IF
IF (icase == 2) THEN
   ! front(): one harmonic at nu_mono (wave_def.f90:54-80), wave_init and wmk_ini
   freq1 = nu_mono
   k_w = wave_number(freq1)
   tf = transfer(igeom, k_w, d_hinge)
   DO i2 = 1, n2
      ctmp = tf * COS(theta_mono * PI / 180.0_rp) * amp_mono * &
             EXP(CMPLX(0.0_rp, -k_w * SIN(theta_mono * PI / 180.0_rp) * (i2 - 1) * dely + ph_mono, RP))
      pre(i2) = REAL(ctmp)
      pim(i2) = AIMAG(ctmp)
      ! y derivative of the paddle amplitude (wavemaking.f90:147-149, no lateral ramp)
      ctmp = CMPLX(0.0_rp, -k_w * SIN(theta_mono * PI / 180.0_rp), RP) * ctmp
      dyre(i2) = REAL(ctmp)
      dyim(i2) = AIMAG(ctmp)
   END DO
   freq(1) = freq1
END IF
!
p%n1 = n1; p%n2 = n2; p%n3 = n3; p%mHOS = mHOS; p%p1 = mHOS
p%p2 = 1
IF (n2 /= 1) p%p2 = mHOS
p%icase = icase; p%igeom = igeom; p%i_wmk = i_wmk; p%iramp = iramp; p%iabsnb = iabsnb; p%precision = prec
p%xlen = xlen; p%ylen = ylen; p%d_hinge = d_hinge; p%Tramp = Tramp; p%T_stop = T_stop; p%T_mono = T_mono
p%nu = C_LOC(nu)
p%n_freq = 0
IF (icase == 2) p%n_freq = 1
p%freq = C_LOC(freq); p%pos_ini_re = C_LOC(pre); p%pos_ini_im = C_LOC(pim)
p%dposdy_ini_re = C_LOC(dyre); p%dposdy_ini_im = C_LOC(dyim); p%nu_x = C_LOC(nux)
CALL hos_check(hos_create(p, ctx), 'hos_create')
IF (icase == 1) THEN
   DO i1 = 1, n1
      eta(i1,:) = aslosh * COS((islosh - 1) * PI / xlen * (i1 - 1) * delx)
   END DO
   CALL hos_check(hos_set_state(ctx, eta, phis, zadd, zadd, zero2, zero2), 'hos_set_state')
ELSE
   CALL hos_check(hos_initialize(ctx), 'hos_initialize')
END IF
!
t_adim = 1.0_rp
IF (idim == 1) t_adim = Tscale
CALL EXECUTE_COMMAND_LINE('mkdir -p Results')
OPEN(2, FILE='Results/vol_energy.dat', STATUS='UNKNOWN')
WRITE(2,'(A)') 'TITLE=" volume and energy evolution "'
IF (idim == 0) THEN
   WRITE(2,'(A)') 'VARIABLES="time*sqrt(g/h)", "voltot", "enepot", "enekin", "enetot"'
ELSE
   WRITE(2,'(A)') 'VARIABLES="time", "voltot", "enepot", "enekin", "enetot"'
END IF
!
time_cur = 0.0_rp
n_rk_tot = 0; n_er_tot = 0; n_acc = 0; n_rej = 0
DO WHILE (time_cur < T_stop + delt_out)
   n_rk_tot = n_rk_tot + n_acc
   n_er_tot = n_er_tot + n_rej
   CALL hos_check(hos_get_physical(ctx, eta, phis, deta, phi_add), 'hos_get_physical')
   CALL hos_check(hos_get_wmk(ctx, alpha, pos_top), 'hos_get_wmk')
   CALL volume_energy()
   IF (igeom == 2 .AND. icase == 2) THEN
      WRITE(2,'(5(ES12.5,X))') time_cur * t_adim, voltot - 0.5_rp * alpha * (1 - d_hinge) * pos_top / xlen &
           - eta(1,1) * alpha * pos_top / xlen, enepot, enekin, enepot + enekin
   ELSE IF (icase == 2) THEN
      WRITE(2,'(5(ES12.5,X))') time_cur * t_adim, voltot - alpha * pos_top / xlen - eta(1,1) * alpha * pos_top / xlen, &
           enepot, enekin, enepot + enekin
   ELSE
      WRITE(2,'(5(ES12.5,X))') time_cur * t_adim, voltot, enepot, enekin, enepot + enekin
   END IF
   WRITE(*,'("    volume = ",ES12.4," t= ",ES11.4," timestep:",ES11.4,2(X,I5))') voltot, time_cur * t_adim, &
        delt * t_adim, n_rk_tot, n_er_tot
   IF (ABS(time_cur - T_stop) <= TINY) EXIT
   time_next = time_cur + delt_out
   IF (time_next > T_stop + delt_out) time_next = T_stop + delt_out
   h_rk = delt
   CALL hos_check(hos_advance(ctx, time_cur, time_next, h_rk, toler, dt_stab, delt_out, n_acc, n_rej), 'hos_advance')
   delt = h_rk
   ! check_slope (initial_condition.f90:687-716)
   CALL hos_check(hos_check_slope(ctx, 8.0_rp, 0.7_rp, fx, fy), 'hos_check_slope')
   IF (fx /= 0 .OR. fy /= 0) THEN
      WRITE(*,'(A)') '******************** Slope of FS elevation too high ******************** '
      WRITE(*,'(A)') 'Trying to filter eta and phis...'
   END IF
END DO
CLOSE(2)
CALL hos_check(hos_destroy(ctx), 'hos_destroy')
!
CONTAINS
!
SUBROUTINE blank()
   READ(unit_in, *)
END SUBROUTINE blank
SUBROUTINE rd_r(v)
   REAL(RP), INTENT(OUT) :: v
   CHARACTER(LEN=N_TOT) :: d
   READ(unit_in, '(A52,ES25.16)') d, v
END SUBROUTINE rd_r
SUBROUTINE rd_i(v)
   INTEGER, INTENT(OUT) :: v
   CHARACTER(LEN=N_TOT) :: d
   READ(unit_in, '(A52,I5)') d, v
END SUBROUTINE rd_i
SUBROUTINE rd_c(v)
   CHARACTER(LEN=*), INTENT(OUT) :: v
   CHARACTER(LEN=N_TOT) :: d
   READ(unit_in, '(A52,A)') d, v
END SUBROUTINE rd_c
!
! k with (2 pi f)^2 = k tanh k by bisection (linear_wave.f90:137-151, :427-466)
REAL(RP) FUNCTION wave_number(f)
   REAL(RP), INTENT(IN) :: f
   REAL(RP) :: lo, hi, x, w2, v
   w2 = (TWOPI * f)**2
   lo = 0.0_rp; hi = 1.0e6_rp; x = hi
   DO WHILE (ABS(lo - hi) > 1.0e-15_rp * ABS(hi))
      x = 0.5_rp * (lo + hi)
      v = x * TANH(x) - w2
      IF (v == 0.0_rp) EXIT
      IF (-w2 * v < 0.0_rp) THEN
         hi = x
      ELSE
         lo = x
      END IF
   END DO
   wave_number = x
END FUNCTION wave_number
!
! paddle transfer function (linear_wave.f90:277-334)
COMPLEX(RP) FUNCTION transfer(ig, k, d)
   INTEGER, INTENT(IN)  :: ig
   REAL(RP), INTENT(IN) :: k, d
   IF (ig == 1) THEN
      IF (k < 50.0_rp) THEN
         transfer = CMPLX(0.0_rp, -(2.0_rp * k + SINH(2.0_rp * k)) / (4.0_rp * SINH(k) * SINH(k)), RP)
      ELSE
         transfer = CMPLX(0.0_rp, -0.5_rp, RP)
      END IF
   ELSE
      IF (k < 50.0_rp .AND. k * d < 50.0_rp) THEN
         transfer = CMPLX(0.0_rp, -k * (1.0_rp - d) * (2.0_rp * k + SINH(2.0_rp * k)) / (4.0_rp * SINH(k)), RP) &
              / (k * (1.0_rp - d) * SINH(k) + COSH(k * d) - COSH(k))
      ELSE IF (k * d < 50.0_rp) THEN
         transfer = CMPLX(0.0_rp, -k * (1.0_rp - d), RP) / (2.0_rp * k * (1.0_rp - d) - 2.0_rp + 4.0_rp * COSH(k * d) * EXP(-k))
      ELSE IF (k * (1.0_rp - d) >= 20.0_rp) THEN
         transfer = CMPLX(0.0_rp, -0.5_rp, RP)
      ELSE
         transfer = CMPLX(0.0_rp, -k * (1.0_rp - d), RP) / (2.0_rp * k * (1.0_rp - d) - 2.0_rp + 2.0_rp * EXP(k * (d - 1.0_rp)))
      END IF
   END IF
END FUNCTION transfer
!
! trapezoidal volume and energy (IO/vol_energy.f90:36-119, :375-420)
SUBROUTINE volume_energy()
   REAL(RP) :: wy(n2), wx(n1), s, sp, sk
   INTEGER :: i, j
   wy = 1.0_rp; wx = 1.0_rp
   IF (n2 > 1) THEN
      wy(1) = 0.5_rp; wy(n2) = 0.5_rp
   END IF
   wx(1) = 0.5_rp; wx(n1) = 0.5_rp
   s = 0.0_rp; sp = 0.0_rp; sk = 0.0_rp
   DO j = 1, n2
      DO i = 1, n1
         s = s + wx(i) * wy(j) * eta(i,j)
         sp = sp + wx(i) * wy(j) * eta(i,j)**2
         sk = sk + wx(i) * wy(j) * (phis(i,j) + phi_add(i,j)) * deta(i,j)
      END DO
   END DO
   voltot = s * delx / xlen
   enepot = sp * delx / 2.0_rp
   enekin = sk * delx / 2.0_rp
   IF (n2 /= 1) THEN
      voltot = voltot / ylen * dely
      enepot = enepot * dely
      enekin = enekin * dely
   END IF
END SUBROUTINE volume_energy
!
END PROGRAM hos_nwt_gpu
