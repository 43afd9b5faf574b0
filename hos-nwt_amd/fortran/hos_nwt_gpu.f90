! Fortran driver of the GPU time-stepping core.
!
! Reads the reference's input_HOS-NWT.dat (same 52-column description + value
! format, IO/input.f90:62-133), nondimensionalises (main/HOS-NWT.f90:99-119),
! builds the paddle motion of a monochromatic wave (icase = 2, front wave or
! Dalrymple's directional wave, piston or hinged flap), of wave files
! (icase = 3 .cfg/.dat, icase = 31/32 Ocean .txt) or a sloshing mode
! (icase = 1), the front beach, the
! RK4 stability bound (main/HOS-NWT.f90:132-196), and runs the time loop of
! main/HOS-NWT.f90:446-581 with the inner adaptive loop on the GPU
! (hos_advance).  Writes the reference's output files (IO/output.f90:37-482)
! with its header echo (IO/input.f90:318-395): vol_energy.dat, and as the
! input flags ask 3d.dat (i3d), HOS_modes.dat (imodes), wmk_motion.dat (iwmk)
! and probes.dat (iprobes = 1, positions read from pro_file); SWENSE output and
! pressure probes are not built.  The reference's sizes are compile-time parameters
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
REAL(RP) :: t_adim, x_adim, alphat
COMPLEX(RP) :: tf, ctmp
INTEGER :: nprobes, nh, nc, rnum_f
! icase = 3 wave components (file2wave, wave/wave_def.f90:181-278): harmonic, frequency,
! amplitude, angle (rad), phase, wave number, transfer function; paddle of the file
INTEGER, ALLOCATABLE :: c_harmo(:)
REAL(RP), ALLOCATABLE :: c_f(:), c_a(:), c_ang(:), c_ph(:), c_k(:)
COMPLEX(RP), ALLOCATABLE :: c_tf(:)
REAL(RP) :: clock_f, depth_f, hinge_f, ly_f
CHARACTER(LEN=16) :: typ_wmk_f
REAL(RP), ALLOCATABLE :: xprobe(:), yprobe(:), eta_probe(:), etam(:,:), phism(:,:), xtop(:), vtop(:)
REAL(RP), ALLOCATABLE, TARGET :: nu(:,:), eta(:,:), phis(:,:), deta(:,:), phi_add(:,:), zero2(:,:), zadd(:,:)
REAL(RP), ALLOCATABLE, TARGET :: freq(:), pre(:), pim(:), dyre(:), dyim(:), nux(:,:)
! dipole wavemaker (igeom = 3): per component frequency, mu(:,1), mu(:,2), mu(:,3)
REAL(RP), ALLOCATABLE, TARGET :: dfr(:), dmxr(:), dmxi(:), dmzr(:), dmzi(:), dzd(:)
! dipole constants (IO/input.f90:137-139; x_dip / h at HOS-NWT.f90:111)
REAL(RP), PARAMETER :: z_dipmono = -0.25_rp, x_dip_si = 2.5_rp
INTEGER, PARAMETER :: n_side = 6
REAL(RP) :: x_dip
! SWENSE output (i_sw = 1): the modes of the last computation (zero before the first)
REAL(RP), ALLOCATABLE :: sw_x(:,:), sw_y(:,:), sw_z(:,:), sw_t(:,:), sw_fs(:,:), sw_fst(:,:), sw_add(:,:), &
     sw_addt(:,:)
INTEGER :: it_out
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
IF (.NOT. ANY(icase == [1, 2, 3, 31, 32, 4, 41])) &
   ERROR STOP 'this driver builds icase = 1, 2, 3, 31, 32, 4 and 41'
IF (ibat /= 2 .AND. ibat /= 3) ERROR STOP 'ibat = 2 (front) or 3 (Dalrymple)'
! adjustments made while reading (IO/input.f90:163-172)
islosh = islosh + 1
IF (n2 == 1) theta_mono = 0.0_rp
theta_mono = theta_mono * PI / 180.0_rp
IF (iramp == 0) Tramp = 0.0_rp
IF (i_wmk < 1 .OR. i_wmk > 3) ERROR STOP 'i_wmk = 1, 2 or 3 is built here'
IF (iprobes == 2) ERROR STOP 'pressure probes (iprobes = 2) are not built here'
IF (i_sw == 1 .AND. .NOT. ((igeom == 1 .OR. igeom == 2) .AND. icase /= 1)) &
   ERROR STOP 'the SWENSE output (i_sw = 1) is built for wavemaker tanks (igeom 1, 2)'
IF (i_sw == 1 .AND. nargs >= 5) THEN
   IF (prec /= 0) ERROR STOP 'the SWENSE output (i_sw = 1) runs in FP64 (precision 0)'
END IF
! output files, before the nondimensionalisation as in main/HOS-NWT.f90:92
CALL init_output()
!
! nondimensionalisation (HOS-NWT.f90:99-119)
Tscale = 1.0_rp / SQRT(G / h)
xlen = xlen / h; ylen = ylen / h; amp_mono = amp_mono / h; d_hinge = d_hinge / h; aslosh = aslosh / h
Tramp = Tramp / Tscale; T_stop = T_stop / Tscale; f_out = f_out * Tscale; nu_mono = nu_mono * Tscale
w_mono = nu_mono * TWOPI; T_mono = 1.0_rp / nu_mono
x_dip = x_dip_si / h
IF (n2 == 1) ylen = 1.0_rp
!
! RK4 stability bound (HOS-NWT.f90:132-196)
c8 = SQRT(8.0_rp / PI**2)
d1 = c8 * xlen / REAL(n1 - 1, RP)
d2 = d1 + 1.0_rp
IF (n2 /= 1) d2 = c8 * ylen / REAL(n2 - 1, RP)
d3 = c8 * 4.0_rp / REAL(2 * (n3 - 1), RP)
IF (icase == 1 .OR. igeom == 3) THEN
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
nh = 1
! Ocean .txt files become <file_name>.cfg/.dat first (initial_condition.f90:180-181)
IF (icase == 31 .OR. icase == 32) CALL ocean_txt2cfg_dat()
IF (icase == 4 .OR. icase == 41) THEN
   CALL spectrum2wave()
ELSE IF (icase == 3 .OR. icase == 31 .OR. icase == 32) THEN
   CALL file2wave()
   ! Dalrymple's method per component in a 3-D basin (initial_condition.f90:192-199)
   IF (ibat == 3 .AND. n2 /= 1) CALL dalrymple_expand(xd_mono)
ELSE IF (icase == 2 .AND. ibat == 3) THEN
   ! one harmonic on the clock 2 nu, rnum 1, Ly = ylen (initial_condition.f90:147-166)
   rnum_f = 1; clock_f = 2.0_rp * nu_mono; depth_f = h; ly_f = ylen; hinge_f = d_hinge
   typ_wmk_f = 'piston'
   IF (igeom == 2) typ_wmk_f = 'hinged'
   nc = 1
   ALLOCATE(c_harmo(1), c_f(1), c_a(1), c_ang(1), c_ph(1), c_k(1), c_tf(1))
   c_harmo = 1; c_f = nu_mono; c_a = amp_mono; c_ang = theta_mono; c_ph = ph_mono
   CALL dalrymple_expand(xd_mono)
END IF
ALLOCATE(freq(nh), pre(n2*nh), pim(n2*nh), dyre(n2*nh), dyim(n2*nh), nux(n1,n2))
freq = 0.0_rp; pre = 0.0_rp; pim = 0.0_rp; dyre = 0.0_rp; dyim = 0.0_rp; nux = 0.0_rp
! beach x derivative as initial_condition.f90:768 builds it (2 c x3), read with i_wmk = 2
IF (iabsnb /= 0 .AND. ABS(xabsf - 1.0_rp) > TINY) THEN
   DO i1 = 1, n1
      IF ((i1 - 1) * delx >= x0) nux(i1,:) = coeffabsf * 2.0_rp * ((i1 - 1) * delx - x0) / x2
   END DO
END IF
IF (igeom == 3 .AND. icase /= 1) THEN
   ! dipoles instead of a paddle (initial_condition.f90:168-170, :204-205)
   IF (icase == 2 .AND. ibat == 2) THEN
      nc = 1
      ALLOCATE(c_harmo(1), c_f(1), c_a(1), c_ang(1), c_ph(1), c_k(1), c_tf(1))
      c_harmo = 1; c_f = nu_mono; c_a = amp_mono; c_ang = theta_mono; c_ph = ph_mono
      c_k = wave_number(nu_mono)
   END IF
   k_w = c_k(1)
   CALL dipoles_ini()
ELSE IF (icase == 2 .AND. ibat == 2) THEN
   ! front(): one harmonic at nu_mono (wave_def.f90:54-80), wave_init and wmk_ini
   freq1 = nu_mono
   k_w = wave_number(freq1)
   tf = transfer(igeom, k_w, d_hinge)
   DO i2 = 1, n2
      ctmp = tf * COS(theta_mono) * amp_mono * &
             EXP(CMPLX(0.0_rp, -k_w * SIN(theta_mono) * (i2 - 1) * dely + ph_mono, RP))
      pre(i2) = REAL(ctmp)
      pim(i2) = AIMAG(ctmp)
      ! y derivative of the paddle amplitude (wavemaking.f90:147-149, no lateral ramp)
      ctmp = CMPLX(0.0_rp, -k_w * SIN(theta_mono), RP) * ctmp
      dyre(i2) = REAL(ctmp)
      dyim(i2) = AIMAG(ctmp)
   END DO
   freq(1) = freq1
ELSE IF (icase == 3 .OR. icase == 31 .OR. icase == 32 .OR. icase == 4 .OR. icase == 41 .OR. &
         (icase == 2 .AND. ibat == 3)) THEN
   CALL wmk_ini()
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
IF (icase == 3 .OR. icase == 31 .OR. icase == 32 .OR. icase == 4 .OR. icase == 41) p%n_freq = nh
p%freq = C_LOC(freq); p%pos_ini_re = C_LOC(pre); p%pos_ini_im = C_LOC(pim)
p%dposdy_ini_re = C_LOC(dyre); p%dposdy_ini_im = C_LOC(dyim); p%nu_x = C_LOC(nux)
IF (igeom == 3 .AND. icase /= 1) THEN
   p%n_freq = 0
   p%n_dip = INT(nc, C_INT)
   p%dip_freq = C_LOC(dfr); p%dip_mux_re = C_LOC(dmxr); p%dip_mux_im = C_LOC(dmxi)
   p%dip_muz_re = C_LOC(dmzr); p%dip_muz_im = C_LOC(dmzi); p%dip_z = C_LOC(dzd)
   p%x_dip = x_dip; p%n_side = n_side
END IF
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
x_adim = 1.0_rp
IF (idim == 1) THEN
   t_adim = Tscale
   x_adim = h
END IF
ALLOCATE(etam(n1,n2), phism(n1,n2), xtop(n2), vtop(n2))
ALLOCATE(sw_x(n1,n2), sw_y(n1,n2), sw_z(n1,n2), sw_t(n1,n2), sw_fs(n1,n2), sw_fst(n1,n2), &
         sw_add(2*(n3-1),n2), sw_addt(2*(n3-1),n2))
sw_x = 0.0_rp; sw_y = 0.0_rp; sw_z = 0.0_rp; sw_t = 0.0_rp; sw_fs = 0.0_rp; sw_fst = 0.0_rp
sw_add = 0.0_rp; sw_addt = 0.0_rp
it_out = 1  ! initialize sets it = 1, the time loop adds 1 before each output (HOS-NWT.f90:450)
!
time_cur = 0.0_rp
n_rk_tot = 0; n_er_tot = 0; n_acc = 0; n_rej = 0
DO WHILE (time_cur < T_stop + delt_out)
   n_rk_tot = n_rk_tot + n_acc
   n_er_tot = n_er_tot + n_rej
   it_out = it_out + 1
   CALL hos_check(hos_get_physical(ctx, eta, phis, deta, phi_add), 'hos_get_physical')
   CALL hos_check(hos_get_wmk(ctx, alpha, pos_top), 'hos_get_wmk')
   CALL volume_energy()
   IF (igeom == 2 .AND. icase /= 1) THEN
      WRITE(2,'(5(ES12.5,X))') time_cur * t_adim, voltot - 0.5_rp * alpha * (1 - d_hinge) * pos_top / xlen &
           - eta(1,1) * alpha * pos_top / xlen, enepot, enekin, enepot + enekin
   ELSE IF (icase /= 1 .AND. igeom /= 3) THEN
      WRITE(2,'(5(ES12.5,X))') time_cur * t_adim, voltot - alpha * pos_top / xlen - eta(1,1) * alpha * pos_top / xlen, &
           enepot, enekin, enepot + enekin
   ELSE
      WRITE(2,'(5(ES12.5,X))') time_cur * t_adim, voltot, enepot, enekin, enepot + enekin
   END IF
   CALL output_time_step()
   WRITE(*,'("    volume = ",ES12.4," t= ",ES11.4," timestep:",ES11.4,2(X,I5))') voltot, time_cur * t_adim, &
        delt * t_adim, n_rk_tot, n_er_tot
   IF (ABS(time_cur - T_stop) <= TINY) EXIT
   time_next = time_cur + delt_out
   IF (time_next > T_stop + delt_out) time_next = T_stop + delt_out
   h_rk = delt
   CALL hos_check(hos_advance(ctx, time_cur, time_next, h_rk, toler, dt_stab, delt_out, n_acc, n_rej), 'hos_advance')
   delt = h_rk
   ! SWENSE: the H2 operator after the attempt that ends on the output time
   ! (runge_kutta.f90:327-333), 10 levels
   IF (i_sw == 1) CALL hos_check(hos_swense_modes(ctx, time_cur, 10_C_INT, sw_x, sw_y, sw_z, sw_t, sw_fs, sw_fst, &
        sw_add, sw_addt), 'hos_swense_modes')
   ! check_slope (initial_condition.f90:687-716)
   CALL hos_check(hos_check_slope(ctx, 8.0_rp, 0.7_rp, fx, fy), 'hos_check_slope')
   IF (fx /= 0 .OR. fy /= 0) THEN
      WRITE(*,'(A)') '******************** Slope of FS elevation too high ******************** '
      WRITE(*,'(A)') 'Trying to filter eta and phis...'
   END IF
END DO
CLOSE(2)
IF (i3d == 1) CLOSE(1)
IF (imodes == 1) CLOSE(5)
IF (iwmk /= 0 .AND. igeom /= 3) CLOSE(8)
IF (iprobes == 1) CLOSE(9)
IF (i_sw == 1) CLOSE(123)
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
! header echo of the input (IO/input.f90:318-395, formats :457-470)
SUBROUTINE write_input(u)
   INTEGER, INTENT(IN) :: u
   CALL wb(u, '----------------------- Tank dimensions -----------------------')
   CALL wr(u, xlen, 'xlen', 'Length(m) of the wavetank')
   CALL wr(u, ylen, 'ylen', 'Beam(m) of the wavetank')
   CALL wr(u, h, 'h', 'Depth(m) of the wavetank')
   CALL wb(u, '------------------------ Select case --------------------------')
   CALL wi(u, icase, 'icase', 'Choice of computed case')
   CALL wb(u, '------------------ Sloshing case (icase = 1) ------------------')
   CALL wi(u, islosh, 'islosh', 'Number of the natural mode')
   CALL wr(u, aslosh, 'aslosh', 'Amplitude (m) of the mode')
   CALL wb(u, '--------------- Monochromatic case (icase = 2) ----------------')
   CALL wr(u, amp_mono, 'amp_mono', 'Amplitude (m) of wave train')
   CALL wr(u, nu_mono, 'nu_mono', 'Frequency (Hz) of wave train')
   CALL wr(u, theta_mono, 'theta_mono', 'Angle (deg) from x-axis')
   CALL wr(u, ph_mono, 'ph_mono', 'Phasis (rad) of wave train')
   CALL wi(u, ibat, 'ibat', 'Directional wmk type')
   CALL wr(u, xd_mono, 'xd_mono', 'Wave target distance (ibat=3)')
   CALL wb(u, '-------------------- File case (icase = 3) --------------------')
   CALL wc(u, file_name, 'file_name', 'Name of frequency file')
   CALL wi(u, i_cut, 'i_cut', 'Frequency cut-off')
   CALL wr(u, nuc_low, 'nuc_low', 'Low cut_off frequency (Hz)')
   CALL wr(u, nuc_high, 'nuc_high', 'High cut_off frequency (Hz) ')
   CALL wb(u, '------------------ Irregular wave (icase=4) -------------------')
   CALL wr(u, Hs, 'Hs', 'Significant wave height (m)')
   CALL wr(u, Tp, 'Tp', 'Peak period (s)')
   CALL wr(u, gamma, 'gamma', 'Shape factor (Jonswap)')
   CALL wi(u, iseed, 'iseed', 'Seed number for random numb. ')
   CALL wb(u, '-------------------- Wavemaker definition ---------------------')
   CALL wi(u, i_wmk, 'i_wmk', 'Nonlinear order of wavemaker')
   CALL wi(u, igeom, 'igeom', 'Type (1: piston, 2: hinged)')
   CALL wr(u, d_hinge, 'd_hinge', 'Rotation axis distance')
   CALL wi(u, iramp, 'iramp', 'Time ramp')
   CALL wr(u, Tramp, 'Tramp', 'Time ramp duration')
   CALL wb(u, '----------------------- Numerical beach -----------------------')
   CALL wi(u, iabsnb, 'iabsnb', 'Absorption numerical beach')
   CALL wr(u, xabsf, 'xabsf', 'Beginning front num. beach')
   CALL wr(u, coeffabsf, 'coeffabsf', 'Absorption strength front')
   CALL wb(u, '------------- Elevation/Velocity-Pressure probes --------------')
   CALL wi(u, iprobes, 'iprobes', 'Use of probes')
   CALL wc(u, pro_file, 'pro_file', 'Filename of probe positions')
   CALL wb(u, '---------------------- Time-integration -----------------------')
   CALL wr(u, T_stop, 'T_stop', 'Duration of the simulation')
   CALL wr(u, toler, 'toler', 'Time tolerance: RK 4(5)')
   CALL wr(u, f_out, 'f_out', 'Output frequency')
   CALL wb(u, '--------------------------- Output ----------------------------')
   CALL wi(u, idim, 'idim', 'Output: 1-dim. ; 0-nondim.')
   CALL wi(u, i3d, 'i3d', 'Free surface plot')
   CALL wi(u, imodes, 'imodes', 'Modes plot')
   CALL wi(u, iwmk, 'iwmk', 'Wavemaker motion plot')
   CALL wi(u, i_sw, 'i_sw', 'Swense output 1=''yes'',0=''no''')
   CALL wb(u, '--------------------- Numerical parameters ---------------------')
   CALL wi(u, n1, 'n1', 'Number of nodes/modes in x')
   CALL wi(u, n2, 'n2', 'Number of nodes/modes in y')
   CALL wi(u, n3, 'n3', 'Number of nodes/modes on wmk')
   CALL wi(u, mHOS, 'mHOS', 'HOS order')
   CALL wi(u, p%p1, 'p1', 'Dealiasing in x-direction')
   CALL wi(u, p%p2, 'p2', 'Dealiasing in y-direction')
   CALL wr(u, p%coeffilt(1), 'coeffilt(1)', 'Filtering ratio in x-modes')
   CALL wr(u, p%coeffilt(2), 'coeffilt(2)', 'Filtering ratio in y-modes')
   CALL wr(u, p%coeffilt(3), 'coeffilt(3)', 'Filtering ratio in z-modes')
END SUBROUTINE write_input
SUBROUTINE wb(u, text)
   INTEGER, INTENT(IN) :: u
   CHARACTER(LEN=*), INTENT(IN) :: text
   WRITE(u,'("#",A)') text
END SUBROUTINE wb
SUBROUTINE wi(u, v, var, text)
   INTEGER, INTENT(IN) :: u, v
   CHARACTER(LEN=*), INTENT(IN) :: var, text
   WRITE(u,'("#",A33,":: ",A13,":: ",I5)') text, var, v
END SUBROUTINE wi
SUBROUTINE wr(u, v, var, text)
   INTEGER, INTENT(IN) :: u
   REAL(RP), INTENT(IN) :: v
   CHARACTER(LEN=*), INTENT(IN) :: var, text
   WRITE(u,'("#",A33,":: ",A13,":: ",ES25.16)') text, var, v
END SUBROUTINE wr
SUBROUTINE wc(u, v, var, text)
   INTEGER, INTENT(IN) :: u
   CHARACTER(LEN=*), INTENT(IN) :: v, var, text
   WRITE(u,'("#",A33,":: ",A13,":: ",A)') text, var, v
END SUBROUTINE wc
!
! init_output (IO/output.f90:37-214): headers, and the probe positions
SUBROUTINE init_output()
   INTEGER :: j, ios
   REAL(RP) :: xr, yr
   CHARACTER(LEN=12) :: num
   CHARACTER(LEN=4000) :: line
   p%p1 = mHOS
   p%p2 = 1
   IF (n2 /= 1) p%p2 = mHOS
   CALL EXECUTE_COMMAND_LINE('mkdir -p Results')
   IF (i3d == 1) THEN
      OPEN(1, FILE='Results/3d.dat', STATUS='UNKNOWN')
      CALL write_input(1)
      WRITE(1,*) 'TITLE=" 3D free surface elevation "'
      IF (idim == 0) THEN
         WRITE(1,'(A)') 'VARIABLES="x/h","y/h","eta/h","phis*sqrt(h/g)/h**2"'
      ELSE
         WRITE(1,'(A)') 'VARIABLES="x","y","eta","phis"'
      END IF
   END IF
   OPEN(2, FILE='Results/vol_energy.dat', STATUS='UNKNOWN')
   CALL write_input(2)
   WRITE(2,'(A)') 'TITLE=" volume and energy evolution "'
   IF (idim == 0) THEN
      WRITE(2,'(A)') 'VARIABLES="time*sqrt(g/h)", "voltot", "enepot", "enekin", "enetot"'
   ELSE
      WRITE(2,'(A)') 'VARIABLES="time", "voltot", "enepot", "enekin", "enetot"'
   END IF
   IF (imodes == 1) THEN
      OPEN(5, FILE='Results/HOS_modes.dat', STATUS='UNKNOWN')
      CALL write_input(5)
      WRITE(5,'(A)') 'TITLE = " 3D spectral potential modes"'
      IF (idim == 0) THEN
         WRITE(5,'(15A)') 'VARIABLES = "kx*h", "ky*h","a_eta/h","a_phis*sqrt(h/g)/h**2","LOG10-a-eta","LOG10-a-phis"'
      ELSE
         WRITE(5,'(15A)') 'VARIABLES = "kx", "ky","a_eta","a_phis","LOG10-a-eta","LOG10-a-phis"'
      END IF
   END IF
   ! modes of the volumic fields for coupling (SWENSE), output.f90:219-228
   IF (i_sw == 1) THEN
      OPEN(123, FILE='Results/modes_HOS_SWENSE.dat', STATUS='REPLACE', FORM='FORMATTED', ACCESS='DIRECT', &
           RECL=18*(n1))
      IF (2 * (n3 - 1) > n1) ERROR STOP 'Change the value of n1 and/or n3 so that n1>=n3_add'
   END IF
   IF (iwmk /= 0 .AND. igeom /= 3) THEN
      OPEN(8, FILE='Results/wmk_motion.dat', STATUS='UNKNOWN')
      CALL write_input(8)
      IF (n2 == 1) THEN
         WRITE(8,*) 'TITLE="Wavemaker motion against time"'
         IF (idim == 0) THEN
            WRITE(8,'(a)') 'VARIABLES="time*sqrt(g/h)", "X-top/h", "Vx-top/sqrt(g*h)"'
         ELSE
            WRITE(8,'(a)') 'VARIABLES="time", "X-top", "Vx-top"'
         END IF
      ELSE
         WRITE(8,*) 'TITLE="Wavemaker motion against time and transverse direction"'
         IF (idim == 0) THEN
            WRITE(8,'(a)') 'VARIABLES="y/h", "X-top/h", "Vx-top/sqrt(g*h)"'
         ELSE
            WRITE(8,'(a)') 'VARIABLES="y", "X-top", "Vx-top"'
         END IF
      END IF
   END IF
   nprobes = 0
   IF (iprobes == 1) THEN
      OPEN(55, FILE=TRIM(pro_file), STATUS='OLD')
      DO
         READ(55, *, IOSTAT=ios)
         IF (ios /= 0) EXIT
         nprobes = nprobes + 1
      END DO
      REWIND(55)
      IF (nprobes > 60) ERROR STOP 'at most 60 wave probes (output format of probes.dat)'
      ALLOCATE(xprobe(nprobes), yprobe(nprobes), eta_probe(nprobes))
      DO j = 1, nprobes
         IF (n2 == 1) THEN
            READ(55, *) xr
            yr = 0.0_rp
         ELSE
            READ(55, *) xr, yr
         END IF
         xprobe(j) = xr / h
         yprobe(j) = yr / h
      END DO
      CLOSE(55)
      OPEN(9, FILE='Results/probes.dat', STATUS='UNKNOWN')
      CALL write_input(9)
      WRITE(9,'(A)') 'TITLE="Probes elevation records versus time"'
      IF (idim == 0) THEN
         line = 'VARIABLES = "time*sqrt(g/h)" '
      ELSE
         line = 'VARIABLES = "time" '
      END IF
      DO j = 1, nprobes
         WRITE(num,'(I0)') j
         IF (idim == 0) THEN
            line = TRIM(line) // ' "p' // TRIM(num) // '/h"'
         ELSE
            line = TRIM(line) // ' "p' // TRIM(num) // '"'
         END IF
      END DO
      WRITE(9,'(A)') TRIM(line) // ' '
   ELSE
      ALLOCATE(xprobe(1), yprobe(1), eta_probe(1))
   END IF
END SUBROUTINE init_output
!
! output_time_step (IO/output.f90:232-482) for the files opened by init_output
! (vol_energy.dat is written in the main loop); tecplot = 11 zone headers
SUBROUTINE output_time_step()
   REAL(RP) :: eta_target, xx, yy
   INTEGER :: j
   IF (iprobes == 1) THEN
      CALL hos_check(hos_probes(ctx, INT(nprobes, C_INT), xprobe, yprobe, eta_probe), 'hos_probes')
      WRITE(9,'(61d13.5)') time_cur * t_adim, (eta_probe(j) * x_adim, j = 1, nprobes)
   END IF
   IF (i3d == 1) THEN
      IF (ABS(time_cur) < TINY) THEN
         WRITE(1,'(A,F6.2,A,I4,A,I4)') 'ZONE SOLUTIONTIME = ', time_cur * t_adim, ', I=', n1, ', J=', n2
         DO i2 = 1, n2
            DO i1 = 1, n1
               WRITE(1,'(2(F10.5,X),2(ES12.5,X))') (i1 - 1) * delx * x_adim, (i2 - 1) * dely * x_adim, &
                    eta(i1,i2) * x_adim, phis(i1,i2) * x_adim**2 / t_adim
            END DO
         END DO
      ELSE
         WRITE(1,'(A,F6.2,A,I4,A,I4)') 'ZONE SOLUTIONTIME = ', time_cur * t_adim, ', D=(1,2), I=', n1, ', J=', n2
         DO i2 = 1, n2
            DO i1 = 1, n1
               eta_target = 0.0_rp
               IF (icase == 2) THEN
                  xx = (i1 - 1) * delx
                  yy = (i2 - 1) * dely
                  IF (igeom == 3) xx = xx - x_dip  ! (output.f90:330-331)
                  eta_target = amp_mono * COS(w_mono * time_cur - k_w * (xx * COS(theta_mono) + yy * SIN(theta_mono)) &
                       + ph_mono)
               END IF
               WRITE(1,'(2(ES12.5,X))') eta(i1,i2) * x_adim, (eta(i1,i2) - eta_target) * x_adim
            END DO
         END DO
      END IF
   END IF
   IF (iwmk == 1 .AND. igeom /= 3) THEN
      CALL hos_check(hos_get_wmk_motion(ctx, alpha, alphat, xtop, vtop), 'hos_get_wmk_motion')
      IF (n2 == 1) THEN
         WRITE(8,'(3d13.5)') time_cur * t_adim, xtop(1) * x_adim, vtop(1) * x_adim / t_adim
      ELSE
         WRITE(8,'(A,F8.2,A,I3)') 'ZONE SOLUTIONTIME = ', time_cur * t_adim, ', I=', n2
         DO i2 = 1, n2
            WRITE(8,'(3d13.5)') (i2 - 1) * dely * x_adim, xtop(i2) * x_adim, vtop(i2) * x_adim / t_adim
         END DO
      END IF
   END IF
   IF (imodes == 1) THEN
      CALL hos_check(hos_get_modes(ctx, etam, phism), 'hos_get_modes')
      etam = etam * x_adim
      phism = phism * x_adim**2 / t_adim
      IF (ABS(time_cur) < TINY) THEN
         WRITE(5,'(A,F6.2,A,I4,A,I4)') 'ZONE SOLUTIONTIME = ', time_cur * t_adim, ', I=', n1, ', J=', n2
         DO i2 = 1, n2
            DO i1 = 1, n1
               xx = (i1 - 1) * PI / xlen
               yy = 0.0_rp
               IF (n2 /= 1) yy = (i2 - 1) * PI / ylen
               WRITE(5,'(6(ES12.5,X))') xx / x_adim, yy / x_adim, etam(i1,i2), phism(i1,i2), &
                    LOG10(MAX(TINY, ABS(etam(i1,i2)))), LOG10(MAX(TINY, ABS(phism(i1,i2))))
            END DO
         END DO
      ELSE
         WRITE(5,'(A,F6.2,A,I4,A,I4)') 'ZONE SOLUTIONTIME = ', time_cur * t_adim, ', D=(1,2), I=', n1, ', J=', n2
         DO i2 = 1, n2
            DO i1 = 1, n1
               WRITE(5,'(4(ES12.5,X))') etam(i1,i2), phism(i1,i2), LOG10(MAX(TINY, ABS(etam(i1,i2)))), &
                    LOG10(MAX(TINY, ABS(phism(i1,i2))))
            END DO
         END DO
      END IF
   END IF
   ! spectral modes for SWENSE (output.f90:417-467): a header record at t = 0, then n2 x 8
   ! records per output time (it_out is the reference's it)
   IF (i_sw == 1) THEN
      IF (ABS(time_cur) < TINY) THEN
         WRITE(123,'(5000(ES17.10,1X))',REC=1) 1.0_rp*n1, 1.0_rp*2*(n3-1), 1.0_rp*n2, 1.0_rp/f_out, T_stop, &
              xlen, ylen, h, (sw_x(i1,1), i1=9,n1)
         WRITE(123,'(5000(ES17.10,1X))',REC=2) (sw_y(i1,1), i1=1,n1)
         WRITE(123,'(5000(ES17.10,1X))',REC=3) (sw_z(i1,1), i1=1,n1)
         WRITE(123,'(5000(ES17.10,1X))',REC=4) (sw_t(i1,1), i1=1,n1)
         WRITE(123,'(5000(ES17.10,1X))',REC=5) (sw_fs(i1,1), i1=1,n1)
         WRITE(123,'(5000(ES17.10,1X))',REC=6) (sw_fst(i1,1), i1=1,n1)
         WRITE(123,'(5000(ES17.10,1X))',REC=7) (sw_add(j,1), j=1,2*(n3-1))
         WRITE(123,'(5000(ES17.10,1X))',REC=8) (sw_addt(j,1), j=1,2*(n3-1))
         DO i2 = 2, n2
            CALL sw_block(8 * (i2 - 1), i2)
         END DO
      END IF
      DO i2 = 1, n2
         CALL sw_block((it_out - 1) * n2 * 8 + 8 * (i2 - 1), i2)
      END DO
   END IF
END SUBROUTINE output_time_step
!
SUBROUTINE sw_block(r0, i2b)
   INTEGER, INTENT(IN) :: r0, i2b
   INTEGER :: j
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+1) (sw_x(j,i2b), j=1,n1)
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+2) (sw_y(j,i2b), j=1,n1)
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+3) (sw_z(j,i2b), j=1,n1)
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+4) (sw_t(j,i2b), j=1,n1)
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+5) (sw_fs(j,i2b), j=1,n1)
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+6) (sw_fst(j,i2b), j=1,n1)
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+7) (sw_add(j,i2b), j=1,2*(n3-1))
   WRITE(123,'(5000(ES17.10,1X))',REC=r0+8) (sw_addt(j,i2b), j=1,2*(n3-1))
END SUBROUTINE sw_block
!
! file2wave (wave/wave_def.f90:181-278, with wave_init :282-329): <file_name>.cfg and
! .dat, components between the cut-off frequencies, nondimensionalised by the file's depth
SUBROUTINE file2wave()
   INTEGER :: u, ios, nf, j, n_start, n_stop, ig
   REAL(RP) :: ramp_f, lo, hi, ly_f, tsc
   CHARACTER(LEN=16) :: typ_ramp_f
   INTEGER, ALLOCATABLE :: hh(:)
   REAL(RP), ALLOCATABLE :: ff(:), aa(:), an(:), ph(:)
   u = 999
   OPEN(u, FILE=TRIM(file_name)//'.cfg', STATUS='OLD')
   READ(u,'(I3)') rnum_f
   READ(u,*) clock_f
   READ(u,*) depth_f
   READ(u,*) ramp_f
   READ(u,*) typ_ramp_f
   READ(u,*) lo
   READ(u,*) hi
   READ(u,*) ly_f
   READ(u,*) typ_wmk_f
   READ(u,*) hinge_f
   CLOSE(u)
   tsc = SQRT(depth_f / G)
   clock_f = clock_f * tsc
   lo = lo * tsc
   hi = hi * tsc
   hinge_f = hinge_f / depth_f
   ly_f = ly_f / depth_f
   OPEN(u, FILE=TRIM(file_name)//'.dat', STATUS='OLD')
   nf = 0
   DO
      READ(u, *, IOSTAT=ios)
      IF (ios /= 0) EXIT
      nf = nf + 1
   END DO
   REWIND(u)
   ALLOCATE(hh(nf), ff(nf), aa(nf), an(nf), ph(nf))
   n_start = 1
   n_stop = 0
   DO j = 1, nf
      READ(u,*) hh(j), ff(j), aa(j), an(j), ph(j)
      IF (ff(j) <= hi / tsc) n_stop = j
      IF (ff(j) < lo / tsc) n_start = n_start + 1
   END DO
   CLOSE(u)
   nc = n_stop - n_start + 1
   IF (nc <= 0) ERROR STOP 'no wave component between the cut-off frequencies'
   ALLOCATE(c_harmo(nc), c_f(nc), c_a(nc), c_ang(nc), c_ph(nc), c_k(nc), c_tf(nc))
   c_harmo = hh(n_start:n_stop)
   c_f = ff(n_start:n_stop) * tsc
   c_a = aa(n_start:n_stop) / depth_f
   c_ang = an(n_start:n_stop) * PI / 180.0_rp
   c_ph = ph(n_start:n_stop)
   ig = 1
   IF (TRIM(typ_wmk_f) == 'hinged') ig = 2
   DO j = 1, nc
      c_k(j) = wave_number(c_f(j))
      c_tf(j) = CMPLX(1.0_rp, 0.0_rp, RP)
      IF (ABS(c_k(j)) > TINY) c_tf(j) = transfer(ig, c_k(j), hinge_f)
      IF (c_f(j) < 1.0E-8_rp) THEN
         c_k(j) = 0.0_rp
         c_tf(j) = CMPLX(1.0_rp, 0.0_rp, RP)
      END IF
   END DO
   ! number of distinct harmonics up to 4 n_max (wavemaking.f90:104-118)
   nh = 0
   DO j = 0, 4 * (2**rnum_f / MAX(1, FLOOR(clock_f * SQRT(G / depth_f))))
      IF (ANY(c_harmo == j)) nh = nh + 1
   END DO
   IF (nh == 0) ERROR STOP 'no wave harmonic below the maximum frequency'
END SUBROUTINE file2wave
!
! Long-crested irregular waves of a JONSWAP (icase 4) or Bretschneider (icase 41) spectrum
! (initialize_wav_cfg, main/initial_condition.f90:212-308): the components the reference
! builds in memory instead of reading a .cfg / .dat pair.  Clock 32 Hz, rnum =
! CEILING(log2(T_stop clock)) with T_stop in seconds; component i = 1 .. NINT(2 2^rnum /
! clock + 1) has harmonic i - 1 and frequency i f_base (f_base = clock / 2^rnum), as there;
! S(f) df with sigma = 0.07 below 1/Tp and 0.09 above (single-precision literals, as there)
! and amplitudes sqrt(2 S df) rescaled so that 4 sqrt(sum S df) = Hs.  Phases: 2 pi u with
! u = (x - 1) / (2^31 - 2) from the Park-Miller minimal-standard sequence x <- 16807 x
! mod (2^31 - 1), x_0 = iseed (1 when iseed mod (2^31 - 1) is 0).  The reference draws them
! with gfortran's SRAND / RAND, which this build does not run; hosnwt/config.py
! (spectrum_components) draws the same sequence as this driver.
SUBROUTINE spectrum2wave()
   INTEGER :: n_pts, j, ig
   INTEGER(8) :: x
   REAL(RP) :: f_base, sigma, hs_int, s
   clock_f = 32.0_rp
   rnum_f = CEILING(LOG10(T_stop * Tscale * clock_f) / LOG10(2.0_rp))
   depth_f = h
   ly_f = ylen * h
   hinge_f = d_hinge * h
   typ_wmk_f = 'piston'
   IF (igeom == 2) typ_wmk_f = 'hinged'
   n_pts = NINT(2 * 2**rnum_f / clock_f + 1)
   f_base = clock_f / 2.0_rp**rnum_f
   nc = n_pts
   ALLOCATE(c_harmo(nc), c_f(nc), c_a(nc), c_ang(nc), c_ph(nc), c_k(nc), c_tf(nc))
   x = MODULO(INT(iseed, 8), 2147483647_8)
   IF (x == 0) x = 1
   hs_int = 0.0_rp
   DO j = 1, nc
      c_harmo(j) = j - 1
      c_f(j) = j * f_base
      c_ang(j) = 0.0_rp
      x = MODULO(16807_8 * x, 2147483647_8)
      c_ph(j) = TWOPI * REAL(x - 1, RP) / 2147483646.0_rp
      IF (c_f(j) < 1.0_rp / Tp) THEN
         sigma = 0.07
      ELSE
         sigma = 0.09
      END IF
      IF (icase == 4) THEN
         s = 1.0_rp / (TWOPI**4) * G**2 / c_f(j)**5 * EXP(-5.0_rp / 4.0_rp * (c_f(j) * Tp)**(-4)) &
             * gamma**(EXP((-(c_f(j) - 1.0_rp / Tp)**2) / (2.0_rp * (sigma / Tp)**2))) * f_base
      ELSE
         s = 5.0_rp / 16.0_rp / (Tp**4 * c_f(j)**5) * Hs * EXP(-5.0_rp / 4.0_rp * (c_f(j) * Tp)**(-4)) * f_base
      END IF
      hs_int = hs_int + s
      c_a(j) = SQRT(2.0_rp * s)
   END DO
   hs_int = 4.0_rp * SQRT(hs_int)
   WRITE(*,'(A,ES12.5,A,ES12.5,A,ES12.5)') 'Hs target =', Hs, ' Tp =', Tp, ' Hs_calc =', hs_int
   c_a = c_a * Hs / hs_int
   ! nondimensional (initial_condition.f90:288-299)
   c_f = c_f * Tscale
   c_a = c_a / h
   clock_f = clock_f * Tscale
   ly_f = ly_f / h
   hinge_f = hinge_f / h
   ig = 1
   IF (TRIM(typ_wmk_f) == 'hinged') ig = 2
   DO j = 1, nc
      c_k(j) = wave_number(c_f(j))
      c_tf(j) = CMPLX(1.0_rp, 0.0_rp, RP)
      IF (ABS(c_k(j)) > TINY) c_tf(j) = transfer(ig, c_k(j), hinge_f)
      IF (c_f(j) < 1.0E-8_rp) THEN
         c_k(j) = 0.0_rp
         c_tf(j) = CMPLX(1.0_rp, 0.0_rp, RP)
      END IF
   END DO
   nh = 0
   DO j = 0, 4 * (2**rnum_f / MAX(1, FLOOR(clock_f * SQRT(G / depth_f))))
      IF (ANY(c_harmo == j)) nh = nh + 1
   END DO
END SUBROUTINE spectrum2wave
!
! Ocean .txt (ocean_txt2cfg_dat, wave/read_ocean_txt.f90:33-228): the run's rnumber, ramp,
! optional rnumber compression and its list of fronts (harmonic, frequency, amplitude,
! angle and phase in rad) become <file_name>.cfg/.dat in the formats file2wave reads.
! Frequencies are harmonic * 32 Hz / 2^rnum; in the ECN basin (icase 31) the phases are
! moved to the basin's centre line, k(f, 5 m) * 29.74 m / 2 * sin(angle).
SUBROUTINE ocean_txt2cfg_dat()
   INTEGER :: u, ios, nl, j, rnum, ramp, rcomp, nfr, first
   LOGICAL :: comp
   REAL(RP) :: clk, fb, lo, hi, depth, kd
   CHARACTER(LEN=200) :: ln
   INTEGER, ALLOCATABLE :: hh(:)
   REAL(RP), ALLOCATABLE :: ff(:), aa(:), an(:), ph(:)
   u = 998
   OPEN(u, FILE=TRIM(file_name)//'.txt', STATUS='OLD')
   nl = 0; comp = .FALSE.; rnum = -1; ramp = 0; rcomp = 1; nfr = -1; first = 0
   DO
      READ(u, '(A)', IOSTAT=ios) ln
      IF (ios /= 0) EXIT
      nl = nl + 1
      IF (INDEX(ln, 'rnumber  =') /= 0) rnum = int_ending_at(ln, 18)
      IF (INDEX(ln, 'ramp     =') /= 0) ramp = int_ending_at(ln, 18)
      IF (INDEX(ln, 'rnumber compression') /= 0) THEN
         comp = .TRUE.
         rcomp = int_ending_at(ln, 62)
      END IF
      IF (INDEX(ln, 'Wave has ') /= 0) THEN
         nfr = int_starting_at(ln, 10)
         first = nl + 2
      END IF
   END DO
   IF (rnum < 0 .OR. nfr < 0) ERROR STOP 'Ocean .txt file: no rnumber or no "Wave has" line'
   REWIND(u)
   DO j = 1, first - 1
      READ(u, '(A)')
   END DO
   ALLOCATE(hh(nfr), ff(nfr), aa(nfr), an(nfr), ph(nfr))
   DO j = 1, nfr
      READ(u, *) hh(j), ff(j), aa(j), an(j), ph(j)
   END DO
   CLOSE(u)
   IF (comp) hh = hh * rcomp
   clk = 32.0_rp
   fb = clk / 2.0_rp**rnum
   ff = hh * fb
   depth = 5.0_rp
   IF (icase == 32) depth = 2.88_rp
   IF (icase == 31) THEN
      DO j = 1, nfr
         kd = wave_number(ff(j) * SQRT(depth / G)) / depth
         ph(j) = ph(j) + kd * 29.74_rp / 2.0_rp * SIN(an(j))
      END DO
   END IF
   an = 180.0_rp * an / PI
   lo = 0.0_rp; hi = 100.0_rp
   IF (i_cut == 1) THEN
      lo = nuc_low; hi = nuc_high
   END IF
   OPEN(u, FILE=TRIM(file_name)//'.cfg', STATUS='REPLACE')
   WRITE(u,'(I2)') rnum
   WRITE(u,'(F5.2)') clk
   WRITE(u,'(F5.2)') depth
   WRITE(u,'(I4)') ramp
   WRITE(u,'(A)') 'lin'
   WRITE(u,'(F5.2)') lo
   WRITE(u,'(F6.2)') hi
   IF (icase == 31) THEN
      WRITE(u,'(F5.2)') 30.0_rp
      WRITE(u,'(A)') 'hinged'
      WRITE(u,'(F6.3)') 2.147_rp
   ELSE
      WRITE(u,'(F5.2)') 5.0_rp
      WRITE(u,'(A)') 'hinged'
      WRITE(u,'(F6.3)') 0.0_rp
   END IF
   CLOSE(u)
   WHERE (aa < 1.0e-98_rp) aa = 0.0_rp
   OPEN(u, FILE=TRIM(file_name)//'.dat', STATUS='REPLACE')
   DO j = 1, nfr
      IF (ff(j) >= lo .AND. ff(j) <= hi) WRITE(u,'(I5,4(X,G20.7))') hh(j), ff(j), aa(j), an(j), ph(j)
   END DO
   CLOSE(u)
END SUBROUTINE ocean_txt2cfg_dat
!
! the integer whose last digit is at column c (read leftwards to a blank)
INTEGER FUNCTION int_ending_at(ln, c)
   CHARACTER(LEN=*), INTENT(IN) :: ln
   INTEGER, INTENT(IN) :: c
   INTEGER :: b
   b = c
   DO WHILE (b > 1 .AND. ln(b:b) /= ' ')
      b = b - 1
   END DO
   READ(ln(b:c), *) int_ending_at
END FUNCTION int_ending_at
!
! the integer whose first digit is at column c (read rightwards to a blank)
INTEGER FUNCTION int_starting_at(ln, c)
   CHARACTER(LEN=*), INTENT(IN) :: ln
   INTEGER, INTENT(IN) :: c
   INTEGER :: e
   e = c
   DO WHILE (e < LEN(ln) .AND. ln(e:e) /= ' ')
      e = e + 1
   END DO
   READ(ln(c:e), *) int_starting_at
END FUNCTION int_starting_at
!
! Dalrymple's directional wavemaker (dalrymple, wave/wave_def.f90:111-177): each component
! (harmonic m, amplitude a, angle theta, phase phi) becomes 2(N+1) components of frequency
! f = m clock / 2^rnum, N = floor(k Ly / pi), at the angles +-asin(mu_n / k), mu_n = n pi / Ly,
! with complex amplitude a e^{i phi} I_n e^{(i sqrt(k^2 - mu_n^2) - i k cos theta) x_d} / 2
! (I_n the projection of e^{i k sin(theta) y} on cos(mu_n y)), so that the wave focuses at
! the distance x_d from the wavemaker.  The phase is taken with the reference's phase()
! (ref_phase).  Then wave_init: k and the transfer function.
SUBROUTINE dalrymple_expand(xd)
   REAL(RP), INTENT(IN) :: xd
   INTEGER :: j, n, nn, m, ntot, ig
   REAL(RP) :: fr, kk, beta, mu
   COMPLEX(RP) :: tmp, In, kn, a
   INTEGER, ALLOCATABLE :: hh(:)
   REAL(RP), ALLOCATABLE :: ff(:), aa(:), an(:), ph(:)
   COMPLEX(RP), PARAMETER :: CI = (0.0_rp, 1.0_rp)
   ntot = 0
   DO j = 1, nc
      fr = c_harmo(j) * clock_f / 2.0_rp**rnum_f
      ntot = ntot + 2 * (FLOOR(wave_number(fr) * ly_f / PI) + 1)
   END DO
   ALLOCATE(hh(ntot), ff(ntot), aa(ntot), an(ntot), ph(ntot))
   m = 0
   DO j = 1, nc
      fr = c_harmo(j) * clock_f / 2.0_rp**rnum_f
      kk = wave_number(fr)
      nn = FLOOR(kk * ly_f / PI)
      beta = kk * SIN(c_ang(j))
      tmp = -EXP(-CI * beta * ly_f)
      DO n = 0, nn
         mu = n * PI / ly_f
         tmp = -tmp
         IF (ABS(c_ang(j)) < TINY) THEN
            In = 0.0_rp
            IF (n == 0) In = 1.0_rp
         ELSE IF (ABS(mu - ABS(beta)) < TINY) THEN
            In = 1.0_rp
         ELSE
            In = 2.0_rp * CI * beta * (tmp - 1.0_rp) / (beta**2 - mu**2) / (MERGE(2.0_rp, 1.0_rp, n == 0) * ly_f)
         END IF
         kn = CI * SQRT(kk**2 - mu**2)
         a = c_a(j) * EXP(CI * c_ph(j)) * In * EXP((kn - CI * kk * COS(c_ang(j))) * xd)
         hh(m + n + 1) = c_harmo(j); hh(m + nn + n + 2) = c_harmo(j)
         ff(m + n + 1) = fr; ff(m + nn + n + 2) = fr
         aa(m + n + 1) = ABS(a / 2.0_rp); aa(m + nn + n + 2) = ABS(a / 2.0_rp)
         an(m + n + 1) = ASIN(mu / kk); an(m + nn + n + 2) = -ASIN(mu / kk)
         ph(m + n + 1) = ref_phase(a); ph(m + nn + n + 2) = ref_phase(a)
      END DO
      m = m + 2 * (nn + 1)
   END DO
   DEALLOCATE(c_harmo, c_f, c_a, c_ang, c_ph, c_k, c_tf)
   nc = ntot
   ALLOCATE(c_harmo(nc), c_f(nc), c_a(nc), c_ang(nc), c_ph(nc), c_k(nc), c_tf(nc))
   c_harmo = hh; c_f = ff; c_a = aa; c_ang = an; c_ph = ph
   ig = 1
   IF (TRIM(typ_wmk_f) == 'hinged') ig = 2
   DO j = 1, nc
      c_k(j) = wave_number(c_f(j))
      c_tf(j) = CMPLX(1.0_rp, 0.0_rp, RP)
      IF (ABS(c_k(j)) > TINY) c_tf(j) = transfer(ig, c_k(j), hinge_f)
      IF (c_f(j) < 1.0E-8_rp) THEN
         c_k(j) = 0.0_rp
         c_tf(j) = CMPLX(1.0_rp, 0.0_rp, RP)
      END IF
   END DO
END SUBROUTINE dalrymple_expand
!
! phase() of wave/wave_def.f90:491-524 as written: 0 whenever |z| > tiny, the argument in
! [0, 2 pi) only below (0 here for z = 0, where the reference divides 0 by 0)
REAL(RP) FUNCTION ref_phase(z)
   COMPLEX(RP), INTENT(IN) :: z
   COMPLEX(RP) :: zt
   REAL(RP) :: pt
   ref_phase = 0.0_rp
   IF (ABS(z) > TINY .OR. ABS(z) == 0.0_rp) RETURN
   zt = z / ABS(z)
   pt = ASIN(AIMAG(zt))
   IF (REAL(zt) >= 0.0_rp) THEN
      ref_phase = pt
      IF (AIMAG(zt) < 0.0_rp) ref_phase = TWOPI + pt
   ELSE
      ref_phase = PI - pt
   END IF
END FUNCTION ref_phase
!
! wmk_ini (wavemaking.f90:119-155): per harmonic, the frequency of its first component and
! pos_ini(y) = sum TF cos(angle) a exp(i(-k sin(angle) y + phase)), and its y derivative
SUBROUTINE wmk_ini()
   INTEGER :: hm, j, n, jf
   COMPLEX(RP) :: term
   n = 0
   DO hm = 0, 4 * (2**rnum_f / MAX(1, FLOOR(clock_f * SQRT(G / depth_f))))
      IF (.NOT. ANY(c_harmo == hm)) CYCLE
      n = n + 1
      jf = 0
      DO j = 1, nc
         IF (c_harmo(j) /= hm) CYCLE
         IF (jf == 0) jf = j
         DO i2 = 1, n2
            term = c_tf(j) * COS(c_ang(j)) * c_a(j) * &
                   EXP(CMPLX(0.0_rp, -c_k(j) * SIN(c_ang(j)) * (i2 - 1) * dely + c_ph(j), RP))
            pre(i2 + n2 * (n - 1)) = pre(i2 + n2 * (n - 1)) + REAL(term)
            pim(i2 + n2 * (n - 1)) = pim(i2 + n2 * (n - 1)) + AIMAG(term)
            term = CMPLX(0.0_rp, -c_k(j) * SIN(c_ang(j)), RP) * term
            dyre(i2 + n2 * (n - 1)) = dyre(i2 + n2 * (n - 1)) + REAL(term)
            dyim(i2 + n2 * (n - 1)) = dyim(i2 + n2 * (n - 1)) + AIMAG(term)
         END DO
      END DO
      freq(n) = c_f(jf)
   END DO
END SUBROUTINE wmk_ini
!
! dipoles_ini (main/resol_wmkr.f90:1404-1455): per wave component the dipole depth from its
! frequency, the dipole transfer function there, mu(:,2) = a TF exp(i(k x_dip + phase)),
! mu(:,1) = -i mu(:,2) tanh(k (z + 1)); the wavenumbers are the dispersion relation's
SUBROUTINE dipoles_ini()
   INTEGER :: j
   REAL(RP) :: z, k, k2, w2, w4
   COMPLEX(RP) :: tfd, mz
   ALLOCATE(dfr(nc), dmxr(nc), dmxi(nc), dmzr(nc), dmzi(nc), dzd(nc))
   DO j = 1, nc
      IF (c_f(j) > 1.0_rp) THEN
         z = -0.35_rp
      ELSE IF (c_f(j) > 0.6_rp) THEN
         z = -0.2_rp
      ELSE IF (c_f(j) > 0.5_rp) THEN
         z = -0.4_rp
      ELSE
         z = -0.6_rp
      END IF
      IF (icase == 2) z = z_dipmono
      k = c_k(j)
      k2 = k * k
      w2 = k * TANH(k)
      ! TF_adim_new, 'dipole' (wave/linear_wave.f90:316-327)
      IF (k <= 20.0_rp) THEN
         w4 = w2 * w2
         tfd = CMPLX((k2 - w4 + w2) / (2.0_rp * SQRT(w2) * (k2 - w4) * COSH(k) * SINH(k * (z + 1.0_rp))), 0.0_rp, RP)
      ELSE
         tfd = CMPLX(SQRT(w2) / (2.0_rp * EXP(k * z) * k2), 0.0_rp, RP)
      END IF
      mz = c_a(j) * tfd * EXP(CMPLX(0.0_rp, k * x_dip + c_ph(j), RP))
      dfr(j) = c_f(j)
      dmzr(j) = REAL(mz); dmzi(j) = AIMAG(mz)
      mz = CMPLX(0.0_rp, -1.0_rp, RP) * mz * TANH(k * (z + 1.0_rp))
      dmxr(j) = REAL(mz); dmxi(j) = AIMAG(mz)
      dzd(j) = z
   END DO
END SUBROUTINE dipoles_ini
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
