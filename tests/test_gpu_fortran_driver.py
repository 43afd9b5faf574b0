"""The Fortran driver (iso_c_binding into libhosnwt) reproduces the oracle's
output files for a reference-format input file.

Results/vol_energy.dat, probes.dat, wmk_motion.dat, 3d.dat and HOS_modes.dat
carry 5 or 6 significant digits (output.f90:201, :109, :801, :101-102, :501-502),
so the values are compared at 2e-4 of each column's scale; the files' layout
(header echo of the input, TITLE / VARIABLES lines, tecplot zone lines) is
checked line by line.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import make_setup, oracle_params
from hosnwt import config as hc
from oracle import hos_oracle as ho

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hos-nwt_amd", "lib", "hos_nwt_gpu")

INPUT = """----------------------- Tank dimensions -----------------------
Length(m) of the wavetank    :: xlen             :: 20.d0
Beam(m) of the wavetank      :: ylen             :: 1.d0
Depth(m) of the wavetank     :: h                :: 1.d0
------------------------ Select case --------------------------
Choice of computed case      :: icase            :: {icase}
------------------ Sloshing case (icase = 1) ------------------
Number of the natural mode   :: islosh           :: 3
Amplitude (m) of the mode    :: aslosh           :: 0.02d0
--------------- Monochromatic case (icase = 2) ----------------
Amplitude (m) of wave train  :: amp_mono         :: {amp}
Frequency (Hz) of wave train :: nu_mono          :: {nu}
Angle (deg) from x-axis      :: theta_mono       :: 0.d0
Phasis (rad) of wave train   :: ph_mono          :: 0.d0
Directional wmk type         :: ibat             :: 2
Wave target distance (ibat=3):: xd_mono          :: 18.d0
-------------------- File case (icase = 3) --------------------
Name of frequency file       :: file_name        :: {file_name}
Frequency cut-off            :: i_cut            :: 0
Low cut_off frequency (Hz)   :: nuc_low          :: 0.3d0
High cut_off frequency (Hz)  :: nuc_high         :: 3.d0
------------------ Irregular wave (icase=4) -------------------
Significant wave height (m)  :: Hs               :: 0.05d0
Peak period (s)              :: Tp               :: 1.d0
Shape factor (Jonswap)       :: gamma            :: 3.3d0
Seed number for random numb. :: iseed            :: 513
-------------------- Wavemaker definition ---------------------
Nonlinear order of wavemaker :: i_wmk            :: {i_wmk}
Type (1: piston, 2: hinged)  :: igeom            :: {igeom}
Rotation axis distance       :: d_hinge          :: {d_hinge}
Time ramp                    :: iramp            :: 1
Time ramp duration (s)       :: Tramp            :: 4.5d0
----------------------- Numerical beach -----------------------
Absorption numerical beach   :: iabsnb           :: 1
Beginning front num. beach   :: xabsf            :: 0.8d0
Absorption strength front    :: coeffabsf        :: 1.d0
------------- Elevation/Velocity-Pressure probes --------------
Use of probes                :: iprobes          :: {iprobes}
Filename of probe positions  :: pro_file         :: prob.inp
---------------------- Time-integration -----------------------
Duration of the simulation   :: T_stop           :: {tstop}
Time tolerance: RK 4(5)      :: toler            :: {toler}
Output frequency             :: f_out            :: {fout}
--------------------------- Output ----------------------------
Output: 1-dim. ; 0-nondim.   :: idim             :: 0
free surface plot            :: i3d              :: {flag}
modes plot                   :: imodes           :: {flag}
wavemaker motion plot        :: iwmk             :: {flag}
Swense output 1='yes',0='no' :: i_sw             :: 0
"""


@pytest.mark.gpu
@pytest.mark.parametrize("i_wmk,toler,tstop", [(1, 1e-7, 3.0), (2, 1e-6, 1.5)])
def test_fortran_driver_matches_oracle(tmp_path, i_wmk, toler, tstop):
    assert os.path.exists(EXE), "build the Fortran driver with hos-nwt_amd/build.sh"
    T, n1, M = 1.5, 64, 3
    s = make_setup(n1=n1, n2=1, M=M, n3=9, xlen=20.0, T=T, T_stop=tstop, f_out=2.0 / T, toler=toler, i_wmk=i_wmk)
    amp = s.cfg.amp_mono
    # 17 significant digits with an explicit decimal point are read back exactly by ES25.16
    # (a field without a decimal point would be scaled by the descriptor's 16 implied digits)
    def num(x):
        return f"{x:.16e}"

    outputs = i_wmk == 1
    text = INPUT.format(amp=num(amp), nu=num(1.0 / T), tstop=num(tstop), fout=num(2.0 / T), i_wmk=i_wmk,
                        toler=num(toler), iprobes=1 if outputs else 0, flag=1 if outputs else 0, icase=2,
                        file_name="none", igeom=1, d_hinge=num(0.0))
    (tmp_path / "input_HOS-NWT.dat").write_text(text)
    xprobes = [2.0, 5.5, 10.25]
    (tmp_path / "prob.inp").write_text("".join(f"{x}\n" for x in xprobes))
    r = subprocess.run([EXE, str(n1), "1", "9", str(M)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    got = np.array(data_rows(tmp_path / "Results" / "vol_energy.dat"))
    orc = ho.Oracle(oracle_params(s))
    st = orc.initialize()
    deta = np.zeros_like(st.eta)
    t, h = 0.0, min(0.95 * s.dt_stab, s.delt_out)
    ref = []
    snaps = []
    while t < s.T_stop + s.delt_out:
        ep, ek, et = orc.energy(st.eta, st.phis + orc.phi_add, deta)
        ref.append((t, ep, ek, et))
        snaps.append(dict(t=t, eta=st.eta.copy(), phis=st.phis.copy(), alpha=orc.alpha, alphat=orc.alphat,
                          x=orc.pos1st[-1, 0], xt=orc.dpos1stdt[-1, 0]))
        if abs(t - s.T_stop) <= 2.2e-16:
            break
        tn = min(t + s.delt_out, s.T_stop + s.delt_out)
        st, t, h, _, _, d = ho.advance(orc, st, t, tn, h, s.cfg.toler, s.dt_stab, s.delt_out)
        deta = d if d is not None else deta
    ref = np.array(ref)
    assert got.shape[0] == ref.shape[0]
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-4, atol=1e-4)
    scale = np.abs(ref[:, 3]).max()
    np.testing.assert_allclose(got[:, 2:5], ref[:, 1:4], rtol=0, atol=2e-4 * scale)
    if outputs:
        check_output_files(tmp_path / "Results", s, orc, snaps, xprobes)


@pytest.mark.gpu
def test_fortran_driver_dipole_matches_oracle(tmp_path):
    """igeom = 3: the driver builds the line dipole of a monochromatic wave (dipoles_ini) and the
    energy record follows the oracle's dipole tank; the volume column is the plain volume."""
    assert os.path.exists(EXE), "build the Fortran driver with hos-nwt_amd/build.sh"
    T, n1, M, tstop, toler = 1.5, 64, 3, 3.0, 1e-7
    s = make_setup(n1=n1, n2=1, M=M, n3=9, xlen=20.0, T=T, T_stop=tstop, f_out=2.0 / T, toler=toler, igeom=3)

    def num(x):
        return f"{x:.16e}"

    text = INPUT.format(amp=num(s.cfg.amp_mono), nu=num(1.0 / T), tstop=num(tstop), fout=num(2.0 / T), i_wmk=1,
                        toler=num(toler), iprobes=0, flag=0, icase=2, file_name="none", igeom=3, d_hinge=num(0.0))
    (tmp_path / "input_HOS-NWT.dat").write_text(text)
    r = subprocess.run([EXE, str(n1), "1", "9", str(M)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    got = np.array(data_rows(tmp_path / "Results" / "vol_energy.dat"))
    orc = ho.Oracle(oracle_params(s))
    st = orc.initialize()
    deta = np.zeros_like(st.eta)
    t, h = 0.0, min(0.95 * s.dt_stab, s.delt_out)
    ref = []
    while t < s.T_stop + s.delt_out:
        ep, ek, et = orc.energy(st.eta, st.phis + orc.phi_add, deta)
        ref.append((t, orc.volume(st.eta), ep, ek, et))
        if abs(t - s.T_stop) <= 2.2e-16:
            break
        tn = min(t + s.delt_out, s.T_stop + s.delt_out)
        st, t, h, _, _, d = ho.advance(orc, st, t, tn, h, s.cfg.toler, s.dt_stab, s.delt_out)
        deta = d if d is not None else deta
    ref = np.array(ref)
    assert got.shape[0] == ref.shape[0]
    assert np.abs(ref[-1, 4]) > 0.0
    np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(got[:, 1], ref[:, 1], rtol=0, atol=2e-4 * max(np.abs(ref[:, 1]).max(), 1e-9))
    scale = np.abs(ref[:, 4]).max()
    np.testing.assert_allclose(got[:, 2:5], ref[:, 2:5], rtol=0, atol=2e-4 * scale)


def swense_records(path, n1):
    """Records of modes_HOS_SWENSE.dat (formatted direct access, RECL = 18 n1; output.f90:219-220,
    :417-467): one list of floats per record."""
    raw = path.read_bytes()
    rl = 18 * n1
    step = rl + 1 if len(raw) % (rl + 1) == 0 and raw[rl:rl + 1] == b"\n" else rl
    assert len(raw) % step == 0
    return [[float(v) for v in raw[i:i + rl].decode().split()] for i in range(0, len(raw), step)]


@pytest.mark.gpu
def test_fortran_driver_writes_swense_modes(tmp_path):
    """i_sw = 1: modes_HOS_SWENSE.dat holds the header record, a zero block at t = 0 and, at
    each output time, the n2 x 8 records of the H2 operator's modes computed after the last
    attempt (10 levels), equal to hos_swense_modes of the same run to the file's 10 digits."""
    assert os.path.exists(EXE), "build the Fortran driver with hos-nwt_amd/build.sh"
    T, n1, M, tstop, toler = 1.5, 64, 3, 1.0, 1e-7
    s = make_setup(n1=n1, n2=1, M=M, n3=9, xlen=20.0, T=T, T_stop=tstop, f_out=2.0 / T, toler=toler)
    text = INPUT.format(amp=f"{s.cfg.amp_mono:.16e}", nu=f"{1.0 / T:.16e}", tstop=f"{tstop:.16e}",
                        fout=f"{2.0 / T:.16e}", i_wmk=1, toler=f"{toler:.16e}", iprobes=0, flag=0, icase=2,
                        file_name="none", igeom=1, d_hinge=f"{0.0:.16e}")
    text = text.replace("i_sw             :: 0", "i_sw             :: 1")
    (tmp_path / "input_HOS-NWT.dat").write_text(text)
    r = subprocess.run([EXE, str(n1), "1", "9", str(M)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    recs = swense_records(tmp_path / "Results" / "modes_HOS_SWENSE.dat", n1)
    n3a = 16
    # header: n1, n3_add, n2, 1/f_out, T_stop, xlen, ylen, h, then modes 9..n1 of the t = 0 block
    hdr = recs[0]
    assert hdr[:3] == [n1, n3a, 1.0]
    np.testing.assert_allclose(hdr[3:8], [1.0 / s.f_out, s.T_stop, s.xlen, 1.0, 1.0], rtol=1e-9)
    from hosnwt.tank import Tank
    tank = Tank(s, precision=0)
    try:
        tank.initialize()
        t, h = 0.0, min(0.95 * s.dt_stab, s.delt_out)
        # the driver's loop (HOS-NWT.f90:446-581): an output, then the advance whose last
        # attempt computes the next output's modes; none before the first (zeros)
        blocks, cur = [], None
        while t < s.T_stop + s.delt_out:
            blocks.append(cur)
            if abs(t - s.T_stop) <= 2.2e-16:
                break
            tn = min(t + s.delt_out, s.T_stop + s.delt_out)
            h, _, _ = tank.advance(tn, h)
            t = tank.time
            cur = tank.swense_modes(t, 10)
        assert len(recs) == 8 * (len(blocks) + 1)
        for ib, b in enumerate(blocks):
            rr = recs[8 * (ib + 1):8 * (ib + 2)]
            if b is None:
                assert all(not any(x) for x in rr)
                continue
            ref = [b["specx"][:, 0], b["specy"][:, 0], b["specz"][:, 0], b["spect"][:, 0], b["FS"][:, 0],
                   b["FSt"][:, 0], b["add"][:, 0], b["addt"][:, 0]]
            for q in range(8):
                got = np.array(rr[q])
                assert got.shape == ref[q].shape, (ib, q)
                np.testing.assert_allclose(got, ref[q], rtol=0, atol=1e-8 * max(np.abs(ref[q]).max(), 1e-30),
                                           err_msg=f"block {ib} record {q}")
            assert np.abs(ref[0]).max() > 0
    finally:
        tank.close()


def data_rows(path):
    """Numeric rows of a Tecplot file (skips the '#' echo, TITLE, VARIABLES and ZONE lines)."""
    rows = []
    for ln in path.read_text().split("\n"):
        t = ln.strip()
        if not t or t.startswith("#") or t.startswith("TITLE") or t.startswith("VARIABLES") or t.startswith("ZONE"):
            continue
        rows.append([float(v.replace("D", "E")) for v in t.split()])
    return rows


def close(got, ref, what):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-4 * max(np.abs(ref).max(), 1e-12), err_msg=what)


def check_output_files(res, s, orc, snaps, xprobes):
    n1 = s.cfg.n1
    tb = orc.tb
    # header echo (IO/input.f90:318-395): 59 '#' lines before TITLE, same in every file
    for name in ("vol_energy.dat", "probes.dat", "wmk_motion.dat", "3d.dat", "HOS_modes.dat"):
        lines = (res / name).read_text().split("\n")
        head = [ln for ln in lines if ln.startswith("#")]
        assert len(head) == 59, (name, len(head))
        assert head[0] == "#----------------------- Tank dimensions -----------------------"
        # '("#",A33,":: ",A13,":: ",ES25.16)': character data shorter than the field is right-justified
        assert head[1] == "#" + "Length(m) of the wavetank".rjust(33) + ":: " + "xlen".rjust(13) + ":: " + \
            "2.0000000000000000E+01".rjust(25)
        assert head[-1].split("::")[1].strip() == "coeffilt(3)"
        assert "TITLE" in lines[59]
        assert lines[60].startswith("VARIABLES")
    # probes.dat: time, eta at the probes = sum_k eta_hat cos(kx x_p) (output.f90:266-276)
    pr = np.array(data_rows(res / "probes.dat"))
    want = np.array([[d["t"]] + [float(np.dot(orc.s2f(d["eta"])[:, 0], np.cos(tb.kx * x))) for x in xprobes]
                     for d in snaps])
    close(pr[:, 0], want[:, 0], "probes time")
    close(pr[:, 1:], want[:, 1:], "probes eta")
    # wmk_motion.dat (2-D): time, alpha X_top, alpha' X_top + alpha X_t,top (output.f90:361-366)
    wm = np.array(data_rows(res / "wmk_motion.dat"))
    close(wm[:, 1], [d["alpha"] * d["x"] for d in snaps], "wmk X-top")
    close(wm[:, 2], [d["alphat"] * d["x"] + d["alpha"] * d["xt"] for d in snaps], "wmk Vx-top")
    # 3d.dat: first zone x, y, eta, phis; later zones eta, eta - eta_target (output.f90:291-338)
    z = data_rows(res / "3d.dat")
    first = np.array(z[:n1])
    close(first[:, 0], tb.cx, "3d x")
    close(first[:, 2], snaps[0]["eta"][:, 0], "3d eta t=0")
    close(first[:, 3], snaps[0]["phis"][:, 0], "3d phis t=0")
    for i, d in enumerate(snaps[1:], start=1):
        blk = np.array(z[n1 * i:n1 * (i + 1)])
        close(blk[:, 0], d["eta"][:, 0], f"3d eta zone {i}")
    zones = [ln for ln in (res / "3d.dat").read_text().split("\n") if ln.startswith("ZONE")]
    assert len(zones) == len(snaps)
    assert zones[0].startswith("ZONE SOLUTIONTIME =   0.00, I=  64, J=   1")
    assert "D=(1,2)" in zones[1]
    # HOS_modes.dat: first zone kx, ky, eta_hat, phis_hat, log10 |.|; later zones without kx, ky
    m = data_rows(res / "HOS_modes.dat")
    first = np.array(m[:n1])
    close(first[:, 0], tb.kx, "modes kx")
    for i, d in enumerate(snaps):
        blk = np.array(m[n1 * i:n1 * (i + 1)])
        c = 2 if i == 0 else 0
        close(blk[:, c], orc.s2f(d["eta"])[:, 0], f"modes eta zone {i}")
        close(blk[:, c + 1], orc.s2f(d["phis"])[:, 0], f"modes phis zone {i}")


@pytest.mark.gpu
def test_fortran_driver_reads_wave_files(tmp_path):
    """icase = 3: the driver reads <file_name>.cfg / .dat (file2wave, wave_def.f90:181-278) for a
    flap and runs to the oracle's volume / energy trace built from the same components."""
    assert os.path.exists(EXE), "build the Fortran driver with hos-nwt_amd/build.sh"
    n1, M, tstop, T = 64, 3, 2.0, 1.5
    comps = hc.jonswap_components(0.05, 1.5, 3.3, n_freq=6, seed=21, rnum=6, clock=8.0)
    d_hinge = 0.3
    s = make_setup(n1=n1, n2=1, M=M, n3=9, xlen=20.0, T=T, T_stop=tstop, f_out=2.0 / T, igeom=2,
                   d_hinge=d_hinge, components=comps, icase=3)

    def num(x):
        return f"{x:.16e}"

    # the reference's wave files: .cfg (rnum as I3, then one value per line), .dat components
    cfg = [f"{comps['rnum']:3d}", num(comps["clock"]), num(1.0), num(0.0), "lin", num(0.0), num(100.0), num(1.0),
           "hinged", num(d_hinge)]
    (tmp_path / "waves.cfg").write_text("\n".join(cfg) + "\n")
    rows = [f"{h:d} {num(f)} {num(a)} {num(an)} {num(p)}" for h, f, a, an, p in
            zip(comps["harmo"], comps["freq"], comps["ampli"], comps["angle"], comps["phase"])]
    (tmp_path / "waves.dat").write_text("\n".join(rows) + "\n")
    text = INPUT.format(amp=num(0.0), nu=num(1.0 / T), tstop=num(tstop), fout=num(2.0 / T), i_wmk=1,
                        toler=num(1e-7), iprobes=0, flag=0, icase=3, file_name="waves", igeom=2,
                        d_hinge=num(d_hinge))
    (tmp_path / "input_HOS-NWT.dat").write_text(text)
    r = subprocess.run([EXE, str(n1), "1", "9", str(M)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    got = np.array(data_rows(tmp_path / "Results" / "vol_energy.dat"))
    orc = ho.Oracle(oracle_params(s))
    st = orc.initialize()
    deta = np.zeros_like(st.eta)
    t, h = 0.0, min(0.95 * s.dt_stab, s.delt_out)
    ref = []
    while t < s.T_stop + s.delt_out:
        ep, ek, et = orc.energy(st.eta, st.phis + orc.phi_add, deta)
        ref.append((t, ep, ek, et))
        if abs(t - s.T_stop) <= 2.2e-16:
            break
        tn = min(t + s.delt_out, s.T_stop + s.delt_out)
        st, t, h, _, _, d = ho.advance(orc, st, t, tn, h, s.cfg.toler, s.dt_stab, s.delt_out)
        deta = d if d is not None else deta
    ref = np.array(ref)
    assert got.shape[0] == ref.shape[0]
    scale = np.abs(ref[:, 3]).max()
    assert scale > 0
    np.testing.assert_allclose(got[:, 2:5], ref[:, 1:4], rtol=0, atol=2e-4 * scale)


@pytest.mark.gpu
@pytest.mark.parametrize("icase", [4, 41])
def test_fortran_driver_builds_internal_spectra(tmp_path, icase):
    """icase = 4 (JONSWAP) / 41 (Bretschneider): the driver builds the components in memory
    (initialize_wav_cfg, initial_condition.f90:212-308) and runs to the oracle's volume /
    energy trace built from hosnwt.config.spectrum_components (the same discretisation and
    the same phase sequence)."""
    assert os.path.exists(EXE), "build the Fortran driver with hos-nwt_amd/build.sh"
    n1, M, tstop, T = 64, 3, 2.0, 1.5
    d_hinge = 0.3
    comps = hc.spectrum_components(icase, 0.05, 1.0, 3.3, 513, tstop)  # Hs, Tp, gamma, iseed of INPUT
    print(icase, "components:", len(comps["freq"]), "rnum", comps["rnum"], flush=True)
    s = make_setup(n1=n1, n2=1, M=M, n3=9, xlen=20.0, T=T, T_stop=tstop, f_out=2.0 / T, igeom=2,
                   d_hinge=d_hinge, components=comps, icase=3)

    def num(x):
        return f"{x:.16e}"

    text = INPUT.format(amp=num(0.0), nu=num(1.0 / T), tstop=num(tstop), fout=num(2.0 / T), i_wmk=1,
                        toler=num(1e-7), iprobes=0, flag=0, icase=icase, file_name="none", igeom=2,
                        d_hinge=num(d_hinge))
    (tmp_path / "input_HOS-NWT.dat").write_text(text)
    r = subprocess.run([EXE, str(n1), "1", "9", str(M)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "Hs_calc" in r.stdout
    got = np.array(data_rows(tmp_path / "Results" / "vol_energy.dat"))
    orc = ho.Oracle(oracle_params(s))
    st = orc.initialize()
    deta = np.zeros_like(st.eta)
    t, h = 0.0, min(0.95 * s.dt_stab, s.delt_out)
    ref = []
    while t < s.T_stop + s.delt_out:
        ep, ek, et = orc.energy(st.eta, st.phis + orc.phi_add, deta)
        ref.append((t, ep, ek, et))
        if abs(t - s.T_stop) <= 2.2e-16:
            break
        tn = min(t + s.delt_out, s.T_stop + s.delt_out)
        st, t, h, _, _, d = ho.advance(orc, st, t, tn, h, s.cfg.toler, s.dt_stab, s.delt_out)
        deta = d if d is not None else deta
    ref = np.array(ref)
    assert got.shape[0] == ref.shape[0]
    scale = np.abs(ref[:, 3]).max()
    assert scale > 0
    np.testing.assert_allclose(got[:, 2:5], ref[:, 1:4], rtol=0, atol=2e-4 * scale)
