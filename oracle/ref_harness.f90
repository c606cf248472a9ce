! ref_harness.f90 -- C-callable harness around the reference's own Fortran kernels.
!
! TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile ('make ref') together with the
! reference's shared/utils/voigt/hestark/opacity/solver/convol sources, read in place from
! the reference tree, into oracle/_ref/libfsref.so.  It fills the reference's module
! globals from plain arrays and calls the reference routines unchanged, so that the C
! oracle (oracle/fs_oracle.c) can be compared with the reference's own arithmetic.  Only
! tests and bench.py's cpu_baseline leg load it.
!
! Limitation: the reference keeps per-depth arrays sized by the first Ntau it sees
! (XlineInit, HeStarkInit), so one process uses one Ntau.

MODULE RefHarness

  USE ISO_C_BINDING
  USE Constants
  USE Config
  USE LineData
  USE AtmosData
  USE SpecData
  USE Voigt
  USE FaradayVoigt
  USE Utils
  USE HeStark
  USE Opacity
  USE Solver
  USE Convol

  IMPLICIT NONE

  INTEGER, SAVE :: ListType = 0

CONTAINS

  SUBROUTINE ref_hvoigt(n, a, v, h) BIND(C, NAME='ref_hvoigt')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: a(n), v(n)
    REAL(C_DOUBLE), INTENT(OUT) :: h(n)
    INTEGER :: i
    CALL VoigtInit
    DO i = 1, n
      h(i) = HVoigt(a(i), v(i))
    END DO
  END SUBROUTINE ref_hvoigt

  SUBROUTINE ref_hvoigt_exact(n, a, v, h) BIND(C, NAME='ref_hvoigt_exact')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: a(n), v(n)
    REAL(C_DOUBLE), INTENT(OUT) :: h(n)
    INTEGER :: i
    DO i = 1, n
      h(i) = HVoigtExact(a(i), v(i))
    END DO
  END SUBROUTINE ref_hvoigt_exact

  SUBROUTINE ref_fvoigt(n, a, v, h, f) BIND(C, NAME='ref_fvoigt')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: a(n), v(n)
    REAL(C_DOUBLE), INTENT(OUT) :: h(n), f(n)
    INTEGER :: i
    CALL FVoigtInit
    DO i = 1, n
      CALL FVoigt(a(i), v(i), h(i), f(i))
    END DO
  END SUBROUTINE ref_fvoigt

  SUBROUTINE ref_wvoigt_exact(n, a, v, h, f) BIND(C, NAME='ref_wvoigt_exact')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: a(n), v(n)
    REAL(C_DOUBLE), INTENT(OUT) :: h(n), f(n)
    INTEGER :: i
    COMPLEX(8) :: w
    DO i = 1, n
      w = WVoigtExact(a(i), v(i))
      h(i) = REAL(w, 8)
      f(i) = AIMAG(w)
    END DO
  END SUBROUTINE ref_wvoigt_exact

  ! Atmosphere, EOS tables, grid and configuration -> module globals; then VoigtInit and
  ! XlineInit (which runs HeStarkInit).  Slots are 1-based EOS indices.
  SUBROUTINE ref_setup(ntau_, nwl_, nmu_, wmode_, wfirst_, dstep_, rstep_, wstart_, wave_, &
                       motype_, rhox, temp, xne, rho, contop, contopstd, &
                       nslot, fracpf, frac, potion_, xamass_, nspc, spcion, spcslot, &
                       slot_h1, slot_he1, slot_hmol, slot_h2, slot_he2, &
                       epsopac_, vwfudge_, h2brd_, vwmol_, vmicro_, molbrd_, xmu_, wmu_) &
                       BIND(C, NAME='ref_setup')
    INTEGER(C_INT), VALUE :: ntau_, nwl_, nmu_, wmode_, motype_, nslot, nspc, molbrd_
    INTEGER(C_INT), VALUE :: slot_h1, slot_he1, slot_hmol, slot_h2, slot_he2
    REAL(C_DOUBLE), VALUE :: wfirst_, dstep_, rstep_, wstart_
    REAL(C_DOUBLE), VALUE :: epsopac_, vwfudge_, h2brd_, vwmol_, vmicro_
    REAL(C_DOUBLE), INTENT(IN) :: wave_(nwl_), rhox(ntau_), temp(ntau_), xne(ntau_), rho(ntau_)
    REAL(C_DOUBLE), INTENT(IN) :: contop(ntau_, nwl_), contopstd(ntau_), xmu_(nmu_), wmu_(nmu_)
    REAL(C_FLOAT), INTENT(IN) :: fracpf(nslot, ntau_), frac(nslot, ntau_), potion_(nslot), xamass_(nslot)
    INTEGER(C_INT), INTENT(IN) :: spcion(nspc), spcslot(nspc)
    INTEGER :: k

    Ntau = ntau_
    Nwl = nwl_
    Nmu = nmu_
    Wmode = wmode_
    Wfirst = wfirst_
    Dstep = dstep_
    Rstep = rstep_
    Wstart = wstart_
    Motype = motype_
    EpsOpac = epsopac_
    VWfudge = vwfudge_
    H2Brd = h2brd_
    VWmol = vwmol_
    Vmicro = vmicro_
    MolBrd = (molbrd_ /= 0)

    IF (ALLOCATED(Atmos)) DEALLOCATE(Atmos)
    ALLOCATE(Atmos(Ntau))
    DO k = 1, Ntau
      Atmos(k)%Rhox = rhox(k)
      Atmos(k)%T = temp(k)
      Atmos(k)%Xne = xne(k)
      Atmos(k)%Xna = 0d0
      Atmos(k)%Rho = rho(k)
    END DO

    IF (ALLOCATED(Wave)) DEALLOCATE(Wave)
    IF (ALLOCATED(ContOpac)) DEALLOCATE(ContOpac)
    IF (ALLOCATED(ContOpacStd)) DEALLOCATE(ContOpacStd)
    IF (ALLOCATED(LineOpac)) DEALLOCATE(LineOpac)
    ALLOCATE(Wave(Nwl), ContOpac(Ntau, Nwl), ContOpacStd(Ntau), LineOpac(Ntau, Nwl))
    Wave = wave_
    ContOpac = contop
    ContOpacStd = contopstd
    LineOpac = 0d0

    IF (ALLOCATED(Fraction)) DEALLOCATE(Fraction)
    IF (ALLOCATED(FractionPF)) DEALLOCATE(FractionPF)
    IF (ALLOCATED(PotIon)) DEALLOCATE(PotIon)
    IF (ALLOCATED(Xamass)) DEALLOCATE(Xamass)
    ALLOCATE(Fraction(nslot, Ntau), FractionPF(nslot, Ntau), PotIon(nslot), Xamass(nslot))
    Fraction = frac
    FractionPF = fracpf
    PotIon = potion_
    Xamass = xamass_

    ! species table: the caller's species, then five entries for the perturber slots
    IF (ALLOCATED(Ion)) DEALLOCATE(Ion)
    IF (ALLOCATED(IdxSpc)) DEALLOCATE(IdxSpc)
    ALLOCATE(Ion(nspc + 5), IdxSpc(nspc + 5))
    Ion(1:nspc) = spcion
    IdxSpc(1:nspc) = spcslot
    Ion(nspc+1:nspc+5) = [1, 1, 1, 2, 2]
    IdxSpc(nspc+1:nspc+5) = [slot_h1, slot_he1, slot_hmol, slot_h2, slot_he2]
    Ix%H1 = nspc + 1
    Ix%He1 = nspc + 2
    Ix%Hmol = nspc + 3
    Ix%H2 = nspc + 4
    Ix%He2 = nspc + 5

    IF (ALLOCATED(Xmu)) DEALLOCATE(Xmu)
    IF (ALLOCATED(Wmu)) DEALLOCATE(Wmu)
    ALLOCATE(Xmu(Nmu), Wmu(Nmu))
    Xmu = xmu_
    Wmu = wmu_

    CALL VoigtInit
    CALL XlineInit
  END SUBROUTINE ref_setup

  ! He I Stark tables built by the reference's HeStarkInit (after ref_setup)
  SUBROUTINE ref_hestark(nhe, wl, elo, gam, shf) BIND(C, NAME='ref_hestark')
    INTEGER(C_INT), INTENT(OUT) :: nhe
    REAL(C_DOUBLE), INTENT(OUT) :: wl(NheTab), elo(NheTab), gam(Ntau, NheTab), shf(Ntau, NheTab)
    nhe = NheTab
    wl = WlcHe
    elo = EloHe
    gam = HeGam
    shf = HeShf
  END SUBROUTINE ref_hestark

  ! A line list -> LinesA/LinesM/LinesZ, Lines, Wlcent, LinSpc, LinFlag (bits 1, 2 as set
  ! in pass 1 of LineopCalc, src/synth.f90:839-844), Ileft = 1, Iright = Nwl.
  SUBROUTINE ref_set_lines(ltype, n, wl, elo, gf, grad, gqst, gvdw, spc, ismol, &
                           npi_off, pi_off, spl_pi, str_pi, si_off, spl_si, str_si) &
                           BIND(C, NAME='ref_set_lines')
    INTEGER(C_INT), VALUE :: ltype, n, npi_off
    REAL(C_DOUBLE), INTENT(IN) :: wl(n)
    REAL(C_FLOAT), INTENT(IN) :: elo(n), gf(n), grad(n), gqst(n), gvdw(n)
    INTEGER(C_INT), INTENT(IN) :: spc(n), ismol(n)
    INTEGER(C_INT), INTENT(IN) :: pi_off(*), si_off(*)
    REAL(C_FLOAT), INTENT(IN) :: spl_pi(*), str_pi(*), spl_si(*), str_si(*)
    INTEGER :: i, a, b

    ListType = ltype
    Nlines = n
    IF (ALLOCATED(LinesA)) DEALLOCATE(LinesA)
    IF (ALLOCATED(LinesM)) DEALLOCATE(LinesM)
    IF (ALLOCATED(LinesZ)) DEALLOCATE(LinesZ)
    IF (ALLOCATED(Wlcent)) DEALLOCATE(Wlcent)
    IF (ALLOCATED(LinSpc)) DEALLOCATE(LinSpc)
    IF (ALLOCATED(LinFlag)) DEALLOCATE(LinFlag)
    IF (ALLOCATED(Ileft)) DEALLOCATE(Ileft)
    IF (ALLOCATED(Iright)) DEALLOCATE(Iright)
    ALLOCATE(Wlcent(n), LinSpc(n), LinFlag(n), Ileft(n), Iright(n))
    LinFlag = 0
    Ileft = 1
    Iright = Nwl
    LinSpc = spc
    IF (ltype == 0) THEN
      ALLOCATE(LinesA(n))
      DO i = 1, n
        LinesA(i)%Wl = wl(i)
        LinesA(i)%Elo = elo(i)
        LinesA(i)%Gf = gf(i)
        LinesA(i)%Grad = grad(i)
        LinesA(i)%Gqst = gqst(i)
        LinesA(i)%Gvdw = gvdw(i)
        LinesA(i)%Ion = 1
        LinesA(i)%El = 'Fe'
        IF (LinesA(i)%Gvdw > 10.0) LinFlag(i) = IBSET(LinFlag(i), 2)
      END DO
      Lines => LinesA
    ELSE IF (ltype == 1) THEN
      ALLOCATE(LinesM(n))
      DO i = 1, n
        LinesM(i)%Wl = wl(i)
        LinesM(i)%Elo = elo(i)
        LinesM(i)%Gf = gf(i)
        LinesM(i)%Grad = grad(i)
      END DO
      Lines => LinesM
    ELSE
      ALLOCATE(LinesZ(n))
      DO i = 1, n
        LinesZ(i)%Wl = wl(i)
        LinesZ(i)%Elo = elo(i)
        LinesZ(i)%Gf = gf(i)
        LinesZ(i)%Grad = grad(i)
        LinesZ(i)%Gqst = gqst(i)
        LinesZ(i)%Gvdw = gvdw(i)
        LinesZ(i)%Ion = 1
        LinesZ(i)%El = 'Fe'
        IF (ismol(i) /= 0) LinFlag(i) = IBSET(LinFlag(i), 1)
        IF (LinesZ(i)%Gvdw > 10.0) LinFlag(i) = IBSET(LinFlag(i), 2)
        IF (npi_off > 0) THEN
          a = pi_off(i) + 1
          b = pi_off(i + 1)
          LinesZ(i)%Npi = b - a + 1
          ALLOCATE(LinesZ(i)%SplPi(b - a + 1), LinesZ(i)%StrPi(b - a + 1))
          LinesZ(i)%SplPi = spl_pi(a:b)
          LinesZ(i)%StrPi = str_pi(a:b)
          a = si_off(i) + 1
          b = si_off(i + 1)
          LinesZ(i)%Nsi = b - a + 1
          ALLOCATE(LinesZ(i)%SplSi(b - a + 1), LinesZ(i)%StrSi(b - a + 1))
          LinesZ(i)%SplSi = spl_si(a:b)
          LinesZ(i)%StrSi = str_si(a:b)
        END IF
      END DO
      Lines => LinesZ
    END IF
    Wlcent = Lines%Wl - Wfirst
  END SUBROUTINE ref_set_lines

  ! XlineSetup + XlineDepth for every line; He I table index folded into LinFlag bits 8-15
  ! as pass 1 does (src/synth.f90:847-857).  Outputs (Ntau, n).
  SUBROUTINE ref_xline(cop, vv, av, flag) BIND(C, NAME='ref_xline')
    REAL(C_DOUBLE), INTENT(OUT) :: cop(Ntau, Nlines), vv(Ntau, Nlines), av(Ntau, Nlines)
    INTEGER(C_INT), INTENT(OUT) :: flag(Nlines)
    INTEGER :: i, k
    TYPE(XlineCoef) :: XC
    ! lines are independent, as in pass 1's OpenMP loop (src/synth.f90:833)
    !$OMP PARALLEL DO SCHEDULE(DYNAMIC, 256) DEFAULT(SHARED) PRIVATE(i, k, XC)
    DO i = 1, Nlines
      CALL XlineSetup(i, XC)
      IF (XC%IxHe > 0) LinFlag(i) = IOR(LinFlag(i), ISHFT(XC%IxHe, 8))
      DO k = 1, Ntau
        CALL XlineDepth(k, XC, cop(k, i), vv(k, i), av(k, i))
      END DO
      flag(i) = LinFlag(i)
    END DO
    !$OMP END PARALLEL DO
  END SUBROUTINE ref_xline

  ! Overrides the per-line flags and windows (after pass 1 of the oracle)
  SUBROUTINE ref_set_windows(flag, il, ir) BIND(C, NAME='ref_set_windows')
    INTEGER(C_INT), INTENT(IN) :: flag(Nlines), il(Nlines), ir(Nlines)
    LinFlag = flag
    Ileft = il
    Iright = ir
  END SUBROUTINE ref_set_windows

  ! Loads block coefficient arrays (Ntau, n)
  SUBROUTINE ref_set_block(n, cop, vv, av) BIND(C, NAME='ref_set_block')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: cop(Ntau, n), vv(Ntau, n), av(Ntau, n)
    IF (ALLOCATED(LinCop)) DEALLOCATE(LinCop)
    IF (ALLOCATED(LinVvgt)) DEALLOCATE(LinVvgt)
    IF (ALLOCATED(LinAvgt)) DEALLOCATE(LinAvgt)
    ALLOCATE(LinCop(Ntau, n), LinVvgt(Ntau, n), LinAvgt(Ntau, n))
    LinCop = cop
    LinVvgt = vv
    LinAvgt = av
  END SUBROUTINE ref_set_block

  ! LineOpacity1 over grid points iwl1..iwl2 with lines L1..L2 of the loaded block
  ! (Loff = L1 - 1), OpenMP over wavelength as at src/synth.f90:1075-1080; adds into out.
  SUBROUTINE ref_lineop1(iwl1, iwl2, l1, l2, out) BIND(C, NAME='ref_lineop1')
    INTEGER(C_INT), VALUE :: iwl1, iwl2, l1, l2
    REAL(C_DOUBLE), INTENT(INOUT) :: out(Ntau, *)
    REAL(8) :: LineOpacL(Ntau)
    INTEGER :: iwl
    !$OMP PARALLEL DO SCHEDULE(DYNAMIC) DEFAULT(SHARED) PRIVATE(iwl, LineOpacL)
    DO iwl = iwl1, iwl2
      CALL LineOpacity1(Wave(iwl), l1, l2, iwl, l1 - 1, LineOpacL)
      out(:, iwl - iwl1 + 1) = out(:, iwl - iwl1 + 1) + LineOpacL
    END DO
    !$OMP END PARALLEL DO
  END SUBROUTINE ref_lineop1

  ! LineOpacity2 of line i at grid index iwl with explicit coefficients
  SUBROUTINE ref_lineop2(wl, i, iwl, cop, vv, av, out) BIND(C, NAME='ref_lineop2')
    REAL(C_DOUBLE), VALUE :: wl
    INTEGER(C_INT), VALUE :: i, iwl
    REAL(C_DOUBLE), INTENT(IN) :: cop(Ntau), vv(Ntau), av(Ntau)
    REAL(C_DOUBLE), INTENT(OUT) :: out(Ntau)
    CALL LineOpacity2(wl, i, iwl, cop, vv, av, out)
  END SUBROUTINE ref_lineop2

  ! AnalyticWindow half-width of line i
  SUBROUTINE ref_analytic_window(i, cop, vv, av, contl, dw) BIND(C, NAME='ref_analytic_window')
    INTEGER(C_INT), VALUE :: i
    REAL(C_DOUBLE), INTENT(IN) :: cop(Ntau), vv(Ntau), av(Ntau), contl(Ntau)
    REAL(C_DOUBLE), INTENT(OUT) :: dw
    CALL AnalyticWindow(i, cop, vv, av, contl, dw)
  END SUBROUTINE ref_analytic_window

  INTEGER(C_INT) FUNCTION ref_waveidx(w) BIND(C, NAME='ref_waveidx')
    REAL(C_DOUBLE), VALUE :: w
    ref_waveidx = WaveIdx(w)
  END FUNCTION ref_waveidx

  REAL(C_DOUBLE) FUNCTION ref_air2vac(w) BIND(C, NAME='ref_air2vac')
    REAL(C_DOUBLE), VALUE :: w
    ref_air2vac = Air2Vac(w)
  END FUNCTION ref_air2vac

  ! Zeeman shift scaling as in MagLineopCalc (src/synth.f90:1193-1194), then the window
  ! search of pass 1 (src/synth.f90:1243-1272) with the reference's MagLineOpacity2 and
  ! RefineIdx, for line i with the given coefficients and continuum at the line centre.
  SUBROUTINE ref_mag_window(i, bsplit, cop, vv, av, contl, il, ir) BIND(C, NAME='ref_mag_window')
    INTEGER(C_INT), VALUE :: i
    REAL(C_DOUBLE), VALUE :: bsplit
    REAL(C_DOUBLE), INTENT(IN) :: cop(Ntau), vv(Ntau), av(Ntau), contl(Ntau)
    INTEGER(C_INT), INTENT(OUT) :: il, ir
    INTEGER :: idw, iwl
    REAL(8) :: SplMax, dW, Wl, MaxOpac, LineOpacL(Ntau)
    CALL VoigtInit
    LinesZ(i)%SplPi = REAL(LinesZ(i)%SplPi * bsplit * (Wlcent(i) + Wfirst)**2, 4)
    LinesZ(i)%SplSi = REAL(LinesZ(i)%SplSi * bsplit * (Wlcent(i) + Wfirst)**2, 4)
    SplMax = MAX(MAXVAL(ABS(LinesZ(i)%SplSi)), MAXVAL(ABS(LinesZ(i)%SplPi)))
    iwl = WaveIdx(Wlcent(i))
    idw = 0
    DO
      dW = (DWstart * DWstep ** idw) * (Wlcent(i) + Wfirst) / 5d3
      Wl = Wlcent(i) + SplMax + dW
      CALL MagLineOpacity2(Wl, i, iwl, cop, vv, av, LineOpacL)
      MaxOpac = MAXVAL(LineOpacL / contl)
      IF (MaxOpac < EpsOpac) THEN
        CALL RefineIdx(i, iwl, idw, cop, vv, av, contl, SplMax, .TRUE.)
        EXIT
      ELSE IF (Wlcent(i) - SplMax - dW < Wave(1) .AND. Wlcent(i) + SplMax + dW > Wave(Nwl)) THEN
        EXIT
      END IF
      idw = idw + 1
    END DO
    il = Ileft(i)
    ir = Iright(i)
  END SUBROUTINE ref_mag_window

  ! MagLineOpacity1 over grid points iwl1..iwl2, lines l1..l2 of the loaded block; out(6,Ntau,*)
  SUBROUTINE ref_maglineop1(iwl1, iwl2, l1, l2, out) BIND(C, NAME='ref_maglineop1')
    INTEGER(C_INT), VALUE :: iwl1, iwl2, l1, l2
    REAL(C_DOUBLE), INTENT(INOUT) :: out(6, Ntau, *)
    REAL(8) :: M(6, Ntau)
    INTEGER :: iwl
    CALL FVoigtInit
    !$OMP PARALLEL DO SCHEDULE(DYNAMIC) DEFAULT(SHARED) PRIVATE(iwl, M)
    DO iwl = iwl1, iwl2
      CALL MagLineOpacity1(Wave(iwl), l1, l2, iwl, l1 - 1, M)
      out(:, :, iwl - iwl1 + 1) = out(:, :, iwl - iwl1 + 1) + M
    END DO
    !$OMP END PARALLEL DO
  END SUBROUTINE ref_maglineop1

  ! LocPrfPP for one opacity column
  SUBROUTINE ref_locprfpp(ploc, xki, src) BIND(C, NAME='ref_locprfpp')
    REAL(C_DOUBLE), INTENT(OUT) :: ploc(Nmu)
    REAL(C_DOUBLE), INTENT(IN) :: xki(Ntau), src(Ntau)
    CALL LocPrfPP(ploc, Xmu, Nmu, xki, src, Atmos%Rhox, Ntau)
  END SUBROUTINE ref_locprfpp

  ! The RadTranCalc loop (src/synth.f90:1850-1871) over grid points iwl1..iwl2 with the
  ! reference's LocPrfPP, using LineOpac columns given in lop(Ntau, *); spl/spc (Nmu, *)
  SUBROUTINE ref_radtran(iwl1, iwl2, lop, spl, spc) BIND(C, NAME='ref_radtran')
    INTEGER(C_INT), VALUE :: iwl1, iwl2
    REAL(C_DOUBLE), INTENT(IN) :: lop(Ntau, *)
    REAL(C_DOUBLE), INTENT(OUT) :: spl(Nmu, *), spc(Nmu, *)
    REAL(8), PARAMETER :: PlkC1 = Hc * Vc * 1d13 / Kbolt
    REAL(8) :: T(Ntau), Rhox(Ntau), Src(Ntau), Xki(Ntau), C1, C2
    INTEGER :: iwl, j
    T = Atmos%T
    Rhox = Atmos%Rhox
    !$OMP PARALLEL DO DEFAULT(SHARED) PRIVATE(iwl, j, C1, C2, Src, Xki)
    DO iwl = iwl1, iwl2
      j = iwl - iwl1 + 1
      C1 = PlkC1 / (Wave(iwl) + Wfirst)
      C2 = EXP(50.7649141D0 - 5.D0 * LOG(Wave(iwl) + Wfirst))
      Src = C2 / (EXP(C1 / T) - 1.D0)
      Xki = ContOpac(:, iwl)
      CALL LocPrfPP(spc(:, j), Xmu, Nmu, Xki, Src, Rhox, Ntau)
      Xki = ContOpac(:, iwl) + lop(:, j)
      CALL LocPrfPP(spl(:, j), Xmu, Nmu, Xki, Src, Rhox, Ntau)
    END DO
    !$OMP END PARALLEL DO
  END SUBROUTINE ref_radtran

  ! StokesPrfPP at grid point iwl with MagLineOpac/LineOpac columns given for that point
  SUBROUTINE ref_stokesprfpp(iwl, mlop, lop, src, sphi2, cphi2, cphi, s2psi, c2psi, stokes) &
                             BIND(C, NAME='ref_stokesprfpp')
    INTEGER(C_INT), VALUE :: iwl
    REAL(C_DOUBLE), INTENT(IN) :: mlop(6, Ntau), lop(Ntau), src(Ntau)
    REAL(C_DOUBLE), INTENT(IN) :: sphi2(Nmu), cphi2(Nmu), cphi(Nmu), s2psi(Nmu), c2psi(Nmu)
    REAL(C_DOUBLE), INTENT(OUT) :: stokes(Nmu, 4)
    IF (ALLOCATED(MagLineOpac)) DEALLOCATE(MagLineOpac)
    ALLOCATE(MagLineOpac(6, Ntau, Nwl))
    MagLineOpac = 0d0
    MagLineOpac(:, :, iwl) = mlop
    LineOpac(:, iwl) = lop
    CALL StokesPrfPP(iwl, src, Atmos%Rhox, sphi2, cphi2, cphi, s2psi, c2psi, stokes)
  END SUBROUTINE ref_stokesprfpp

  ! SphereInit on the current set-up for Radius and layer heights; returns the ray geometry
  SUBROUTINE ref_sphere(radius_, height_, ntmu, graz, map, rhoxmu_, muloc_) BIND(C, NAME='ref_sphere')
    REAL(C_DOUBLE), VALUE :: radius_
    REAL(C_DOUBLE), INTENT(IN) :: height_(Ntau)
    INTEGER(C_INT), INTENT(OUT) :: ntmu(Nmu), graz(Nmu), map(2*Ntau, Nmu)
    REAL(C_DOUBLE), INTENT(OUT) :: rhoxmu_(2*Ntau, Nmu), muloc_(2*Ntau, Nmu)
    Radius = radius_
    Atmos%Height = height_
    IF (ALLOCATED(NtauMu)) DEALLOCATE(NtauMu, GrazMu, MapMu, RhoxMu, MuLoc)
    CALL SphereInit
    ntmu = NtauMu
    graz = MERGE(1, 0, GrazMu)
    map = MapMu
    rhoxmu_ = RhoxMu
    muloc_ = MuLoc
  END SUBROUTINE ref_sphere

  ! LocPrfSph for one opacity column on the rays of the last ref_sphere call
  SUBROUTINE ref_locprfsph(ploc, xki, src) BIND(C, NAME='ref_locprfsph')
    REAL(C_DOUBLE), INTENT(OUT) :: ploc(Nmu)
    REAL(C_DOUBLE), INTENT(IN) :: xki(Ntau), src(Ntau)
    CALL LocPrfSph(ploc, xki, src)
  END SUBROUTINE ref_locprfsph

  ! StokesPrfSph at wavelength iwl with the given MagLineOpac/LineOpac columns and the
  ! per-point field geometry (2*Ntau, Nmu) of each angle
  SUBROUTINE ref_stokesprfsph(iwl, mlop, lop, src, sphi2, cphi2, cphi, s2psi, c2psi, stokes) &
                              BIND(C, NAME='ref_stokesprfsph')
    INTEGER(C_INT), VALUE :: iwl
    REAL(C_DOUBLE), INTENT(IN) :: mlop(6, Ntau), lop(Ntau), src(Ntau)
    REAL(C_DOUBLE), INTENT(IN), DIMENSION(2*Ntau, Nmu) :: sphi2, cphi2, cphi, s2psi, c2psi
    REAL(C_DOUBLE), INTENT(OUT) :: stokes(Nmu, 4)
    IF (ALLOCATED(MagLineOpac)) DEALLOCATE(MagLineOpac)
    ALLOCATE(MagLineOpac(6, Ntau, Nwl))
    MagLineOpac = 0d0
    MagLineOpac(:, :, iwl) = mlop
    LineOpac(:, iwl) = lop
    CALL StokesPrfSph(iwl, src, sphi2, cphi2, cphi, s2psi, c2psi, stokes)
  END SUBROUTINE ref_stokesprfsph

  SUBROUTINE ref_solve4x4(a, b, x) BIND(C, NAME='ref_solve4x4')
    REAL(C_DOUBLE), INTENT(IN) :: a(4, 4), b(4)
    REAL(C_DOUBLE), INTENT(OUT) :: x(4)
    CALL Solve4x4(a, b, x)
  END SUBROUTINE ref_solve4x4

  SUBROUTINE ref_zeeman(jlo, jup, glo, gup, npi, spi, rpi, nsig, ssig, rsig) BIND(C, NAME='ref_zeeman')
    REAL(C_FLOAT), VALUE :: jlo, jup, glo, gup
    INTEGER(C_INT), INTENT(OUT) :: npi, nsig
    REAL(C_FLOAT), INTENT(OUT) :: spi(*), rpi(*), ssig(*), rsig(*)
    REAL(4), ALLOCATABLE :: s1(:), r1(:), s2(:), r2(:)
    INTEGER :: n1, n2
    CALL ZeemanPattern(jlo, jup, glo, gup, n1, s1, r1, n2, s2, r2)
    npi = n1
    nsig = n2
    spi(1:n1) = s1
    rpi(1:n1) = r1
    ssig(1:n2) = s2
    rsig(1:n2) = r2
  END SUBROUTINE ref_zeeman

  SUBROUTINE ref_gauss_mu(quad, n, x, w) BIND(C, NAME='ref_gauss_mu')
    INTEGER(C_INT), VALUE :: quad, n
    REAL(C_DOUBLE), INTENT(OUT) :: x(n), w(n)
    CALL GaussMu(quad, n, x, w)
  END SUBROUTINE ref_gauss_mu

  SUBROUTINE ref_mu_quadrature(quad, n, x, w) BIND(C, NAME='ref_mu_quadrature')
    INTEGER(C_INT), VALUE :: quad, n
    REAL(C_DOUBLE), INTENT(INOUT) :: x(n)
    REAL(C_DOUBLE), INTENT(OUT) :: w(n)
    REAL(8) :: r(n + 1)
    CALL MuQuadrature(quad, n, x, w, r, .TRUE.)
  END SUBROUTINE ref_mu_quadrature

  ! DiskInteg (src/convol.f90:67-246) on intensity spectra spl, spc (Nmu, Nwl): sets the
  ! module variables it reads and returns SfL, SfC
  SUBROUTINE ref_diskinteg(nmu_, nwl_, wmode_, wfirst_, wlast_, dstep_, quad, xmu_, vsini_, vmacro_, vgau_, &
                           spl_, spc_, sfl_, sfc_) BIND(C, NAME='ref_diskinteg')
    INTEGER(C_INT), VALUE :: nmu_, nwl_, wmode_, quad
    REAL(C_DOUBLE), VALUE :: wfirst_, wlast_, dstep_, vsini_, vmacro_, vgau_
    REAL(C_DOUBLE), INTENT(IN) :: xmu_(nmu_), spl_(nmu_, nwl_), spc_(nmu_, nwl_)
    REAL(C_DOUBLE), INTENT(OUT) :: sfl_(nwl_), sfc_(nwl_)
    Nmu = nmu_
    Nwl = nwl_
    Wmode = wmode_
    Wfirst = wfirst_
    Wlast = wlast_
    Dstep = dstep_
    DiskQuad = quad
    Vsini = vsini_
    Vmacro = vmacro_
    Vgau = vgau_
    IF (ALLOCATED(Xmu)) DEALLOCATE(Xmu)
    IF (ALLOCATED(Wmu)) DEALLOCATE(Wmu)
    ALLOCATE(Xmu(Nmu), Wmu(Nmu))
    Xmu = xmu_
    IF (ALLOCATED(SpL)) DEALLOCATE(SpL)
    IF (ALLOCATED(SpC)) DEALLOCATE(SpC)
    IF (ALLOCATED(SfL)) DEALLOCATE(SfL)
    IF (ALLOCATED(SfC)) DEALLOCATE(SfC)
    ALLOCATE(SpL(Nmu, Nwl), SpC(Nmu, Nwl))
    SpL = spl_
    SpC = spc_
    CALL DiskInteg
    sfl_ = SfL
    sfc_ = SfC
  END SUBROUTINE ref_diskinteg

  ! ConvolGau (src/convol.f90:249-300) on one spectrum
  SUBROUTINE ref_convolgau(nw, s, fwhm, wmode_, wfirst_, wlast_, dstep_) BIND(C, NAME='ref_convolgau')
    INTEGER(C_INT), VALUE :: nw, wmode_
    REAL(C_DOUBLE), VALUE :: fwhm, wfirst_, wlast_, dstep_
    REAL(C_DOUBLE), INTENT(INOUT) :: s(nw)
    Wmode = wmode_
    Wfirst = wfirst_
    Wlast = wlast_
    Dstep = dstep_
    CALL ConvolGau(nw, s, fwhm)
  END SUBROUTINE ref_convolgau

  REAL(C_DOUBLE) FUNCTION ref_semi_area(r, v) BIND(C, NAME='ref_semi_area')
    REAL(C_DOUBLE), VALUE :: r, v
    ref_semi_area = SemiArea(r, v)
  END FUNCTION ref_semi_area

  ! Vac2Air and the bound searches of ReadHDF (src/utils.f90:701-737, 1007-1017)
  REAL(C_DOUBLE) FUNCTION ref_vac2air(w) BIND(C, NAME='ref_vac2air')
    REAL(C_DOUBLE), VALUE :: w
    ref_vac2air = Vac2Air(w)
  END FUNCTION ref_vac2air

  ! REAL(4) 10.0**x as the reference's compiler evaluates it in scalar code (the line loop
  ! of ReadHDF, src/ioutils.f90:1056-1076); one call per value keeps it scalar
  SUBROUTINE ref_pow10(x, y) BIND(C, NAME='ref_pow10')
    REAL(C_FLOAT), VALUE :: x
    REAL(C_FLOAT), INTENT(OUT) :: y
    y = 10.0**x
  END SUBROUTINE ref_pow10

  SUBROUTINE ref_bounds(n, a, x1, x2, i1, i2) BIND(C, NAME='ref_bounds')
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), INTENT(IN) :: a(n)
    REAL(C_DOUBLE), VALUE :: x1, x2
    INTEGER(C_INT), INTENT(OUT) :: i1, i2
    i1 = LowerBoundReal8(a, x1)
    i2 = UpperBoundReal8(a, x2)
  END SUBROUTINE ref_bounds

  SUBROUTINE ref_mu_quadrature_r(quad, n, x, w, r) BIND(C, NAME='ref_mu_quadrature_r')
    INTEGER(C_INT), VALUE :: quad, n
    REAL(C_DOUBLE), INTENT(INOUT) :: x(n)
    REAL(C_DOUBLE), INTENT(OUT) :: w(n), r(n + 1)
    CALL MuQuadrature(quad, n, x, w, r, .FALSE.)
  END SUBROUTINE ref_mu_quadrature_r

  INTEGER(C_INT) FUNCTION ref_nthreads() BIND(C, NAME='ref_nthreads')
    INTEGER :: n
    n = 1
    !$OMP PARALLEL
    !$OMP MASTER
    !$ n = omp_get_num_threads_f()
    !$OMP END MASTER
    !$OMP END PARALLEL
    ref_nthreads = n
  CONTAINS
    INTEGER FUNCTION omp_get_num_threads_f()
      !$ USE omp_lib, ONLY: omp_get_num_threads
      omp_get_num_threads_f = 1
      !$ omp_get_num_threads_f = omp_get_num_threads()
    END FUNCTION omp_get_num_threads_f
  END FUNCTION ref_nthreads

END MODULE RefHarness
