! fsgpu_host.f90 -- Fortran host side of libfsgpu.so.
!
! Binds the C-ABI of include/fsgpu.h with ISO_C_BINDING interfaces and provides drop-in
! replacements for FastSpec's hot-path drivers.  The module reads the same module globals
! as the routines it replaces (Config, LineData, AtmosData, SpecData, HeStark) and is meant
! to be compiled inside the reference tree, next to src/synth.f90:
!
!   GpuInit            device set-up after Init and ContopCalc (src/fastspec.f90:192,224)
!   GpuLineopCalc      replaces LineopCalc(ifile, NlinesLim)      src/synth.f90:747
!   GpuMagLineopCalc   replaces MagLineopCalc(ifile, NlinesLim)   src/synth.f90:1098
!   GpuRadTranCalc     replaces RadTranCalc                       src/synth.f90:1825
!   GpuMagRadTranCalc  replaces MagRadTranCalc                    src/synth.f90:1904
!   GpuFlux            the flux sums of DiskInteg (CImode 2 without convolution)
!   GpuIntegConvol     replaces IntegConvol(CImode)               src/convol.f90:27
!   GpuGetLineop/GpuSetLineop   LineOpac to/from the host (SAVE flag 2, OPACITY keyword)
!
! The line-opacity drivers leave LinFlag, Wlcent, LinSpc and LinStr allocated when the
! list's line strengths are requested, so that the caller's WriteLinStr (src/ioutils.f90:1826)
! writes fs_linst.bin unchanged; GpuLineopRelease frees them afterwards.
MODULE FsgpuHost

  USE ISO_C_BINDING
  USE Constants
  USE Config
  USE LineData
  USE AtmosData
  USE SpecData
  USE Utils, ONLY: ErrStop, IsMolecule
  USE HeStark, ONLY: HeGam, HeShf, WlcHe, EloHe, NheTab, HeMatchTol, HeEloTol

  IMPLICIT NONE

  INTEGER(C_INT32_T), PARAMETER :: FSG_FP64 = 0, FSG_FP32 = 1

  ! mirror of fsg_setup
  TYPE, BIND(C) :: FsgSetup
    INTEGER(C_INT32_T) :: ntau, nwl, nmu, wmode
    REAL(C_DOUBLE) :: wfirst, dstep, rstep, wstart
    TYPE(C_PTR) :: wave
    INTEGER(C_INT32_T) :: motype
    TYPE(C_PTR) :: rhox, temp, xne, rho, contop, contop_std
    INTEGER(C_INT32_T) :: pf_ld, nslot
    TYPE(C_PTR) :: fraction_pf, pot_ion, xamass, frac_h1, frac_he1, frac_hmol
    INTEGER(C_INT32_T) :: nspc
    TYPE(C_PTR) :: spc_ion, spc_slot
    INTEGER(C_INT32_T) :: he1_slot, nhe
    TYPE(C_PTR) :: he_wl, he_elo
    REAL(C_DOUBLE) :: he_wl_tol, he_elo_tol
    TYPE(C_PTR) :: he_gam, he_shf
    REAL(C_DOUBLE) :: eps_opac, vwfudge, h2brd, vwmol, vmicro
    INTEGER(C_INT32_T) :: molbrd
    TYPE(C_PTR) :: xmu, wmu
    INTEGER(C_INT32_T) :: precision, batched, device
    REAL(C_DOUBLE) :: radius
    TYPE(C_PTR) :: height
  END TYPE FsgSetup

  ! mirror of fsg_lines
  TYPE, BIND(C) :: FsgLines
    INTEGER(C_INT64_T) :: n
    TYPE(C_PTR) :: wl, elo, gf, grad, gqst, gvdw, spc, is_mol
    TYPE(C_PTR) :: pi_off, spl_pi, str_pi, si_off, spl_si, str_si
  END TYPE FsgLines

  ! mirror of fsg_lineop_stats
  TYPE, BIND(C) :: FsgStats
    INTEGER(C_INT64_T) :: nlines_in, nlines_use, nwide, nblocks, nsegments, evals
    REAL(C_DOUBLE) :: ms_pass1, ms_opacity
    INTEGER(C_INT64_T) :: evals_core
    REAL(C_DOUBLE) :: ms_far, ms_core
    INTEGER(C_INT32_T) :: core_tile, far_ring
  END TYPE FsgStats

  INTERFACE
    INTEGER(C_INT) FUNCTION fsg_init(Setup, Ctx) BIND(C, NAME='fsg_init')
      IMPORT :: C_INT, C_PTR, FsgSetup
      TYPE(FsgSetup), INTENT(IN) :: Setup
      TYPE(C_PTR), INTENT(OUT) :: Ctx
    END FUNCTION fsg_init
    INTEGER(C_INT) FUNCTION fsg_finalize(Ctx) BIND(C, NAME='fsg_finalize')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: Ctx
    END FUNCTION fsg_finalize
    TYPE(C_PTR) FUNCTION fsg_last_error() BIND(C, NAME='fsg_last_error')
      IMPORT :: C_PTR
    END FUNCTION fsg_last_error
    INTEGER(C_INT) FUNCTION fsg_lineop(Ctx, ListType, Lines, NlinesLim, LinStrOut, FlagsOut, &
                                       IleftOut, IrightOut, Stats) BIND(C, NAME='fsg_lineop')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, FsgLines, FsgStats
      TYPE(C_PTR), VALUE :: Ctx
      INTEGER(C_INT32_T), VALUE :: ListType
      TYPE(FsgLines), INTENT(IN) :: Lines
      INTEGER(C_INT64_T), VALUE :: NlinesLim
      TYPE(C_PTR), VALUE :: LinStrOut, FlagsOut, IleftOut, IrightOut
      TYPE(FsgStats), INTENT(OUT) :: Stats
    END FUNCTION fsg_lineop
    INTEGER(C_INT) FUNCTION fsg_maglineop(Ctx, Lines, NlinesLim, Bfield, Bframe, LinStrOut, FlagsOut, &
                                          IleftOut, IrightOut, Stats) BIND(C, NAME='fsg_maglineop')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, FsgLines, FsgStats
      TYPE(C_PTR), VALUE :: Ctx
      TYPE(FsgLines), INTENT(IN) :: Lines
      INTEGER(C_INT64_T), VALUE :: NlinesLim
      REAL(C_DOUBLE), INTENT(IN) :: Bfield(*)
      INTEGER(C_INT32_T), VALUE :: Bframe
      TYPE(C_PTR), VALUE :: LinStrOut, FlagsOut, IleftOut, IrightOut
      TYPE(FsgStats), INTENT(OUT) :: Stats
    END FUNCTION fsg_maglineop
    INTEGER(C_INT) FUNCTION fsg_get_lineop(Ctx, Out) BIND(C, NAME='fsg_get_lineop')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: Ctx
      REAL(C_DOUBLE), INTENT(OUT) :: Out(*)
    END FUNCTION fsg_get_lineop
    INTEGER(C_INT) FUNCTION fsg_set_lineop(Ctx, In) BIND(C, NAME='fsg_set_lineop')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: Ctx
      REAL(C_DOUBLE), INTENT(IN) :: In(*)
    END FUNCTION fsg_set_lineop
    INTEGER(C_INT) FUNCTION fsg_radtran(Ctx, SpLOut, SpCOut) BIND(C, NAME='fsg_radtran')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: Ctx
      REAL(C_DOUBLE), INTENT(OUT) :: SpLOut(*), SpCOut(*)
    END FUNCTION fsg_radtran
    INTEGER(C_INT) FUNCTION fsg_magradtran(Ctx, Bfield, Nbfield, Bframe, SpZOut, SpCOut) &
                                           BIND(C, NAME='fsg_magradtran')
      IMPORT :: C_INT, C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: Ctx
      REAL(C_DOUBLE), INTENT(IN) :: Bfield(*)
      INTEGER(C_INT32_T), VALUE :: Nbfield, Bframe
      REAL(C_DOUBLE), INTENT(OUT) :: SpZOut(*), SpCOut(*)
    END FUNCTION fsg_magradtran
    INTEGER(C_INT) FUNCTION fsg_flux(Ctx, SfLOut, SfCOut) BIND(C, NAME='fsg_flux')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: Ctx
      REAL(C_DOUBLE), INTENT(OUT) :: SfLOut(*), SfCOut(*)
    END FUNCTION fsg_flux
    INTEGER(C_INT) FUNCTION fsg_integ_convol(Ctx, Cimode, Vsini_, Vmacro_, Vgau_) BIND(C, NAME='fsg_integ_convol')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: Ctx
      INTEGER(C_INT32_T), VALUE :: Cimode
      REAL(C_DOUBLE), VALUE :: Vsini_, Vmacro_, Vgau_
    END FUNCTION fsg_integ_convol
    INTEGER(C_INT) FUNCTION fsg_get_spectra(Ctx, SpLOut, SpZOut, SpCOut) BIND(C, NAME='fsg_get_spectra')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: Ctx, SpLOut, SpZOut, SpCOut
    END FUNCTION fsg_get_spectra
    INTEGER(C_SIZE_T) FUNCTION c_strlen(S) BIND(C, NAME='strlen')
      IMPORT :: C_SIZE_T, C_PTR
      TYPE(C_PTR), VALUE :: S
    END FUNCTION c_strlen
  END INTERFACE

  TYPE(C_PTR), SAVE :: GpuCtx = C_NULL_PTR
  INTEGER, SAVE :: GpuPrecision = FSG_FP32

CONTAINS

  ! Message of the last failed library call
  FUNCTION GpuError() RESULT(Msg)
    CHARACTER(:), ALLOCATABLE :: Msg
    TYPE(C_PTR) :: p
    CHARACTER(KIND=C_CHAR), POINTER :: c(:)
    INTEGER :: n, k
    p = fsg_last_error()
    n = INT(c_strlen(p))
    ALLOCATE(CHARACTER(n) :: Msg)
    IF (n == 0) RETURN
    CALL C_F_POINTER(p, c, [n])
    DO k = 1, n
      Msg(k:k) = c(k)
    END DO
  END FUNCTION GpuError

  SUBROUTINE GpuCheck(Rc)
    INTEGER(C_INT), INTENT(IN) :: Rc
    IF (Rc /= 0) CALL ErrStop(GpuError())
  END SUBROUTINE GpuCheck

  ! Device set-up from the host's module globals, after Init and ContopCalc.
  ! Precision: FSG_FP32 (batched path) or FSG_FP64 (plain path).
  SUBROUTINE GpuInit(Precision, Device)
    INTEGER, INTENT(IN), OPTIONAL :: Precision, Device
    REAL(8), ALLOCATABLE, TARGET :: Rhox(:), T(:), Xne(:), Rho(:), Cstd(:), HeW(:), HeE(:), HeG(:,:), HeS(:,:)
    REAL(8), ALLOCATABLE, TARGET :: Hgt(:)
    REAL(4), ALLOCATABLE, TARGET :: Fh1(:), Fhe1(:), Fhmol(:)
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: SpcIon(:), SpcSlot(:)
    INTEGER :: k, Nspc_, Hmol
    INTEGER :: Prec, Dev

    Prec = FSG_FP32
    IF (PRESENT(Precision)) Prec = Precision
    Dev = 0
    IF (PRESENT(Device)) Dev = Device
    GpuPrecision = Prec

    ALLOCATE(Rhox(Ntau), T(Ntau), Xne(Ntau), Rho(Ntau), Cstd(Ntau), Fh1(Ntau), Fhe1(Ntau), Fhmol(Ntau))
    ALLOCATE(Hgt(Ntau))
    DO k = 1, Ntau
      Hgt(k) = Atmos(k)%Height
      Rhox(k) = Atmos(k)%Rhox
      T(k) = Atmos(k)%T
      Xne(k) = Atmos(k)%Xne
      Rho(k) = Atmos(k)%Rho
      Fh1(k) = Fraction(IdxSpc(Ix%H1), k)
      Fhe1(k) = Fraction(IdxSpc(Ix%He1), k)
      Hmol = IdxSpc(Ix%Hmol)
      Fhmol(k) = 0.0
      IF (Hmol > 0) Fhmol(k) = Fraction(Hmol, k)
    END DO
    Cstd = 1d0
    IF (Motype == 1) Cstd = ContOpacStd
    Nspc_ = SIZE(Ion)
    ALLOCATE(SpcIon(Nspc_), SpcSlot(Nspc_))
    SpcIon = Ion
    SpcSlot = IdxSpc
    ALLOCATE(HeW(NheTab), HeE(NheTab), HeG(Ntau, NheTab), HeS(Ntau, NheTab))
    HeW = WlcHe
    HeE = EloHe
    HeG = HeGam
    HeS = HeShf
    CALL GpuInitArrays(Wave, Rhox, T, Xne, Rho, ContOpac, Cstd, FractionPF, PotIon, Xamass, &
                       Fh1, Fhe1, Fhmol, SpcIon, SpcSlot, HeW, HeE, HeG, HeS, Xmu, Wmu, Hgt, Prec, Dev)

  CONTAINS

    ! the C pointers are taken from TARGET dummies and used only during this call
    SUBROUTINE GpuInitArrays(Wave_, Rhox_, T_, Xne_, Rho_, Cont_, Cstd_, Fpf_, Pot_, Xam_, &
                             Fh1_, Fhe1_, Fhmol_, Ion_, Slot_, HeW_, HeE_, HeG_, HeS_, Xmu_, Wmu_, Hgt_, &
                             Prec_, Dev_)
      REAL(8), INTENT(IN), TARGET :: Wave_(*), Rhox_(*), T_(*), Xne_(*), Rho_(*), Cont_(*), Cstd_(*)
      REAL(8), INTENT(IN), TARGET :: HeW_(*), HeE_(*), HeG_(*), HeS_(*), Xmu_(*), Wmu_(*), Hgt_(*)
      REAL(4), INTENT(IN), TARGET :: Fpf_(*), Pot_(*), Xam_(*), Fh1_(*), Fhe1_(*), Fhmol_(*)
      INTEGER(C_INT32_T), INTENT(IN), TARGET :: Ion_(*), Slot_(*)
      INTEGER, INTENT(IN) :: Prec_, Dev_
      TYPE(FsgSetup) :: S

      S%ntau = Ntau
      S%nwl = Nwl
      S%nmu = Nmu
      S%wmode = Wmode
      S%wfirst = Wfirst
      S%dstep = Dstep
      S%rstep = 0d0
      IF (Wmode == 1) S%rstep = Rstep
      S%wstart = Wstart
      S%wave = C_LOC(Wave_)
      S%motype = Motype
      S%rhox = C_LOC(Rhox_)
      S%temp = C_LOC(T_)
      S%xne = C_LOC(Xne_)
      S%rho = C_LOC(Rho_)
      S%contop = C_LOC(Cont_)
      S%contop_std = C_LOC(Cstd_)
      S%pf_ld = SIZE(FractionPF, 1)
      S%nslot = SIZE(PotIon)
      S%fraction_pf = C_LOC(Fpf_)
      S%pot_ion = C_LOC(Pot_)
      S%xamass = C_LOC(Xam_)
      S%frac_h1 = C_LOC(Fh1_)
      S%frac_he1 = C_LOC(Fhe1_)
      S%frac_hmol = C_LOC(Fhmol_)
      S%nspc = SIZE(Ion)
      S%spc_ion = C_LOC(Ion_)
      S%spc_slot = C_LOC(Slot_)
      S%he1_slot = IdxSpc(Ix%He1)
      S%nhe = NheTab
      S%he_wl = C_LOC(HeW_)
      S%he_elo = C_LOC(HeE_)
      S%he_wl_tol = HeMatchTol
      S%he_elo_tol = HeEloTol
      S%he_gam = C_LOC(HeG_)
      S%he_shf = C_LOC(HeS_)
      S%eps_opac = EpsOpac
      S%vwfudge = VWfudge
      S%h2brd = H2Brd
      S%vwmol = VWmol
      S%vmicro = Vmicro
      S%molbrd = MERGE(1, 0, MolBrd)
      S%xmu = C_LOC(Xmu_)
      S%wmu = C_LOC(Wmu_)
      S%precision = Prec_
      S%batched = MERGE(1, 0, Prec_ == FSG_FP32)
      S%device = Dev_
      ! spherical models: Radius > 0 and the layer heights (src/solver.f90:51)
      S%radius = Radius
      S%height = C_NULL_PTR
      IF (Radius > 0d0) S%height = C_LOC(Hgt_)
      CALL GpuCheck(fsg_init(S, GpuCtx))
    END SUBROUTINE GpuInitArrays

  END SUBROUTINE GpuInit

  SUBROUTINE GpuFinalize
    INTEGER(C_INT) :: rc
    IF (C_ASSOCIATED(GpuCtx)) rc = fsg_finalize(GpuCtx)
    GpuCtx = C_NULL_PTR
  END SUBROUTINE GpuFinalize

  ! Structure-of-arrays copy of the current line list (Lines => LinesA/M/Z)
  SUBROUTINE LinesSoA(ifile, Wl, Elo, Gf, Grad, Gqst, Gvdw, Spc, IsMol)
    INTEGER, INTENT(IN) :: ifile
    REAL(8), ALLOCATABLE, INTENT(OUT) :: Wl(:)
    REAL(4), ALLOCATABLE, INTENT(OUT) :: Elo(:), Gf(:), Grad(:), Gqst(:), Gvdw(:)
    INTEGER(C_INT32_T), ALLOCATABLE, INTENT(OUT) :: Spc(:)
    INTEGER(C_INT8_T), ALLOCATABLE, INTENT(OUT) :: IsMol(:)
    INTEGER :: i

    ALLOCATE(Wl(Nlines), Elo(Nlines), Gf(Nlines), Grad(Nlines), Gqst(Nlines), Gvdw(Nlines), &
             Spc(Nlines), IsMol(Nlines))
    Gqst = 0.0
    Gvdw = 0.0
    IsMol = 0
    DO i = 1, Nlines
      Wl(i) = Lines(i)%Wl
      Elo(i) = Lines(i)%Elo
      Gf(i) = Lines(i)%Gf
      Grad(i) = Lines(i)%Grad
    END DO
    SELECT CASE (LineListType(ifile))
      CASE (0)
        Gqst = LinesA%Gqst
        Gvdw = LinesA%Gvdw
      CASE (2)
        Gqst = LinesZ%Gqst
        Gvdw = LinesZ%Gvdw
        DO i = 1, Nlines
          IF (IsMolecule(LinesZ(i)%El)) IsMol(i) = 1
        END DO
    END SELECT
    ! line -> global species index (src/synth.f90:797-810)
    IF (ALLOCATED(LinSpcAll(ifile)%Idx)) THEN
      IF (SIZE(LinSpcAll(ifile)%Idx) /= Nlines) CALL ErrStop('GpuLineopCalc: line-to-species map size mismatch')
      Spc = LinSpcAll(ifile)%Idx
    ELSE
      Spc = SUM(Nspc(1:ifile))
    END IF
  END SUBROUTINE LinesSoA

  ! Replaces LineopCalc(ifile, NlinesLim): the list's line opacity is added to the
  ! device-resident LineOpac.  With DoLinStr the per-line arrays that WriteLinStr reads
  ! are left allocated.
  SUBROUTINE GpuLineopCalc(ifile, NlinesLim, DoLinStr)
    INTEGER, INTENT(IN) :: ifile, NlinesLim
    LOGICAL, INTENT(OUT) :: DoLinStr
    REAL(8), ALLOCATABLE, TARGET :: Wl(:)
    REAL(4), ALLOCATABLE, TARGET :: Elo(:), Gf(:), Grad(:), Gqst(:), Gvdw(:), Str(:)
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: Spc(:), Flags(:), Il(:), Ir(:)
    INTEGER(C_INT8_T), ALLOCATABLE, TARGET :: IsMol(:)
    TYPE(FsgLines) :: L
    TYPE(FsgStats) :: St
    TYPE(C_PTR) :: pstr

    DoLinStr = ((LineListType(ifile) == 0 .OR. LineListType(ifile) == 2) .AND. OutFlags(1) >= 1) .OR. &
               (LineListType(ifile) == 1 .AND. OutFlags(1) == 2)
    CALL LinesSoA(ifile, Wl, Elo, Gf, Grad, Gqst, Gvdw, Spc, IsMol)
    ALLOCATE(Flags(Nlines), Il(Nlines), Ir(Nlines), Str(Nlines))
    L%n = Nlines
    L%wl = C_LOC(Wl)
    L%elo = C_LOC(Elo)
    L%gf = C_LOC(Gf)
    L%grad = C_LOC(Grad)
    L%gqst = C_LOC(Gqst)
    L%gvdw = C_LOC(Gvdw)
    L%spc = C_LOC(Spc)
    L%is_mol = C_NULL_PTR
    IF (LineListType(ifile) == 2) L%is_mol = C_LOC(IsMol)
    L%pi_off = C_NULL_PTR
    L%spl_pi = C_NULL_PTR
    L%str_pi = C_NULL_PTR
    L%si_off = C_NULL_PTR
    L%spl_si = C_NULL_PTR
    L%str_si = C_NULL_PTR
    pstr = C_NULL_PTR
    IF (DoLinStr) pstr = C_LOC(Str)
    CALL GpuCheck(fsg_lineop(GpuCtx, INT(LineListType(ifile), C_INT32_T), L, INT(NlinesLim, C_INT64_T), &
                             pstr, C_LOC(Flags), C_LOC(Il), C_LOC(Ir), St))
    NlinesUse = INT(St%nlines_use)
    IF (Verbose >= 1 .AND. St%nwide > 0) WRITE(*, '(A, I0, A)', ADVANCE='NO') 'Wide lines: ', St%nwide, ' :: '
    IF (DoLinStr) CALL KeepForWriteLinStr(Wl, Flags, Spc, Str)
  END SUBROUTINE GpuLineopCalc

  ! Replaces MagLineopCalc(ifile, NlinesLim) for a Zeeman (.z) list with BFIELD set.
  SUBROUTINE GpuMagLineopCalc(ifile, NlinesLim, DoLinStr)
    INTEGER, INTENT(IN) :: ifile, NlinesLim
    LOGICAL, INTENT(OUT) :: DoLinStr
    REAL(8), ALLOCATABLE, TARGET :: Wl(:), Bf(:)
    REAL(4), ALLOCATABLE, TARGET :: Elo(:), Gf(:), Grad(:), Gqst(:), Gvdw(:), Str(:)
    REAL(4), ALLOCATABLE, TARGET :: SplPi(:), StrPi(:), SplSi(:), StrSi(:)
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: Spc(:), Flags(:), Il(:), Ir(:), PiOff(:), SiOff(:)
    INTEGER(C_INT8_T), ALLOCATABLE, TARGET :: IsMol(:)
    TYPE(FsgLines) :: L
    TYPE(FsgStats) :: St
    TYPE(C_PTR) :: pstr
    INTEGER :: i, np, ns

    DoLinStr = OutFlags(1) >= 1
    CALL LinesSoA(ifile, Wl, Elo, Gf, Grad, Gqst, Gvdw, Spc, IsMol)
    ! Zeeman patterns in CSR form (src/shared.f90:473-481)
    ALLOCATE(PiOff(Nlines + 1), SiOff(Nlines + 1))
    PiOff(1) = 0
    SiOff(1) = 0
    DO i = 1, Nlines
      PiOff(i + 1) = PiOff(i) + LinesZ(i)%Npi
      SiOff(i + 1) = SiOff(i) + LinesZ(i)%Nsi
    END DO
    ALLOCATE(SplPi(MAX(1, PiOff(Nlines + 1))), StrPi(MAX(1, PiOff(Nlines + 1))))
    ALLOCATE(SplSi(MAX(1, SiOff(Nlines + 1))), StrSi(MAX(1, SiOff(Nlines + 1))))
    DO i = 1, Nlines
      np = LinesZ(i)%Npi
      ns = LinesZ(i)%Nsi
      SplPi(PiOff(i) + 1:PiOff(i) + np) = LinesZ(i)%SplPi
      StrPi(PiOff(i) + 1:PiOff(i) + np) = LinesZ(i)%StrPi
      SplSi(SiOff(i) + 1:SiOff(i) + ns) = LinesZ(i)%SplSi
      StrSi(SiOff(i) + 1:SiOff(i) + ns) = LinesZ(i)%StrSi
    END DO
    ALLOCATE(Flags(Nlines), Il(Nlines), Ir(Nlines), Str(Nlines), Bf(SIZE(Bfield)))
    Bf = Bfield
    L%n = Nlines
    L%wl = C_LOC(Wl)
    L%elo = C_LOC(Elo)
    L%gf = C_LOC(Gf)
    L%grad = C_LOC(Grad)
    L%gqst = C_LOC(Gqst)
    L%gvdw = C_LOC(Gvdw)
    L%spc = C_LOC(Spc)
    L%is_mol = C_LOC(IsMol)
    L%pi_off = C_LOC(PiOff)
    L%spl_pi = C_LOC(SplPi)
    L%str_pi = C_LOC(StrPi)
    L%si_off = C_LOC(SiOff)
    L%spl_si = C_LOC(SplSi)
    L%str_si = C_LOC(StrSi)
    pstr = C_NULL_PTR
    IF (DoLinStr) pstr = C_LOC(Str)
    CALL GpuCheck(fsg_maglineop(GpuCtx, L, INT(NlinesLim, C_INT64_T), Bf, INT(Bframe, C_INT32_T), &
                                pstr, C_LOC(Flags), C_LOC(Il), C_LOC(Ir), St))
    NlinesUse = INT(St%nlines_use)
    IF (Verbose >= 1 .AND. St%nwide > 0) WRITE(*, '(A, I0, A)', ADVANCE='NO') 'Wide lines: ', St%nwide, ' :: '
    IF (DoLinStr) CALL KeepForWriteLinStr(Wl, Flags, Spc, Str)
  END SUBROUTINE GpuMagLineopCalc

  ! LinFlag, Wlcent, LinSpc, LinStr as WriteLinStr reads them (src/ioutils.f90:1836-1866)
  SUBROUTINE KeepForWriteLinStr(Wl, Flags, Spc, Str)
    REAL(8), INTENT(IN) :: Wl(:)
    INTEGER(C_INT32_T), INTENT(IN) :: Flags(:), Spc(:)
    REAL(4), INTENT(IN) :: Str(:)
    CALL GpuLineopRelease
    ALLOCATE(LinFlag(Nlines), Wlcent(Nlines), LinStr(Nlines))
    IF (ALLOCATED(LinSpc)) DEALLOCATE(LinSpc)
    ALLOCATE(LinSpc(Nlines))
    LinFlag = Flags
    Wlcent = Wl - Wfirst
    LinSpc = Spc
    LinStr = Str
  END SUBROUTINE KeepForWriteLinStr

  SUBROUTINE GpuLineopRelease
    IF (ALLOCATED(LinFlag)) DEALLOCATE(LinFlag)
    IF (ALLOCATED(Wlcent)) DEALLOCATE(Wlcent)
    IF (ALLOCATED(LinStr)) DEALLOCATE(LinStr)
  END SUBROUTINE GpuLineopRelease

  ! Replaces RadTranCalc: allocates and fills SpC, SpL(Nmu, Nwl)
  SUBROUTINE GpuRadTranCalc
    IF (ALLOCATED(SpC)) DEALLOCATE(SpC)
    IF (ALLOCATED(SpL)) DEALLOCATE(SpL)
    ALLOCATE(SpC(Nmu, Nwl), SpL(Nmu, Nwl))
    CALL GpuCheck(fsg_radtran(GpuCtx, SpL, SpC))
  END SUBROUTINE GpuRadTranCalc

  ! Replaces MagRadTranCalc: allocates and fills SpC(Nmu, Nwl), SpZ(Nmu, 4, Nwl)
  SUBROUTINE GpuMagRadTranCalc
    IF (ALLOCATED(SpC)) DEALLOCATE(SpC)
    IF (ALLOCATED(SpZ)) DEALLOCATE(SpZ)
    ALLOCATE(SpC(Nmu, Nwl), SpZ(Nmu, 4, Nwl))
    CALL GpuCheck(fsg_magradtran(GpuCtx, Bfield, INT(SIZE(Bfield), C_INT32_T), INT(Bframe, C_INT32_T), SpZ, SpC))
  END SUBROUTINE GpuMagRadTranCalc

  ! Flux sums Sf = sum_mu Wmu I_mu of the last formal solution (no convolution)
  SUBROUTINE GpuFlux
    IF (ALLOCATED(SfL)) DEALLOCATE(SfL)
    IF (ALLOCATED(SfC)) DEALLOCATE(SfC)
    ALLOCATE(SfL(Nwl), SfC(Nwl))
    CALL GpuCheck(fsg_flux(GpuCtx, SfL, SfC))
  END SUBROUTINE GpuFlux

  ! Replaces IntegConvol(CImode) with the module's Vsini, Vmacro, Vgau: CImode 1 broadens the
  ! intensity spectra (SpL, SpC or SpZ, SpC) with the Gaussian instrumental profile, CImode 2
  ! disk-integrates them into SfL, SfC with rotation, macroturbulence and the instrumental
  ! profile (Stokes I after the polarised solution, src/fastspec.f90:437-440)
  SUBROUTINE GpuIntegConvol(CImode)
    INTEGER, INTENT(IN) :: CImode
    CALL GpuCheck(fsg_integ_convol(GpuCtx, INT(CImode, C_INT32_T), Vsini, Vmacro, Vgau))
    IF (CImode == 2) THEN
      CALL GpuFlux
    ELSE IF (Magnetic) THEN
      CALL GetSpectra(SpZ, SpC, .TRUE.)
    ELSE
      CALL GetSpectra(SpL, SpC, .FALSE.)
    END IF
  CONTAINS
    SUBROUTINE GetSpectra(SpOut, SpcOut, Polarised)
      REAL(8), INTENT(OUT), TARGET :: SpOut(*), SpcOut(*)
      LOGICAL, INTENT(IN) :: Polarised
      IF (Polarised) THEN
        CALL GpuCheck(fsg_get_spectra(GpuCtx, C_NULL_PTR, C_LOC(SpOut), C_LOC(SpcOut)))
      ELSE
        CALL GpuCheck(fsg_get_spectra(GpuCtx, C_LOC(SpOut), C_NULL_PTR, C_LOC(SpcOut)))
      END IF
    END SUBROUTINE GetSpectra
  END SUBROUTINE GpuIntegConvol

  ! LineOpac(Ntau, Nwl) from the device (SAVE flag 2) and to it (OPACITY keyword)
  SUBROUTINE GpuGetLineop
    CALL GpuCheck(fsg_get_lineop(GpuCtx, LineOpac))
  END SUBROUTINE GpuGetLineop

  SUBROUTINE GpuSetLineop
    CALL GpuCheck(fsg_set_lineop(GpuCtx, LineOpac))
  END SUBROUTINE GpuSetLineop

END MODULE FsgpuHost
