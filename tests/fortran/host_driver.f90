! host_driver.f90 -- test driver of the Fortran host binding (fastspec_amd/fortran/fsgpu_host.f90).
!
! TEST INFRASTRUCTURE.  Reads a case file written by tests/fortran_case.py into the
! reference's module globals (as Init/EosSolve/ContopCalc would leave them), runs
! GpuInit, GpuLineopCalc, GpuRadTranCalc and GpuFlux, and writes the flux, the line flags
! and the line strengths to out.bin.  It then leaves the results, fetched into the module
! arrays by the binding, in the reference's output files fs_lineop.bin, fs_spec.bin,
! fs_flux.bin and fs_linst.bin, with the unformatted stream records of WriteOut and
! WriteLinStr (src/ioutils.f90:1751-1875; that file needs the HDF5 Fortran modules, which
! this image lacks, so its records are repeated here).
PROGRAM HostDriver
  USE Constants
  USE Config
  USE LineData
  USE AtmosData
  USE SpecData
  USE HeStark, ONLY: HeStarkInit
  USE FsgpuHost
  IMPLICIT NONE
  INTEGER :: hdr(14), k, prec
  REAL(8) :: dh(9)
  REAL(8), ALLOCATABLE :: a_rhox(:), a_t(:), a_xne(:), a_rho(:), l_wl(:)
  REAL(4), ALLOCATABLE :: l_elo(:), l_gf(:), l_grad(:), l_gqst(:), l_gvdw(:)
  INTEGER, ALLOCATABLE :: l_spc(:)
  INTEGER :: nslot, nspc_, nl, ltype
  LOGICAL :: DoLinStr
  CHARACTER(256) :: arg
  INTEGER, ALLOCATABLE :: keep(:)
  INTEGER :: i

  CALL GET_COMMAND_ARGUMENT(1, arg)
  prec = FSG_FP32
  IF (COMMAND_ARGUMENT_COUNT() >= 2) THEN
    CALL GET_COMMAND_ARGUMENT(2, arg)
    READ(arg, *) prec
    CALL GET_COMMAND_ARGUMENT(1, arg)
  END IF
  OPEN(10, FILE=TRIM(arg), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(10) hdr
  READ(10) dh
  Ntau = hdr(1); Nwl = hdr(2); Nmu = hdr(3); Wmode = hdr(4); Motype = hdr(5)
  nslot = hdr(6); nspc_ = hdr(7); nl = hdr(8); ltype = hdr(9)
  Wfirst = dh(1); Dstep = dh(2); Rstep = dh(3); Wstart = dh(4)
  EpsOpac = dh(5); VWfudge = dh(6); H2Brd = dh(7); VWmol = dh(8); Vmicro = dh(9)
  ALLOCATE(Wave(Nwl), a_rhox(Ntau), a_t(Ntau), a_xne(Ntau), a_rho(Ntau))
  ALLOCATE(ContOpac(Ntau, Nwl), ContOpacStd(Ntau), LineOpac(Ntau, Nwl), Xmu(Nmu), Wmu(Nmu))
  READ(10) Wave, a_rhox, a_t, a_xne, a_rho, ContOpac, ContOpacStd, Xmu, Wmu
  ALLOCATE(Atmos(Ntau))
  DO k = 1, Ntau
    Atmos(k)%Rhox = a_rhox(k); Atmos(k)%T = a_t(k); Atmos(k)%Xne = a_xne(k); Atmos(k)%Rho = a_rho(k)
    Atmos(k)%Xna = 0d0
  END DO
  ALLOCATE(FractionPF(nslot, Ntau), Fraction(nslot, Ntau), PotIon(nslot), Xamass(nslot))
  READ(10) FractionPF, Fraction, PotIon, Xamass
  ALLOCATE(Ion(nspc_ + 5), IdxSpc(nspc_ + 5))
  READ(10) Ion(1:nspc_), IdxSpc(1:nspc_)
  Ion(nspc_+1:nspc_+5) = [1, 1, 1, 2, 2]
  IdxSpc(nspc_+1:nspc_+5) = hdr(10:14)
  Ix%H1 = nspc_ + 1; Ix%He1 = nspc_ + 2; Ix%Hmol = nspc_ + 3; Ix%H2 = nspc_ + 4; Ix%He2 = nspc_ + 5
  ALLOCATE(l_wl(nl), l_elo(nl), l_gf(nl), l_grad(nl), l_gqst(nl), l_gvdw(nl), l_spc(nl))
  READ(10) l_wl, l_elo, l_gf, l_grad, l_gqst, l_gvdw, l_spc
  CLOSE(10)
  LineOpac = 0d0
  Nlines = nl
  Nfiles = 1
  ALLOCATE(LineListType(1), LinSpcAll(1), Nspc(1))
  LineListType(1) = ltype
  Nspc(1) = 0
  LinSpcAll(1)%Idx = l_spc
  ALLOCATE(Elemc(nspc_ + 5))
  Elemc = 'Fe'
  ALLOCATE(LinesA(nl))
  DO k = 1, nl
    LinesA(k)%Wl = l_wl(k); LinesA(k)%Elo = l_elo(k); LinesA(k)%Gf = l_gf(k); LinesA(k)%Grad = l_grad(k)
    LinesA(k)%Gqst = l_gqst(k); LinesA(k)%Gvdw = l_gvdw(k); LinesA(k)%Ion = 1; LinesA(k)%El = 'Fe'
  END DO
  Lines => LinesA
  OutFlags = [1, 0, 0, 1, 1]
  CALL HeStarkInit

  ! optional arguments 3-5: Vsini, Vmacro, Vgau for a disk integration with convolution
  IF (COMMAND_ARGUMENT_COUNT() >= 5) THEN
    CALL GET_COMMAND_ARGUMENT(3, arg)
    READ(arg, *) Vsini
    CALL GET_COMMAND_ARGUMENT(4, arg)
    READ(arg, *) Vmacro
    CALL GET_COMMAND_ARGUMENT(5, arg)
    READ(arg, *) Vgau
  END IF

  CALL GpuInit(prec)
  CALL GpuLineopCalc(1, nl, DoLinStr)
  CALL GpuRadTranCalc
  IF (Vsini > 0d0 .OR. Vmacro > 0d0 .OR. Vgau > 0d0) THEN
    CALL GpuIntegConvol(2)
  ELSE
    CALL GpuFlux
  END IF
  OPEN(11, FILE='out.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(11) Nwl, NlinesUse, MERGE(1, 0, DoLinStr)
  WRITE(11) SfL, SfC, SpL
  IF (DoLinStr) WRITE(11) LinFlag, LinStr
  CLOSE(11)
  CALL GpuGetLineop
  OPEN(12, FILE='fs_lineop.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(12) Nwl, Ntau
  WRITE(12) Wave + Wfirst, LineOpac
  WRITE(12) Atmos%Rhox, Atmos%T, Atmos%Xne, Atmos%Xna, Atmos%Rho
  CLOSE(12)
  OPEN(12, FILE='fs_spec.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(12) Nwl, Nmu
  WRITE(12) Xmu
  WRITE(12) Wave + Wfirst
  WRITE(12) SpL, SpC
  CLOSE(12)
  OPEN(12, FILE='fs_flux.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(12) Nwl
  WRITE(12) Wave + Wfirst
  WRITE(12) SfL, SfC
  CLOSE(12)
  IF (DoLinStr) THEN
    keep = PACK([(i, i=1, Nlines)], .NOT. BTEST(LinFlag, 0))
    OPEN(12, FILE='fs_linst.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(12) SIZE(keep)
    WRITE(12) Elemc(LinSpc(keep))
    WRITE(12) Ion(LinSpc(keep))
    WRITE(12) Wlcent(keep) + Wfirst
    WRITE(12) LinStr(keep)
    CLOSE(12)
  END IF
  CALL GpuLineopRelease
  CALL GpuFinalize
  WRITE(*, '(A, I0)') 'host_driver: done, lines used ', NlinesUse
END PROGRAM HostDriver
