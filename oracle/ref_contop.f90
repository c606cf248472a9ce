! ref_contop.f90 -- TEST INFRASTRUCTURE: a C-callable harness around the reference's own
! continuum opacity (Xcont and ContopInit of src/contop.f90, compiled unchanged from the
! reference sources by oracle/Makefile).  It sets the Contop module's depth state as
! ContopSetup does (src/synth.f90:574-639) for a small EOS table, then evaluates Xcont at the
! wavelengths given, as ContopCalc's loop does (src/synth.f90:663-669, without Air2Vac: the
! caller passes vacuum wavelengths).
MODULE RefContop
  USE ISO_C_BINDING
  USE Constants, ONLY: Kbolt, Hc, erg2ev
  USE AtmosData, ONLY: Ntau, Fraction, FractionPF
  USE Contop, ONLY: Xcont, ContopInit, CT => T, CTlog => Tlog, CTkev => Tkev, CTk => Tk, CHkt => Hkt, &
    CEhvkt => Ehvkt, CStim => Stim, CXne => Xne, CXna => Xna, CRho => Rho, Ifop, IxH1, IxH2, IxHe1, &
    IxHe2, IxHe3, IxHmin, IxHmol, IxC1, IxC2, IxN1, IxO1, IxMg1, IxMg2, IxAl1, IxSi1, IxSi2, IxCa2, &
    IxFe1, IxCH, IxOH, IxNH, IxHot
  IMPLICIT NONE

CONTAINS

  ! EOS slots of the table: 1 H I, 2 H II, 3 H-, 4 He I, 5 H2, 6 He II, 7 He III, 8 C I,
  ! 9 Mg I, 10 Al I, 11 Si I, 12 Fe I, 13 CH, 14 OH, 15 NH, 16 every other species (empty),
  ! 17 C II, 18 Mg II, 19 Si II, 20 N I, 21 O I, 22 Ca II, and from 23 on HOTOP's stages I-VI
  ! of C, N, O, Ne, Mg, Si, S, Fe (slot 22 + 6 (element - 1) + stage).  fpf, frac: FractionPF
  ! and Fraction (70, ntau).  ifop_(20): the source switches (OpFlags).  out(ntau, n): Copac
  ! at each wavelength wl (vacuum, A).
  SUBROUTINE ref_xcont(ntau_, temp, xne, xna, rho, fpf, frac, ifop_, n, wl, out) BIND(C, NAME='ref_xcont')
    INTEGER(C_INT), VALUE :: ntau_, n
    REAL(C_DOUBLE), INTENT(IN) :: temp(ntau_), xne(ntau_), xna(ntau_), rho(ntau_), wl(n)
    REAL(C_FLOAT), INTENT(IN) :: fpf(70, ntau_), frac(70, ntau_)
    INTEGER(C_INT), INTENT(IN) :: ifop_(20)
    REAL(C_DOUBLE), INTENT(OUT) :: out(ntau_, n)
    REAL(8) :: Copac(ntau_)
    INTEGER :: i, k

    Ntau = ntau_
    IF (ALLOCATED(Fraction)) DEALLOCATE(Fraction)
    IF (ALLOCATED(FractionPF)) DEALLOCATE(FractionPF)
    ALLOCATE(Fraction(70, Ntau), FractionPF(70, Ntau))
    Fraction = frac
    FractionPF = fpf
    IF (ALLOCATED(CT)) DEALLOCATE(CT, CTlog, CTkev, CTk, CHkt, CEhvkt, CStim, CXne, CXna, CRho)
    ALLOCATE(CT(Ntau), CTlog(Ntau), CTkev(Ntau), CTk(Ntau), CHkt(Ntau), CEhvkt(Ntau), CStim(Ntau), &
             CXne(Ntau), CXna(Ntau), CRho(Ntau))
    CT = temp
    CXne = xne
    CXna = xna
    CRho = rho
    CTlog = LOG(CT)
    CTk = Kbolt * CT
    CTkev = CTk * erg2ev
    CHkt = Hc / CTk
    Ifop = ifop_
    IxH1 = 1; IxH2 = 2; IxHmin = 3; IxHe1 = 4; IxHmol = 5; IxHe2 = 6; IxHe3 = 7
    IxC1 = 8; IxMg1 = 9; IxAl1 = 10; IxSi1 = 11; IxFe1 = 12; IxCH = 13; IxOH = 14; IxNH = 15
    IxC2 = 17; IxMg2 = 18; IxSi2 = 19; IxN1 = 20; IxO1 = 21; IxCa2 = 22
    DO k = 1, 8
      DO i = 1, 6
        IxHot(i, k) = 22 + 6 * (k - 1) + i
      END DO
    END DO
    CALL ContopInit
    DO i = 1, n
      CALL Xcont(wl(i), Copac)
      out(:, i) = Copac
    END DO
  END SUBROUTINE ref_xcont

END MODULE RefContop
