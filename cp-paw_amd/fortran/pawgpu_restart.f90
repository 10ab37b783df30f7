!
! Restart records of the wave-function step in the reference's restart-file
! format, on host arrays (no GPU calls; WAVESGPU$WRITE / WAVESGPU$READ in
! pawgpu.f90 move the arrays between these routines and the GPU set).
!
! The file is Fortran unformatted sequential, as the reference's
! WRITERESTART / READRESTART (src/paw_ionew.f90:12-275) use it: separator
! records (SEPARATOR_TYPE, src/paw_report.f90:107-125) open each section and
! carry its record count. The WAVES section (WAVES_WRITEPSI,
! src/paw_waves2.f90:3168-3350, and WAVES_WRITELAMBDA, :3920-4040), here for
! one k-point and one spin:
!   NKPT, NSPIN, RBAS(3,3), NWAVE
!   KEY='PSI', NGG, NDIM, NB, ILOGICAL (bit 0: super wave functions)
!   XK(3) (relative), IGVEC(3,NGG)
!   PSI(NGG) per super band, for PSI0 and (NWAVE=2) PSIM
!   KEY='LAMBDA', NB, NDIM, NSPIN, NLAMBDA
!   LAMBDA(NB,NB) for RLAM0 and (NLAMBDA=2) RLAMM
! Reading follows WAVES_READPSI / WAVES_READLAMBDA (:5147-5513, :4041-4267):
! plane waves are matched by their integer coordinates in the file's
! reciprocal basis (WAVES_MAPG, :4268-4399; absent ones become zero),
! min(NBH,NBH') bands and the first spin are copied, PSIM = PSI0 when
! NWAVE = 1, and the multipliers fill the leading min(NB,NB') block of zeroed
! matrices. Routines return IERR /= 0 and a message instead of stopping.
!
MODULE PAWGPU_RESTART
  IMPLICIT NONE
  TYPE SEPARATOR_TYPE
    SEQUENCE
    INTEGER(4) :: NREC
    CHARACTER(32) :: ID
    CHARACTER(32) :: NAME
    CHARACTER(64) :: VERSION
    CHARACTER(256) :: NOTES
  END TYPE SEPARATOR_TYPE
  CHARACTER(7), PARAMETER :: RESTART_VERSION = 'AUG1996'
  TYPE(SEPARATOR_TYPE), PARAMETER :: RESTART_HEADER = &
 &    SEPARATOR_TYPE(0, 'HEADER', 'NONE', RESTART_VERSION, 'DEFAULT HEADER')
  TYPE(SEPARATOR_TYPE), PARAMETER :: RESTART_EOF = &
 &    SEPARATOR_TYPE(0, 'ENDOFFILE', 'NONE', RESTART_VERSION, 'DEFAULT EOF')
  TYPE(SEPARATOR_TYPE), PARAMETER :: RESTART_WAVES = &
 &    SEPARATOR_TYPE(0, 'WAVES', 'NONE', RESTART_VERSION, 'NONE')
  REAL(8), PARAMETER, PRIVATE :: TWOPI = 6.283185307179586476925286766559D0
CONTAINS
  ! records after the WAVES separator (WAVES_WRITEPSI called with NREC=-1)
  PURE FUNCTION RESTART_WAVES_NREC(NBH, NWAVE, NLAMBDA) RESULT(N)
    INTEGER(4), INTENT(IN) :: NBH, NWAVE, NLAMBDA
    INTEGER(4) :: N
    N = 1 + 2 + NWAVE * NBH + 1 + NLAMBDA
  END FUNCTION RESTART_WAVES_NREC

  ! reciprocal lattice vectors as columns: GBAS^T RBAS = 2 pi
  SUBROUTINE RESTART_GBAS(RBAS, GBAS)
    REAL(8), INTENT(IN) :: RBAS(3, 3)
    REAL(8), INTENT(OUT) :: GBAS(3, 3)
    REAL(8) :: DET
    INTEGER(4) :: I, J, I1, I2, J1, J2
    DO J = 1, 3
      J1 = MOD(J, 3) + 1
      J2 = MOD(J + 1, 3) + 1
      DO I = 1, 3
        I1 = MOD(I, 3) + 1
        I2 = MOD(I + 1, 3) + 1
        ! cofactor of RBAS(I,J)
        GBAS(I, J) = RBAS(I1, J1) * RBAS(I2, J2) - RBAS(I1, J2) * RBAS(I2, J1)
      END DO
    END DO
    DET = SUM(RBAS(:, 1) * GBAS(:, 1))
    GBAS = GBAS * (TWOPI / DET)
  END SUBROUTINE RESTART_GBAS

  SUBROUTINE RESTART_INVERT3(A, AINV)
    REAL(8), INTENT(IN) :: A(3, 3)
    REAL(8), INTENT(OUT) :: AINV(3, 3)
    REAL(8) :: C(3, 3)
    CALL RESTART_GBAS(A, C)
    ! RESTART_GBAS gives 2 pi times the inverse transposed
    AINV = TRANSPOSE(C) / TWOPI
  END SUBROUTINE RESTART_INVERT3

  ! WAVES_MAPG: MAP(IG) = index of the file's vector on the grid point of
  ! GVECB(:,IG), 0 if the file has none
  SUBROUTINE RESTART_MAPG(NGA, GBASA, GVECA, NGB, GVECB, MAP, IERR, MSG)
    INTEGER(4), INTENT(IN) :: NGA, NGB
    REAL(8), INTENT(IN) :: GBASA(3, 3), GVECA(3, NGA), GVECB(3, NGB)
    INTEGER(4), INTENT(OUT) :: MAP(NGB), IERR
    CHARACTER(*), INTENT(OUT) :: MSG
    REAL(8) :: GINV(3, 3), XG(3), XK(3)
    INTEGER(4), ALLOCATABLE :: IV(:, :), GRID(:, :, :)
    INTEGER(4) :: MING(3), MAXG(3), IG, I(3)
    IERR = 0
    MSG = ' '
    MAP = 0
    IF (NGA < 1) RETURN
    CALL RESTART_INVERT3(GBASA, GINV)
    XG = MATMUL(GINV, GVECA(:, 1))
    XK = XG - REAL(NINT(XG), KIND=8)
    ALLOCATE(IV(3, NGA))
    MING = HUGE(1_4)
    MAXG = -HUGE(1_4)
    DO IG = 1, NGA
      XG = MATMUL(GINV, GVECA(:, IG)) - XK
      IV(:, IG) = NINT(XG)
      IF (SUM((XG - REAL(IV(:, IG), KIND=8))**2) > 1.D-6) THEN
        IERR = 1
        MSG = 'G-VECTORS CANNOT BE REPRESENTED AS SUM OF K-VECTOR AND RECIPROCAL LATTICE VECTORS'
        RETURN
      END IF
      MING = MIN(MING, IV(:, IG))
      MAXG = MAX(MAXG, IV(:, IG))
    END DO
    ALLOCATE(GRID(MING(1):MAXG(1), MING(2):MAXG(2), MING(3):MAXG(3)))
    GRID = 0
    DO IG = 1, NGA
      IF (GRID(IV(1, IG), IV(2, IG), IV(3, IG)) /= 0) THEN
        IERR = 1
        MSG = 'TWO G-VECTORS ARE MAPPED ONTO THE SAME POINT'
        RETURN
      END IF
      GRID(IV(1, IG), IV(2, IG), IV(3, IG)) = IG
    END DO
    DO IG = 1, NGB
      I = NINT(MATMUL(GINV, GVECB(:, IG)) - XK)
      IF (ANY(I < MING) .OR. ANY(I > MAXG)) CYCLE
      MAP(IG) = GRID(I(1), I(2), I(3))
    END DO
  END SUBROUTINE RESTART_MAPG

  ! the WAVES section: separator and records
  SUBROUTINE RESTART_WRITE_WAVES(NFIL, RBAS, NB, TSUPER, NGG, XK, IGVEC, NWAVE, PSI0, PSIM, &
 &                               NLAMBDA, RLAM0, RLAMM)
    INTEGER(4), INTENT(IN) :: NFIL, NB, NGG, NWAVE, NLAMBDA
    LOGICAL(4), INTENT(IN) :: TSUPER
    REAL(8), INTENT(IN) :: RBAS(3, 3), XK(3)
    INTEGER(4), INTENT(IN) :: IGVEC(3, NGG)
    COMPLEX(8), INTENT(IN) :: PSI0(NGG, *), PSIM(NGG, *)
    COMPLEX(8), INTENT(IN) :: RLAM0(NB, NB), RLAMM(NB, NB)
    TYPE(SEPARATOR_TYPE) :: SEP
    CHARACTER(8) :: KEY
    INTEGER(4) :: NBH, IB, ILOGICAL
    NBH = NB
    IF (TSUPER) NBH = (NB + 1) / 2
    SEP = RESTART_WAVES
    SEP%NREC = RESTART_WAVES_NREC(NBH, NWAVE, NLAMBDA)
    WRITE(NFIL) SEP
    WRITE(NFIL) 1_4, 1_4, RBAS, NWAVE
    KEY = 'PSI'
    ILOGICAL = 0
    IF (TSUPER) ILOGICAL = 1
    WRITE(NFIL) KEY, NGG, 1_4, NB, ILOGICAL
    WRITE(NFIL) XK, IGVEC
    DO IB = 1, NBH
      WRITE(NFIL) PSI0(:, IB)
    END DO
    IF (NWAVE == 2) THEN
      DO IB = 1, NBH
        WRITE(NFIL) PSIM(:, IB)
      END DO
    END IF
    KEY = 'LAMBDA'
    WRITE(NFIL) KEY, NB, 1_4, 1_4, NLAMBDA
    IF (NLAMBDA >= 1) WRITE(NFIL) RLAM0
    IF (NLAMBDA >= 2) WRITE(NFIL) RLAMM
  END SUBROUTINE RESTART_WRITE_WAVES

  ! the records of a WAVES section (after its separator) into the present set:
  ! NGL plane waves GVEC(3,NGL) of the present cell, NB bands, relative k
  ! point XK; PSI0/PSIM hold the present wave functions on entry (kept for
  ! bands beyond the file's); NLAMBDA returns min(history on file, 2)
  SUBROUTINE RESTART_READ_WAVES(NFIL, NB, TSUPER, NGL, GVEC, XK, PSI0, PSIM, NWAVE, NLAMBDA, &
 &                              RLAM0, RLAMM, IERR, MSG)
    INTEGER(4), INTENT(IN) :: NFIL, NB, NGL
    LOGICAL(4), INTENT(IN) :: TSUPER
    REAL(8), INTENT(IN) :: GVEC(3, NGL), XK(3)
    COMPLEX(8), INTENT(INOUT) :: PSI0(NGL, *), PSIM(NGL, *)
    INTEGER(4), INTENT(OUT) :: NWAVE, NLAMBDA, IERR
    COMPLEX(8), INTENT(OUT) :: RLAM0(NB, NB), RLAMM(NB, NB)
    CHARACTER(*), INTENT(OUT) :: MSG
    INTEGER(4) :: NKPT_, NSPIN_, NGG_, NDIM_, NB_, NBH_, NBH, ILOGICAL, NBL, NDIML, NSPINL, NLAM_
    INTEGER(4) :: IW, IS, IB, IG, I, NBA, IOS
    REAL(8) :: RBAS_(3, 3), GBAS_(3, 3), K_(3)
    CHARACTER(8) :: KEY
    LOGICAL(4) :: TSUPER_
    INTEGER(4), ALLOCATABLE :: IGVEC_(:, :), MAP(:)
    REAL(8), ALLOCATABLE :: GVEC_(:, :)
    COMPLEX(8), ALLOCATABLE :: PSIIN(:), LAMBDA(:, :)
    IERR = 0
    MSG = ' '
    NLAMBDA = 0
    RLAM0 = (0.D0, 0.D0)
    RLAMM = (0.D0, 0.D0)
    READ(NFIL, IOSTAT=IOS) NKPT_, NSPIN_, RBAS_, NWAVE
    IF (IOS /= 0) GOTO 9999
    IF (NKPT_ /= 1) THEN
      IERR = 1
      MSG = 'KPOINT SETS ARE FIXED: NKPT ON FILE DIFFERS FROM 1'
      RETURN
    END IF
    READ(NFIL, IOSTAT=IOS) KEY, NGG_, NDIM_, NB_, ILOGICAL
    IF (IOS /= 0) GOTO 9999
    IF (KEY /= 'PSI') THEN
      IERR = 1
      MSG = 'ID IS NOT "PSI": FILE IS CORRUPTED'
      RETURN
    END IF
    TSUPER_ = BTEST(ILOGICAL, 0)
    ALLOCATE(IGVEC_(3, NGG_), GVEC_(3, NGG_), MAP(NGL), PSIIN(NGG_ * NDIM_))
    READ(NFIL, IOSTAT=IOS) K_, IGVEC_
    IF (IOS /= 0) GOTO 9999
    IF (SUM(ABS(K_ - XK)) > 1.D-8) THEN
      IERR = 1
      MSG = 'INCONSISTENT K-POINTS'
      RETURN
    END IF
    IF (TSUPER_ .NEQV. TSUPER) THEN
      IERR = 1
      MSG = 'TRANSFORMATION BETWEEN REGULAR AND SUPER WAVE FUNCTIONS NOT IMPLEMENTED'
      RETURN
    END IF
    IF (NDIM_ /= 1) THEN
      IERR = 1
      MSG = 'TRANSFORMATION BETWEEN SCALAR AND SPINOR WAVE FUNCTIONS NOT IMPLEMENTED'
      RETURN
    END IF
    CALL RESTART_GBAS(RBAS_, GBAS_)
    DO IG = 1, NGG_
      GVEC_(:, IG) = MATMUL(GBAS_, REAL(IGVEC_(:, IG), KIND=8) + K_)
    END DO
    CALL RESTART_MAPG(NGG_, GBAS_, GVEC_, NGL, GVEC, MAP, IERR, MSG)
    IF (IERR /= 0) RETURN
    NBH_ = NB_
    IF (TSUPER_) NBH_ = (NB_ + 1) / 2
    NBH = NB
    IF (TSUPER) NBH = (NB + 1) / 2
    DO IW = 1, NWAVE
      DO IS = 1, NSPIN_
        DO IB = 1, NBH_
          READ(NFIL, IOSTAT=IOS) PSIIN
          IF (IOS /= 0) GOTO 9999
          IF (IS /= 1 .OR. IB > NBH) CYCLE
          DO IG = 1, NGL
            IF (IW == 1) THEN
              PSI0(IG, IB) = (0.D0, 0.D0)
              IF (MAP(IG) /= 0) PSI0(IG, IB) = PSIIN(MAP(IG))
            ELSE
              PSIM(IG, IB) = (0.D0, 0.D0)
              IF (MAP(IG) /= 0) PSIM(IG, IB) = PSIIN(MAP(IG))
            END IF
          END DO
        END DO
      END DO
    END DO
    IF (NWAVE == 1) PSIM(:, 1:MIN(NBH, NBH_)) = PSI0(:, 1:MIN(NBH, NBH_))
    READ(NFIL, IOSTAT=IOS) KEY, NBL, NDIML, NSPINL, NLAM_
    IF (IOS /= 0) GOTO 9999
    IF (KEY /= 'LAMBDA') THEN
      IERR = 1
      MSG = 'ID IS NOT "LAMBDA": FILE IS CORRUPTED'
      RETURN
    END IF
    ALLOCATE(LAMBDA(NBL, NBL))
    NBA = MIN(NB, NBL)
    DO I = 1, NLAM_
      DO IS = 1, NSPINL
        READ(NFIL, IOSTAT=IOS) LAMBDA
        IF (IOS /= 0) GOTO 9999
        IF (IS /= 1) CYCLE
        IF (I == 1) RLAM0(1:NBA, 1:NBA) = LAMBDA(1:NBA, 1:NBA)
        IF (I == 2) RLAMM(1:NBA, 1:NBA) = LAMBDA(1:NBA, 1:NBA)
      END DO
    END DO
    NLAMBDA = MIN(NLAM_, 2)
    RETURN
9999 CONTINUE
    IERR = 2
    WRITE(MSG, '(A,I0)') 'ERROR WHILE READING FROM FILE: IOSTAT=', IOS
  END SUBROUTINE RESTART_READ_WAVES

  ! RESTART$SKIP: read a separator and pass over its records
  SUBROUTINE RESTART_SKIP(NFIL, ID, IERR)
    INTEGER(4), INTENT(IN) :: NFIL
    CHARACTER(*), INTENT(OUT) :: ID
    INTEGER(4), INTENT(OUT) :: IERR
    TYPE(SEPARATOR_TYPE) :: SEP
    INTEGER(4) :: I, IOS
    IERR = 0
    READ(NFIL, IOSTAT=IOS) SEP
    IF (IOS /= 0) THEN
      IERR = 2
      RETURN
    END IF
    ID = SEP%ID
    IF (SEP%NREC < 0 .OR. SEP%NREC > 10000000) THEN
      IERR = 1
      RETURN
    END IF
    DO I = 1, SEP%NREC
      READ(NFIL, IOSTAT=IOS)
      IF (IOS /= 0) THEN
        IERR = 2
        RETURN
      END IF
    END DO
  END SUBROUTINE RESTART_SKIP
END MODULE PAWGPU_RESTART
