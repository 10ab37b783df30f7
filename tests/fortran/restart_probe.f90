!
! Test program for the restart records (module PAWGPU_RESTART), compiled and
! run by tests/test_restart.py on the CPU:
!   restart_probe write FILE   write HEADER, an ATOMS section of two records,
!                              the WAVES section of the fixed data below, and
!                              ENDOFFILE
!   restart_probe read FILE    read the WAVES section of FILE into the present
!                              set below (skipping other sections) and print
!                              NWAVE, NLAMBDA, PSI0, PSIM, RLAM0, RLAMM
! Fixed data (the test restates it): RBAS columns (6,0,0) (1,7,0) (0.5,1,8);
! file G set IGVEC(:,IG) from the triple loop below; NB=5 super bands (NBH=3)
!   PSI0(IG,IB) = (IG + 0.25 IB, -0.5 IG IB), PSIM(IG,IB) = (2 IG, IB)
!   RLAM0(I,J) = (I + 10 J, 0.5 I), RLAMM(I,J) = (-I, J)
! Present set: NB=4 (NBH=2), G vectors = the file's in reverse order followed by
! (9,9,9) (absent from the file); present wave functions (-1,-1) on entry.
!
PROGRAM RESTART_PROBE
  USE PAWGPU_RESTART
  IMPLICIT NONE
  CHARACTER(512) :: MODE, FNAME
  INTEGER(4), PARAMETER :: NB = 5, NBH = 3, NGG = 27, NBP = 4, NBHP = 2, NGP = NGG + 1
  REAL(8) :: RBAS(3, 3), GBAS(3, 3), XK(3)
  INTEGER(4) :: IGVEC(3, NGG), IG, IB, I, J, K, IU, IO, NWAVE, NLAMBDA, IERR
  COMPLEX(8) :: PSI0(NGG, NBH), PSIM(NGG, NBH), RLAM0(NB, NB), RLAMM(NB, NB)
  COMPLEX(8) :: P0(NGP, NBHP), PM(NGP, NBHP), L0(NBP, NBP), LM(NBP, NBP)
  REAL(8) :: GVEC(3, NGP)
  TYPE(SEPARATOR_TYPE) :: SEP
  CHARACTER(256) :: MSG
  CHARACTER(32) :: SKIPID
  LOGICAL(4) :: TDONE

  CALL GET_COMMAND_ARGUMENT(1, MODE)
  CALL GET_COMMAND_ARGUMENT(2, FNAME)
  RBAS = RESHAPE((/6.D0, 0.D0, 0.D0, 1.D0, 7.D0, 0.D0, 0.5D0, 1.D0, 8.D0/), (/3, 3/))
  IG = 0
  DO I = -1, 1
    DO J = -1, 1
      DO K = -1, 1
        IG = IG + 1
        IGVEC(:, IG) = (/K, J - K, I + 2 * J/)
      END DO
    END DO
  END DO
  XK = 0.D0

  IF (TRIM(MODE) == 'write') THEN
    DO IB = 1, NBH
      DO IG = 1, NGG
        PSI0(IG, IB) = CMPLX(IG + 0.25D0 * IB, -0.5D0 * IG * IB, KIND=8)
        PSIM(IG, IB) = CMPLX(2.D0 * IG, REAL(IB, 8), KIND=8)
      END DO
    END DO
    DO J = 1, NB
      DO I = 1, NB
        RLAM0(I, J) = CMPLX(I + 10.D0 * J, 0.5D0 * I, KIND=8)
        RLAMM(I, J) = CMPLX(-REAL(I, 8), REAL(J, 8), KIND=8)
      END DO
    END DO
    OPEN(NEWUNIT=IU, FILE=TRIM(FNAME), FORM='UNFORMATTED', ACCESS='SEQUENTIAL', STATUS='REPLACE')
    WRITE(IU) RESTART_HEADER
    SEP = SEPARATOR_TYPE(2, 'ATOMS', 'NONE', RESTART_VERSION, 'NONE')
    WRITE(IU) SEP
    WRITE(IU) 7_4, 1.5D0
    WRITE(IU) (/1.D0, 2.D0, 3.D0/)
    CALL RESTART_WRITE_WAVES(IU, RBAS, NB, .TRUE., NGG, XK, IGVEC, 2, PSI0, PSIM, 2, RLAM0, RLAMM)
    WRITE(IU) RESTART_EOF
    CLOSE(IU)
  ELSE IF (TRIM(MODE) == 'read') THEN
    CALL RESTART_GBAS(RBAS, GBAS)
    DO IG = 1, NGG
      GVEC(:, IG) = MATMUL(GBAS, REAL(IGVEC(:, NGG + 1 - IG), KIND=8))
    END DO
    GVEC(:, NGP) = MATMUL(GBAS, (/9.D0, 9.D0, 9.D0/))
    P0 = (-1.D0, -1.D0)
    PM = (-1.D0, -1.D0)
    OPEN(NEWUNIT=IU, FILE=TRIM(FNAME), FORM='UNFORMATTED', ACCESS='SEQUENTIAL', STATUS='OLD')
    READ(IU) SEP
    IF (SEP%ID /= RESTART_HEADER%ID) STOP 'FORMAT OF RESTART FILE NOT RECOGNIZED'
    TDONE = .FALSE.
    DO
      READ(IU, IOSTAT=IO) SEP
      IF (IO /= 0) STOP 'no ENDOFFILE'
      IF (SEP%ID == RESTART_EOF%ID) EXIT
      IF (SEP%ID == RESTART_WAVES%ID .AND. .NOT. TDONE) THEN
        CALL RESTART_READ_WAVES(IU, NBP, .TRUE., NGP, GVEC, XK, P0, PM, NWAVE, NLAMBDA, L0, LM, IERR, MSG)
        IF (IERR /= 0) THEN
          WRITE(0, '(A)') TRIM(MSG)
          STOP 1
        END IF
        TDONE = .TRUE.
      ELSE
        BACKSPACE(IU)
        CALL RESTART_SKIP(IU, SKIPID, IERR)
        IF (IERR /= 0) STOP 'FILE STRUCTURE CORRUPTED'
        WRITE(*, '(A,A)') 'skipped ', TRIM(SKIPID)
      END IF
    END DO
    CLOSE(IU)
    WRITE(*, '(A,2I4)') 'counts ', NWAVE, NLAMBDA
    DO IB = 1, NBHP
      DO IG = 1, NGP
        WRITE(*, '(A,2I4,4ES26.17E3)') 'psi ', IB, IG, P0(IG, IB), PM(IG, IB)
      END DO
    END DO
    DO J = 1, NBP
      DO I = 1, NBP
        WRITE(*, '(A,2I4,4ES26.17E3)') 'lam ', I, J, L0(I, J), LM(I, J)
      END DO
    END DO
  ELSE
    STOP 'usage: restart_probe write|read FILE'
  END IF
END PROGRAM RESTART_PROBE
