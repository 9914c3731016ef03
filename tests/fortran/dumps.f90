! dumps.f90 -- the reference's dump and initial_sol statements as a test
! program (test infrastructure for tests/test_textio_host.py): reads a grid,
! PHI (Fortran layout, COMPLEX(8)) and X, Y, Z from the stream file given as
! the first argument, writes every dump kind with the FORMATs and loops of
! SRC/cgpe1d.f90:513-516,547-549, SRC/cgpe2d.f90:521-528,560-565 and
! SRC/cgpe3d.f90:534-563,596-602, then reads the file named by the second
! argument list-directed as initial_sol (SRC/cgpe3d.f90:391-409 and its 1D/2D
! forms) and writes the PHI, X, Y it reads to the third argument.
PROGRAM DUMPS
  IMPLICIT NONE
  INTEGER :: DIM, MS, NX, NY, NZ, I, J, K, L, NX2, NY2, NZ2
  COMPLEX(8), ALLOCATABLE :: PHI(:,:,:,:)
  REAL(8), ALLOCATABLE :: X(:), Y(:), Z(:)
  REAL(8) :: TMP(12)
  CHARACTER(LEN=512) :: FIN, FSOL, FOUT
  CALL GET_COMMAND_ARGUMENT(1, FIN)
  CALL GET_COMMAND_ARGUMENT(2, FSOL)
  CALL GET_COMMAND_ARGUMENT(3, FOUT)
  OPEN (20, FILE=TRIM(FIN), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
  READ (20) DIM, MS, NX, NY, NZ
  ALLOCATE (PHI(NX, NY, NZ, MS), X(NX), Y(NY), Z(NZ))
  READ (20) PHI, X, Y, Z
  CLOSE (20)
  NX2 = NX/2 + 1
  NY2 = NY/2 + 1
  NZ2 = NZ/2 + 1
  SELECT CASE (DIM)
    CASE (1)
      OPEN (7, FILE='f_1d_tmp.dat', STATUS='replace', FORM='FORMATTED', ACTION='WRITE')
      DO I = 1, NX
        WRITE (7,'(F11.5,10F14.6)') X(I), (ABS(PHI(I,1,1,L))**2, ATAN2(AIMAG(PHI(I,1,1,L)), REAL(PHI(I,1,1,L))), L = 1, MS)
      END DO
      CLOSE (7)
      OPEN (8, FILE='f_1d_sol.dat', STATUS='replace', FORM='FORMATTED', ACTION='WRITE')
      DO I = 1, NX
        WRITE (8,'(F11.5,10F18.10)') X(I), (ABS(PHI(I,1,1,L))**2, ATAN2(AIMAG(PHI(I,1,1,L)), REAL(PHI(I,1,1,L))), L = 1, MS)
      END DO
      CLOSE (8)
    CASE (2)
      OPEN (7, FILE='f_2d.dat', STATUS='replace', FORM='FORMATTED', ACTION='WRITE')
      DO I = 1, NX
        DO J = 1, NY
          WRITE (7,'(2F8.2,10F10.5)') X(I), Y(J), (ABS(PHI(I,J,1,L))**2, &
                 ATAN2(AIMAG(PHI(I,J,1,L)), REAL(PHI(I,J,1,L))), L = 1, MS)
        END DO
        WRITE (7,'(11F10.2)')
      END DO
      CLOSE (7)
    CASE (3)
      OPEN (7, FILE='f_3d.dat', STATUS='replace', FORM='FORMATTED', ACTION='WRITE')
      DO I = 1, NX
        DO J = 1, NY
          DO K = 1, NZ
            WRITE (7,'(10F16.10)') (ABS(PHI(I,J,K,L))**2, ATAN2(AIMAG(PHI(I,J,K,L)), REAL(PHI(I,J,K,L))), L = 1, MS)
          END DO
        END DO
      END DO
      CLOSE (7)
      OPEN (11, FILE='f_3d_xy.dat', STATUS='replace', FORM='FORMATTED', ACTION='WRITE')
      DO I = 1, NX
        DO J = 1, NY
          WRITE (11,'(2F10.4,10F14.8)') X(I), Y(J), (ABS(PHI(I,J,NZ2,L))**2, &
                 ATAN2(AIMAG(PHI(I,J,NZ2,L)), REAL(PHI(I,J,NZ2,L))), L = 1, MS)
        END DO
        WRITE (11,'(2F10.4,10F14.8)')
      END DO
      CLOSE (11)
      OPEN (12, FILE='f_3d_xz.dat', STATUS='replace', FORM='FORMATTED', ACTION='WRITE')
      DO I = 1, NX
        DO J = 1, NZ
          WRITE (12,'(2F10.4,10F14.8)') X(I), Z(J), (ABS(PHI(I,NY2,J,L))**2, &
                 ATAN2(AIMAG(PHI(I,NY2,J,L)), REAL(PHI(I,NY2,J,L))), L = 1, MS)
        END DO
        WRITE (12,'(2F10.4,10F14.8)')
      END DO
      CLOSE (12)
  END SELECT
  ! initial_sol, list-directed
  OPEN (1, FILE=TRIM(FSOL), STATUS='OLD', FORM='FORMATTED', ACTION='READ')
  DO I = 1, NX
    DO J = 1, NY
      DO K = 1, NZ
        SELECT CASE (DIM)
          CASE (1)
            READ (1,*) X(I), (TMP(L), L = 1, 2*MS)
          CASE (2)
            READ (1,*) X(I), Y(J), (TMP(L), L = 1, 2*MS)
          CASE (3)
            READ (1,*) (TMP(L), L = 1, 2*MS)
        END SELECT
        DO L = 1, MS
          PHI(I,J,K,L) = SQRT(TMP(2*L-1))*EXP(CMPLX(0.0D0, TMP(2*L), KIND=8))
        END DO
      END DO
    END DO
  END DO
  CLOSE (1)
  OPEN (21, FILE=TRIM(FOUT), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='replace', ACTION='WRITE')
  WRITE (21) PHI, X, Y
  CLOSE (21)
END PROGRAM DUMPS
