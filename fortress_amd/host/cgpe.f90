! cgpe.f90 -- drop-in host for FORTRESS's cgpe.x on MI355X.
!
! Reads ./INPUT/inputND.dat and writes ./OUTPUT/* exactly as the reference
! drivers do (SRC/cgpe1d.f90:105-563, SRC/cgpe2d.f90:108-579,
! SRC/cgpe3d.f90:113-617): same records, same file names and FORMATs, the
! same STP/NSTP events, CONV/CONV_EN test and EXIT semantics.  The time loop
! runs on the GPU through libfgpe; the host only crosses the boundary at
! events (observables every STP, dumps every NSTP) and at the end.
!
! Usage:  cgpe.x [1D|2D|3D] [fp64|fp32]     (default 3D fp64, like the
! reference Makefile's default DIMENSION=3D; fp32 selects FP32 propagation with
! FP64 reductions).  The environment variable FGPE_PRECISION=fp32 does the same.
!
! Binary checkpoints (opt-in, no counterpart in the reference; without these
! variables every file is as the reference writes it):
!   FGPE_CHECKPOINT=path, FGPE_CHECKPOINT_EVERY=k   write a checkpoint of the
!     run (state, Newton state, II, EN_OLD) to path every k steps;
!   FGPE_RESUME=path   start from such a checkpoint of a run with the same
!     input (imaginary or real time; initial_sol is then not read) at its II:
!     the files then hold the rows from there on, and the run continues
!     bit-identically to one that never stopped (given the same event steps).
PROGRAM CGPE
  USE, INTRINSIC :: ISO_C_BINDING
  USE FGPE
  IMPLICIT NONE

  REAL(8), PARAMETER :: EPS = 1.0D-40
  TYPE(FGPE_INPUT) :: INP
  TYPE(FGPE_PARAMS) :: PRM
  TYPE(C_PTR) :: H
  INTEGER :: DIM, SPIN, MS, NX, NY, NZ, NITER, STP, NSTP, THREADS, OPTION_FPC, SWITCH_IM, SWITCH_SOC
  INTEGER :: NATOMS, I, J, K, L, II, IIPRINT, IOSTAT, NEXTEV, RC, NX2, NY2, NZ2
  REAL(8) :: DX, DY, DZ, DT, M, A0, A2, A4, NUX, NUY, NUZ, TOL, GAMMAX, GAMMAY, GAMMAZ, MAG
  REAL(8) :: AOSC, TNORM, EN, EN_OLD, MZ, CONV, CONV_EN, DMAX, START, FINISH
  REAL(8) :: NORM(5), RMS(5), MU(5), CEN(5)
  REAL(8), ALLOCATABLE :: X(:), Y(:), Z(:)
  COMPLEX(8), ALLOCATABLE :: PHI(:,:,:,:)
  CHARACTER(LEN=1) :: SPIN_ID
  CHARACTER(LEN=8) :: ARG
  CHARACTER(LEN=40) :: INFILE
  CHARACTER(LEN=2) :: TAG
  CHARACTER(LEN=10) :: DATE, TIME, ZONE
  INTEGER :: VALUES(8), COUNT0, COUNT1, RATE
  INTEGER(C_INT32_T) :: PREC
  LOGICAL :: FEXIST, EVSTP, EVNSTP, RESUMED
  CHARACTER(LEN=1024) :: CKPATH, RSPATH
  CHARACTER(LEN=32) :: EVERYSTR
  INTEGER :: CKEVERY
  TYPE(FGPE_LOOP_STATE) :: LS

  CALL SYSTEM_CLOCK(COUNT0, RATE)
  CALL DATE_AND_TIME(DATE, TIME, ZONE, VALUES)

  DIM = 3
  PREC = FGPE_FP64
  CALL GET_ENVIRONMENT_VARIABLE('FGPE_PRECISION', ARG)
  IF (TRIM(ARG) == 'fp32' .OR. TRIM(ARG) == 'FP32') PREC = FGPE_FP32
  CKPATH = ' '
  RSPATH = ' '
  EVERYSTR = ' '
  CKEVERY = 0
  CALL GET_ENVIRONMENT_VARIABLE('FGPE_CHECKPOINT', CKPATH)
  CALL GET_ENVIRONMENT_VARIABLE('FGPE_CHECKPOINT_EVERY', EVERYSTR)
  CALL GET_ENVIRONMENT_VARIABLE('FGPE_RESUME', RSPATH)
  IF (LEN_TRIM(EVERYSTR) > 0) READ (EVERYSTR,*) CKEVERY
  IF (LEN_TRIM(CKPATH) == 0) CKEVERY = 0
  RESUMED = LEN_TRIM(RSPATH) > 0
  DO I = 1, COMMAND_ARGUMENT_COUNT()
    CALL GET_COMMAND_ARGUMENT(I, ARG)
    SELECT CASE (TRIM(ARG))
      CASE ('1D', '1d', '1')
        DIM = 1
      CASE ('2D', '2d', '2')
        DIM = 2
      CASE ('3D', '3d', '3')
        DIM = 3
      CASE ('fp32', 'FP32')
        PREC = FGPE_FP32
      CASE ('fp64', 'FP64')
        PREC = FGPE_FP64
    END SELECT
  END DO

  OPEN (100, FILE='err_msg.dat', STATUS='UNKNOWN', FORM='FORMATTED', ACTION='WRITE')
  WRITE (INFILE, '(A,I1,A)') './INPUT/input', DIM, 'D.dat'
  INQUIRE (FILE=TRIM(INFILE), EXIST=FEXIST)
  IF (FEXIST) THEN
    OPEN (10, FILE=TRIM(INFILE), STATUS='OLD', FORM='FORMATTED', ACTION='READ')
  ELSE
    ! same text as SRC/cgpe3d.f90:228-229 (cgpe1d/cgpe2d name their own file),
    ! one record whatever the line length of list-directed output
    WRITE (100,'(1X,A)') "'input"//ACHAR(48 + DIM)//"D.dat' does not exist."
    WRITE (*,'(1X,A)') "'input"//ACHAR(48 + DIM)//"D.dat' does not exist."
    STOP
  END IF

  ! ---- records of INPUT/inputND.dat (SRC/cgpe3d.f90:239-266)
  SWITCH_SOC = 0
  IOSTAT = 0
  OPTION_FPC = 4
  MAG = 0.0D0
  DY = 1.0D0
  DZ = 1.0D0
  NY = 1
  NZ = 1
  GAMMAY = 0.0D0
  GAMMAZ = 0.0D0
  A4 = 0.0D0
  READ (10,*) SPIN
  READ (10,*) NITER
  READ (10,*) STP, NSTP
  READ (10,*) THREADS
  SELECT CASE (DIM)
    CASE (1)
      READ (10,*) NX
      READ (10,*) DX, DT
    CASE (2)
      READ (10,*) NX, NY
      READ (10,*) DX, DY, DT
    CASE (3)
      READ (10,*) NX, NY, NZ
      READ (10,*) DX, DY, DZ, DT
  END SELECT
  READ (10,*) M
  IF (SPIN == 1) THEN
    READ (10,*) A0, A2
  ELSE
    READ (10,*) A0, A2, A4
  END IF
  READ (10,*) NATOMS
  READ (10,*) NUX, NUY, NUZ
  READ (10,*) TOL
  READ (10,*) SWITCH_IM
  SELECT CASE (DIM)
    CASE (1)
      READ (10,*) GAMMAX
    CASE (2)
      READ (10,*) GAMMAX, GAMMAY
    CASE (3)
      READ (10,*) GAMMAX, GAMMAY, GAMMAZ
  END SELECT
  IF (ABS(GAMMAX) > EPS .OR. ABS(GAMMAY) > EPS .OR. ABS(GAMMAZ) > EPS) SWITCH_SOC = 1
  IF (SWITCH_SOC == 0) THEN
    READ (10,*) OPTION_FPC
  ELSE
    READ (10,*)
  END IF
  IF (SWITCH_IM == 1 .AND. SWITCH_SOC == 0) THEN
    READ (10,*) MAG
  ELSE
    READ (10,*)
  END IF
  CLOSE (10)

  MS = 2*SPIN + 1
  NX2 = NX/2 + 1
  NY2 = NY/2 + 1
  NZ2 = NZ/2 + 1
  INP%DIM = DIM
  INP%SPIN = SPIN
  INP%NITER = NITER
  INP%STP = STP
  INP%NSTP = NSTP
  INP%THREADS = THREADS
  INP%N = [NX, NY, NZ]
  INP%D = [DX, DY, DZ]
  INP%DT = DT
  INP%MASS = M
  INP%A = [A0, A2, A4]
  INP%NATOMS = NATOMS
  INP%NU = [NUX, NUY, NUZ]
  INP%TOL = TOL
  INP%SWITCH_IM = SWITCH_IM
  INP%GAMMA = [GAMMAX, GAMMAY, GAMMAZ]
  INP%OPTION_FPC = OPTION_FPC
  INP%MAG = MAG
  RC = FGPE_DERIVE(INP, PREC, PRM, AOSC)
  IF (RC /= 0) CALL FAIL_STOP(RC)
  WRITE (*,*) 'MAXIMUM_THREAD_NUMBER_SET = ', THREADS

  IF (SPIN == 1) THEN
    SPIN_ID = '1'
  ELSE
    SPIN_ID = '2'
  END IF
  IF (SWITCH_IM == 1) THEN
    TAG = 'im'
    OPEN (9, FILE='./OUTPUT/convergence.dat', STATUS='UNKNOWN', FORM='FORMATTED', ACTION='WRITE')
  ELSE
    TAG = 're'
    OPEN (1, FILE='./INPUT/initial_sol_spin'//SPIN_ID//'.dat', IOSTAT=IOSTAT, STATUS='OLD', FORM='FORMATTED', &
          ACTION='READ')
  END IF
  OPEN (2, FILE='./OUTPUT/file1_'//TAG//'_spin'//SPIN_ID//'.dat', STATUS='UNKNOWN', FORM='FORMATTED', ACTION='WRITE')
  OPEN (3, FILE='./OUTPUT/file2_'//TAG//'_spin'//SPIN_ID//'.dat', STATUS='UNKNOWN', FORM='FORMATTED', ACTION='WRITE')
  OPEN (4, FILE='./OUTPUT/file3_'//TAG//'_spin'//SPIN_ID//'.dat', STATUS='UNKNOWN', FORM='FORMATTED', ACTION='WRITE')
  OPEN (8, FILE='./OUTPUT/solution_file_'//TAG//'_spin'//SPIN_ID//'.dat', STATUS='UNKNOWN', FORM='FORMATTED', &
        ACTION='WRITE')

  WRITE (2,1007) VALUES(3), VALUES(2), VALUES(1), VALUES(5), VALUES(6), VALUES(7)
  CALL HEADER()

  ! ---- grids for the dumps (SRC/cgpe3d.f90:729-744)
  ALLOCATE (X(NX), Y(NY), Z(NZ))
  DO I = 1, NX
    X(I) = (-1.0D0 + 2.0D0*(I - 1.0D0)/NX)*(NX*DX)/2.0D0
  END DO
  DO I = 1, NY
    Y(I) = (-1.0D0 + 2.0D0*(I - 1.0D0)/NY)*(NY*DY)/2.0D0
  END DO
  DO I = 1, NZ
    Z(I) = (-1.0D0 + 2.0D0*(I - 1.0D0)/NZ)*(NZ*DZ)/2.0D0
  END DO
  IF (DIM < 2) Y = 0.0D0
  IF (DIM < 3) Z = 0.0D0
  ALLOCATE (PHI(NX, NY, NZ, MS))

  ! ---- real time needs initial_sol (SRC/cgpe3d.f90:383-390): checked after the
  ! file1 header, as there, and before the device is touched
  IF (SWITCH_IM /= 1 .AND. IOSTAT /= 0 .AND. .NOT. RESUMED) THEN
    WRITE (*,'(1X,A)') 'Check if the "initial_sol_spin'//SPIN_ID//'.dat" is available'
    WRITE (*,'(1X,A)') 'if not, user can read the "solution_file_im_spin'//SPIN_ID//'.dat" file as an initial solution.'
    WRITE (100,'(1X,A)') 'Check if the "initial_sol_spin'//SPIN_ID//'.dat" is available'
    WRITE (100,'(1X,A)') 'if not, user can read the "solution_file_im_spin'//SPIN_ID//'.dat" file as an initial solution.'
    STOP
  END IF

  RC = FGPE_CREATE(PRM, H)
  IF (RC /= 0) CALL FAIL_STOP(RC)
  RC = FGPE_INITIAL_GUESS(H)
  IF (RC /= 0) CALL FAIL_STOP(RC)

  ! ---- real time: initial state from initial_sol (SRC/cgpe3d.f90:383-410)
  IF (SWITCH_IM /= 1 .AND. .NOT. RESUMED) THEN
    ! SRC/cgpe3d.f90:391-409 (2D SRC/cgpe2d.f90:382-397, 1D SRC/cgpe1d.f90:
    ! 377-388), parsed on the input's thread count: X (and Y) come from the file
    CLOSE (1)
    RC = FGPE_READ_INITIAL_SOL('./INPUT/initial_sol_spin'//SPIN_ID//'.dat'//C_NULL_CHAR, INT(DIM, C_INT32_T), &
                               INP%N, INT(MS, C_INT32_T), PHI, X, Y, INT(THREADS, C_INT32_T))
    IF (RC /= 0) CALL FAIL_STOP(RC)
    RC = FGPE_UPLOAD(H, PHI)
    IF (RC /= 0) CALL FAIL_STOP(RC)
  END IF

  IF (RESUMED) THEN
    ! ---- resume: state, Newton state, II and EN_OLD from the checkpoint
    RC = FGPE_CHECKPOINT_LOAD(H, TRIM(RSPATH)//C_NULL_CHAR, LS)
    IF (RC /= 0) CALL FAIL_STOP(RC)
    II = INT(LS%II)
    EN_OLD = LS%EN_OLD
    RC = FGPE_LAST_NORMS(H, NORM)
    IF (RC /= 0) CALL FAIL_STOP(RC)
  ELSE
    IF (SWITCH_IM == 1) THEN
      RC = FGPE_NORMALIZE(H, TNORM)
      IF (RC /= 0) CALL FAIL_STOP(RC)
    END IF

    ! ---- initial observables (SRC/cgpe3d.f90:416-431)
    RC = FGPE_NORMC(H, NORM)
    IF (RC /= 0) CALL FAIL_STOP(RC)
    RC = FGPE_OBSERVABLES(H, EN, MU, MZ, RMS)
    IF (RC /= 0) CALL FAIL_STOP(RC)
    RC = FGPE_CENTER_ABS(H, CEN)
    IF (RC /= 0) CALL FAIL_STOP(RC)
    CALL ROW_INITIAL()
    EN_OLD = 0.0D0
    II = 0
    WRITE (3,'(F14.6,6F18.10)') 2.0D0*II*DT, EN/2.0D0, (RMS(L), L = 1, MS)
    WRITE (4,'(F14.6,7F18.10)') 2.0D0*II*DT, (NORM(L), L = 1, MS), SUM(NORM(1:MS)), MZ
    ! EN_OLD starts at 0 (SRC/cgpe3d.f90:307) and becomes EN at the end of every
    ! step (529), so the first CONV_EN test sees the initial EN unless STP = 1
    IF (STP > 1) EN_OLD = EN
  END IF

  ! ---- time loop, advanced between events (SRC/cgpe3d.f90:451-567)
  IIPRINT = NITER
  DO WHILE (II < NITER)
    NEXTEV = MIN(NITER, (II/STP + 1)*STP, (II/NSTP + 1)*NSTP)
    IF (CKEVERY > 0) NEXTEV = MIN(NEXTEV, (II/CKEVERY + 1)*CKEVERY)
    EVSTP = MOD(NEXTEV, STP) == 0
    EVNSTP = MOD(NEXTEV, NSTP) == 0
    IF (SWITCH_IM == 1 .AND. EVSTP) THEN
      ! PHI_OLD is the state after step NEXTEV-1 (SRC/cgpe3d.f90:530);
      ! before the first step it is zero (306)
      IF (NEXTEV - 1 > II) THEN
        RC = FGPE_RUN(H, INT(NEXTEV - 1 - II, C_INT64_T))
        IF (RC /= 0) CALL FAIL_STOP(RC)
        II = NEXTEV - 1
      END IF
      IF (II > 0) THEN
        RC = FGPE_KEEP_OLD(H)
        IF (RC /= 0) CALL FAIL_STOP(RC)
      END IF
    END IF
    RC = FGPE_RUN(H, INT(NEXTEV - II, C_INT64_T))
    IF (RC /= 0) CALL FAIL_STOP(RC)
    II = NEXTEV
    RC = FGPE_LAST_NORMS(H, NORM)
    IF (RC /= 0) CALL FAIL_STOP(RC)
    IF (EVSTP) THEN
      RC = FGPE_OBSERVABLES(H, EN, MU, MZ, RMS)
      IF (RC /= 0) CALL FAIL_STOP(RC)
      IF (SWITCH_IM == 1) THEN
        RC = FGPE_CONV_MAX(H, DMAX)
        IF (RC /= 0) CALL FAIL_STOP(RC)
        CONV = DMAX/(STP*2.0D0*DT)
        CONV_EN = ABS(EN - EN_OLD)/(STP*2.0D0*DT)
        WRITE (9,*) 2.0D0*II*DT, CONV, CONV_EN
        IF (CONV <= TOL .OR. CONV_EN <= TOL) THEN
          IIPRINT = II - 1
          EXIT
        END IF
      END IF
      WRITE (3,'(F14.6,6F18.10)') 2.0D0*II*DT, EN/2.0D0, (RMS(L), L = 1, MS)
      WRITE (4,'(F14.6,7F18.10)') 2.0D0*II*DT, (NORM(L), L = 1, MS), SUM(NORM(1:MS)), MZ
      EN_OLD = EN
    END IF
    IF (EVNSTP) THEN
      RC = FGPE_CENTER_ABS(H, CEN)
      IF (RC /= 0) CALL FAIL_STOP(RC)
      CALL ROW_NSTP()
      RC = FGPE_DOWNLOAD(H, PHI)
      IF (RC /= 0) CALL FAIL_STOP(RC)
      CALL DUMP_TMP()
    END IF
    IF (CKEVERY > 0) THEN
      IF (MOD(II, CKEVERY) == 0) THEN
        LS%II = II
        LS%EN_OLD = EN_OLD
        LS%USER = 0.0D0
        RC = FGPE_CHECKPOINT_SAVE(H, TRIM(CKPATH)//C_NULL_CHAR, LS)
        IF (RC /= 0) CALL FAIL_STOP(RC)
      END IF
    END IF
  END DO

  ! ---- final observables and solution (SRC/cgpe3d.f90:569-602)
  RC = FGPE_OBSERVABLES(H, EN, MU, MZ, RMS)
  IF (RC /= 0) CALL FAIL_STOP(RC)
  RC = FGPE_CENTER_ABS(H, CEN)
  IF (RC /= 0) CALL FAIL_STOP(RC)
  CALL ROW_FINAL(IIPRINT)
  RC = FGPE_DOWNLOAD(H, PHI)
  IF (RC /= 0) CALL FAIL_STOP(RC)
  ! unit 8 (opened at the start, as in the reference) is closed before the
  ! file is written through the library
  CLOSE (8)
  CALL DUMP_SOLUTION()
  CALL FGPE_DESTROY(H)
  PRINT *, 'Program execution complete'
  CALL DATE_AND_TIME(DATE, TIME, ZONE, VALUES)
  CALL SYSTEM_CLOCK(COUNT1)
  START = 0.0D0
  FINISH = DBLE(COUNT1 - COUNT0)/DBLE(RATE)
  WRITE (2,1008) VALUES(3), VALUES(2), VALUES(1), VALUES(5), VALUES(6), VALUES(7)
  IF (DIM == 1) THEN
    WRITE (2,1109) (FINISH - START)
  ELSE
    WRITE (2,1009) (FINISH - START)
  END IF
  CLOSE (100)

1007 FORMAT("Started on DAY/MONTH/YEAR, HOURS:MINUTES:SECONDS = ",I2,'/',I2,'/',I4,',',I3,':',I2,':',I2)
1008 FORMAT("Ended on DAY/MONTH/YEAR, HOURS:MINUTES:SECONDS = ",I2,'/',I2,'/',I4,',',I3,':',I2,':',I2)
1009 FORMAT("Time elapsed = ",F18.8, ' in seconds')
1109 FORMAT("Time elapsed = ",F18.6, ' in seconds')

CONTAINS

  SUBROUTINE FAIL_STOP(CODE)
    INTEGER(C_INT), INTENT(IN) :: CODE
    IF (CODE == FGPE_ENEWTON) THEN
      ! SRC/cgpe3d.f90:1556-1557, as one record
      WRITE (100,'(1X,A,1X,ES22.15E3,A)') 'Newton-Raphason subroutine could not calculate the roots within', &
            1.0D-6, ' tolerance'
    ELSE
      WRITE (100,'(A,I0,2A)') ' libfgpe error ', CODE, ': ', TRIM(FGPE_MESSAGE())
      WRITE (*,'(A,I0,2A)') ' libfgpe error ', CODE, ': ', TRIM(FGPE_MESSAGE())
    END IF
    STOP
  END SUBROUTINE FAIL_STOP

  ! file1 header: 1D SRC/cgpe1d.f90:334-365, 2D SRC/cgpe2d.f90:338-369,
  ! 3D SRC/cgpe3d.f90:346-379
  SUBROUTINE HEADER()
    WRITE (2,'(A)') ' '
    WRITE (2,899) THREADS, THREADS
    SELECT CASE (DIM)
      CASE (1)
        WRITE (2,1900) SWITCH_IM, OPTION_FPC, SWITCH_SOC, GAMMAX, TOL
        WRITE (2,1901) PRM%ALPHA(1), PRM%ALPHA(2), PRM%ALPHA(3)
        WRITE (2,'(A)') ' '
        WRITE (2,1902) NX
        WRITE (2,1903) NITER, NSTP, STP
        WRITE (2,1905) DX
        WRITE (2,1906) DT, MAG
      CASE (2)
        WRITE (2,2900) SWITCH_IM, OPTION_FPC, SWITCH_SOC, GAMMAX, GAMMAY, TOL
        WRITE (2,901) PRM%ALPHA(1), PRM%ALPHA(2), PRM%ALPHA(3)
        WRITE (2,'(A)') ' '
        WRITE (2,2902) NX, NY
        WRITE (2,903) NITER, NSTP, STP
        WRITE (2,2905) DX, DY
        WRITE (2,2906) DT, MAG
      CASE (3)
        WRITE (2,900) SWITCH_IM, OPTION_FPC, SWITCH_SOC, GAMMAX, GAMMAY, GAMMAZ, TOL
        WRITE (2,901) PRM%ALPHA(1), PRM%ALPHA(2), PRM%ALPHA(3)
        WRITE (2,'(A)') ' '
        WRITE (2,902) NX, NY, NZ
        WRITE (2,903) NITER, NSTP, STP
        WRITE (2,904) DX, DY, DZ
        WRITE (2,905) DT, MAG
    END SELECT
    WRITE (2,'(A)') ' '
    WRITE (2,908) AOSC
    IF (SPIN == 1) THEN
      IF (DIM == 3) THEN
        WRITE (2,906) A0, A2
        WRITE (2,907) PRM%TAU(1)/2.0D0, PRM%TAU(2)/2.0D0
      ELSE
        WRITE (2,1904) A0, A2
        WRITE (2,1908) PRM%TAU(1)/2.0D0, PRM%TAU(2)/2.0D0
      END IF
    ELSE
      WRITE (2,911) A0, A2, A4
      WRITE (2,912) PRM%TAU(1)/2.0D0, PRM%TAU(2)/2.0D0, PRM%TAU(3)/2.0D0
    END IF
    WRITE (2,'(A)') ' '
899 FORMAT(' OPENMP_THREADS = ', I3,', FFTW_THREADS = ', I3)
900 FORMAT(' SWITCH_IM = ', I2,', OPTION_FPC = ', I2, ', SWITCH_SOC = ', &
           I2, ', GAMMAX = ', F6.2, ', GAMMAY = ', F6.2, ', GAMMAZ = ', F6.2, ', TOL = ', ES9.2)
901 FORMAT(' Anisotropy ALPHAX =',F12.6,', ALPHAY = ',F12.6, ', ALPHAZ = ',F12.6)
902 FORMAT(' No. of space steps NX = ',I8, ', NY = ', I8, ', NZ = ', I8)
903 FORMAT(' No. of time steps : NITER = ',I9,', NSTP = ',I9,', STP = ',I9)
904 FORMAT(' Spatial step-size DX = ',F10.6,', DY = ',F10.6,', DZ = ',F10.6)
905 FORMAT(' Temporal step-size  DT = ', F10.6,', MAG = ', F12.6)
907 FORMAT(' TAU(0) = ', F12.6,', TAU(1) = ', F12.6, ' in dimensionless units ')
906 FORMAT(' A0 = ', F12.6, ', A2 = ', F12.6, ' (all in units of Bohr radius)')
908 FORMAT(' AOSC = ', E12.4)
911 FORMAT(' A0 = ', F12.6, ', A2 = ', F12.6, ', A4 = ', F12.6, ' (all in units of Bohr radius) ')
912 FORMAT(' TAU(0) = ', F12.6,', TAU(1) = ', F12.6,', TAU(2) = ', F12.6, ' (in dimensionless units)')
1900 FORMAT(' SWITCH_IM = ',I2,', OPTION_FPC = ', I2, ', SWITCH_SOC = ', I2, ', GAMMAX = ', F6.2, ', TOL = ', ES9.2)
1901 FORMAT(' Anisotropy ALPHAX = ',F12.6,', ALPHAY = ',F12.6,', ALPHAZ = ',F12.6)
1902 FORMAT(' No. of space steps NX = ',I8)
1903 FORMAT(' No. of time steps : NITER = ',I9,', NSTP = ',I9,', STP = ',I7)
1904 FORMAT(' A0 = ', F12.6, ', A2 = ', F12.6, ',  (all in units of Bohr radius) ')
1905 FORMAT(' Space Step DX = ',F10.6)
1906 FORMAT(' Time Step  DT = ', F10.6, ', MAG =',F12.6)
1908 FORMAT(' TAU(0) = ', F12.6,', TAU(1) = ', F12.6,' (in dimensionless units)')
2900 FORMAT(' SWITCH_IM = ', I2,', OPTION_FPC = ', I2, ', SWITCH_SOC = ', &
            I2, ', GAMMAX = ', F6.2, ', GAMMAY = ', F6.2, ', TOL = ', ES9.2)
2902 FORMAT(' No. of space steps NX = ',I8, ', NY = ', I8)
2905 FORMAT(' Space Step DX = ',F10.6,', DY = ',F10.6)
2906 FORMAT(' Time Step  DT = ', F10.6,', MAG = ',F12.6)
  END SUBROUTINE HEADER

  ! "Initial" rows with their rulers: formats 1001-1003/1011-1013 of each driver
  SUBROUTINE ROW_INITIAL()
    IF (SPIN == 1) THEN
      CALL RULER()
      IF (DIM == 3) THEN
        WRITE (2,1002)
      ELSE
        WRITE (2,1002)
      END IF
      CALL RULER()
      WRITE (2,1003) SUM(NORM(1:MS)), EN/2.0D0, (MU(L)/2.0D0, L = 1, MS), (CEN(L), L = 1, MS)
    ELSE
      CALL RULER()
      SELECT CASE (DIM)
        CASE (1)
          WRITE (2,1112)
        CASE (2)
          WRITE (2,2112)
        CASE (3)
          WRITE (2,3112)
      END SELECT
      CALL RULER()
      WRITE (2,1013) SUM(NORM(1:MS)), EN/2.0D0, (MU(L)/2.0D0, L = 1, MS), (CEN(L), L = 1, MS)
    END IF
1002 FORMAT (23X, 'Norm', 7X,'EN', 7X, 'MU1', 7X, 'MU2', 7X, 'MU3', 6X, 'phi1', 6X, 'phi2', 6X,'phi3')
1003 FORMAT ('Initial: ', 8X, 1F11.4, 4F10.5, 3F10.5)
1112 FORMAT (23X, 'Norm', 7X,'EN', 7X, 'MU1', 7X, 'MU2', 7X, 'MU3', 7X, 'MU4', 7X, 'MU5', 7X, &
             'phi1', 6X, 'phi2', 6X,'phi3',5X, 'phi4', 6X,'phi5')
2112 FORMAT (23X, 'Norm', 7X,'EN', 7X, 'MU1', 7X, 'MU2', 7X, 'MU3', 7X, 'MU4', 7X, 'MU5', 7X, &
             'phi1', 7X, 'phi2', 5X,'phi3',5X, 'phi4', 5X,'phi5')
3112 FORMAT (23X, 'Norm', 7X,'EN', 7X, 'MU1', 7X, 'MU2', 7X, 'MU3', 7X, 'MU4', 7X, 'MU5', 7X, &
             'phi1', 7X, 'phi2', 5X,'phi3',5X, 'phi4', 6X,'phi5')
1013 FORMAT ('Initial : ', 7X, 1F11.4, 6F10.5, 5F10.5)
  END SUBROUTINE ROW_INITIAL

  SUBROUTINE RULER()
    IF (SPIN == 1) THEN
      IF (DIM == 3) THEN
        WRITE (2,'(23X, 74(''-''))')
      ELSE
        WRITE (2,'(23X, 76(''-''))')
      END IF
    ELSE
      WRITE (2,'(22X, 117(''-''))')
    END IF
  END SUBROUTINE RULER

  SUBROUTINE ROW_NSTP()
    IF (DIM == 3) THEN
      WRITE (2,3005) SUM(NORM(1:MS)), EN/2.0D0, (MU(L)/2.0D0, L = 1, MS), (CEN(L), L = 1, MS)
    ELSE
      WRITE (2,1005) SUM(NORM(1:MS)), EN/2.0D0, (MU(L)/2.0D0, L = 1, MS), (CEN(L), L = 1, MS)
    END IF
1005 FORMAT('After NSTP iter.    :', 1F7.4, 6F10.5, 5F10.5)
3005 FORMAT('After NSTP iter.:    ', 1F7.4, 6F10.5, 5F10.5)
  END SUBROUTINE ROW_NSTP

  SUBROUTINE ROW_FINAL(IT)
    INTEGER, INTENT(IN) :: IT
    CALL RULER()
    WRITE (2,1006) IT, SUM(NORM(1:MS)), EN/2.0D0, (MU(L)/2.0D0, L = 1, MS), (CEN(L), L = 1, MS)
    CALL RULER()
    WRITE (2,'(A)') ' '
1006 FORMAT('After ', I8 ,' iter.:', 1F7.4, 6F10.5, 5F10.5)
  END SUBROUTINE ROW_FINAL

  ! transient dumps every NSTP: 1D SRC/cgpe1d.f90:510-518, 2D
  ! SRC/cgpe2d.f90:519-530, 3D (full, xy and xz slices) SRC/cgpe3d.f90:532-566
  SUBROUTINE DUMP_TMP()
    ! NSTP dumps: 1D SRC/cgpe1d.f90:513-516, 2D SRC/cgpe2d.f90:521-528,
    ! 3D SRC/cgpe3d.f90:534-563 (full grid, then the z = NZ/2+1 and y = NY/2+1
    ! slices), formatted on the input's thread count
    SELECT CASE (DIM)
      CASE (1)
        CALL DUMP('./OUTPUT/tmp_solution_file_spin'//SPIN_ID//'.dat', FGPE_DUMP_1D_TMP)
      CASE (2)
        CALL DUMP('./OUTPUT/tmp_solution_file_spin'//SPIN_ID//'.dat', FGPE_DUMP_2D)
      CASE (3)
        CALL DUMP('./OUTPUT/tmp_solution_file_spin'//SPIN_ID//'.dat', FGPE_DUMP_3D)
        CALL DUMP('./OUTPUT/tmp_solution_file_spin'//SPIN_ID//'_xy.dat', FGPE_DUMP_3D_XY)
        CALL DUMP('./OUTPUT/tmp_solution_file_spin'//SPIN_ID//'_xz.dat', FGPE_DUMP_3D_XZ)
    END SELECT
  END SUBROUTINE DUMP_TMP

  ! final solution: 1D SRC/cgpe1d.f90:547-549, 2D SRC/cgpe2d.f90:560-565,
  ! 3D SRC/cgpe3d.f90:596-602
  SUBROUTINE DUMP_SOLUTION()
    SELECT CASE (DIM)
      CASE (1)
        CALL DUMP('./OUTPUT/solution_file_'//TAG//'_spin'//SPIN_ID//'.dat', FGPE_DUMP_1D_SOL)
      CASE (2)
        CALL DUMP('./OUTPUT/solution_file_'//TAG//'_spin'//SPIN_ID//'.dat', FGPE_DUMP_2D)
      CASE (3)
        CALL DUMP('./OUTPUT/solution_file_'//TAG//'_spin'//SPIN_ID//'.dat', FGPE_DUMP_3D)
    END SELECT
  END SUBROUTINE DUMP_SOLUTION

  SUBROUTINE DUMP(PATH, KIND)
    CHARACTER(LEN=*), INTENT(IN) :: PATH
    INTEGER(C_INT32_T), INTENT(IN) :: KIND
    RC = FGPE_WRITE_DUMP(PATH//C_NULL_CHAR, KIND, INP%N, INT(MS, C_INT32_T), PHI, X, Y, Z, &
                         INT(THREADS, C_INT32_T))
    IF (RC /= 0) CALL FAIL_STOP(RC)
  END SUBROUTINE DUMP
END PROGRAM CGPE
