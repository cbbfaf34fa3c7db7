!==========================================================
! Stand-alone host program of the drop-in build (cmf_dropin_devdiag): the
! reference's initialisation and time loop (CMF_DRV_INPUT / CMF_DRV_INIT /
! CMF_DRV_ADVANCE / CMF_DRV_END, as src/MAIN_cmf.F90 drives them) with the
! device physics and the device-side output diagnostics linked in.
!
! usage: cmf_dropin_devdiag <runoff.bin> [ignored arguments]
!   runoff.bin : records of float64 runoff, one per forcing interval DTIN,
!                repeated cyclically (the file reading of CMF_FORCING_GET is not
!                part of this program). Without LINTERP: NX*NY values [m3/s]
!                assigned to D2RUNOFF of each cell directly. With LINTERP (namelist
!                NFORCE, input matrix CINPMAT): NXIN*NYIN values in the input unit
!                (DROFUNIT), passed as the first field of PBUFF to the forcing put
!                of the main loop of src/MAIN_cmf.F90: CMF_HIP_FORCING_PUT
!                (interpolation on the device), or the reference's CMF_FORCING_PUT
!                on the host when CMF_HIP_FORCING=host is set; with
!                CMF_HIP_FORCING=resident every record is copied to HBM before the
!                time loop (CMF_HIP_FORCING_RECORDS) and each forcing step selects
!                its record there (CMF_HIP_FORCING_PUT_RECORD)
!   stdout     : "STEPTIME <step> <seconds>" per outer step (time inside
!                CMF_DRV_ADVANCE, and with LINTERP inside the forcing put too),
!                then "RUNTIME <seconds> <nsteps> <nseqall>"
! Output and restart files are written by the reference's own writers as the
! namelist (input_cmf.nam) asks.
!==========================================================
PROGRAM CMF_DROPIN_MAIN
USE PARKIND1,                ONLY: JPIM, JPRB, JPRD
USE YOS_CMF_INPUT,           ONLY: NX, NY, NXIN, NYIN, DT, DTIN, LOUTPUT, IFRQ_OUT
USE YOS_CMF_TIME,            ONLY: NSTEPS
USE YOS_CMF_MAP,             ONLY: NSEQALL, I1SEQX, I1SEQY
USE YOS_CMF_PROG,            ONLY: D2RUNOFF
USE CMF_DRV_CONTROL_MOD,     ONLY: CMF_DRV_INPUT, CMF_DRV_INIT, CMF_DRV_END
USE CMF_DRV_ADVANCE_MOD,     ONLY: CMF_DRV_ADVANCE
USE CMF_CTRL_FORCING_MOD,    ONLY: CMF_FORCING_PUT, LINTERP
USE CMF_CTRL_PHYSICS_MOD,    ONLY: CMF_HIP_FORCING_PUT, CMF_HIP_FORCING_RECORDS, CMF_HIP_FORCING_PUT_RECORD
IMPLICIT NONE
CHARACTER(LEN=512)           :: CROF
CHARACTER(LEN=16)            :: CPUT
INTEGER(KIND=JPIM)           :: ISTEP, NPERREC, IREC, NREC, ISEQ, IU, MX, MY
REAL(KIND=JPRD), ALLOCATABLE :: ROF(:,:)
REAL(KIND=JPRB), ALLOCATABLE :: ZBUFF(:,:,:), ZALL(:,:,:)
LOGICAL                      :: LNEW, LHOSTPUT, LRESIDENT
INTEGER(KIND=8)              :: T0, T1, RATE, TSUM, NBYTES

CALL GET_COMMAND_ARGUMENT(1, CROF)
CALL CMF_DRV_INPUT
!! the output predicate of CMF_DRV_ADVANCE (drv_advance:121) takes MOD(JHOUR,
!! IFRQ_OUT) even with LOUTPUT off, when the output namelist left IFRQ_OUT at 0
IF( .NOT. LOUTPUT .AND. IFRQ_OUT <= 0 ) IFRQ_OUT = 24
CALL CMF_DRV_INIT

MX = NX;  MY = NY
IF( LINTERP )THEN
  MX = NXIN;  MY = NYIN
  ALLOCATE(ZBUFF(NXIN,NYIN,1))
ENDIF
CALL GET_ENVIRONMENT_VARIABLE('CMF_HIP_FORCING', CPUT)
LHOSTPUT = TRIM(CPUT) == 'host'
LRESIDENT = LINTERP .AND. TRIM(CPUT) == 'resident'
ALLOCATE(ROF(MX,MY))
OPEN(NEWUNIT=IU, FILE=TRIM(CROF), FORM='UNFORMATTED', ACCESS='STREAM', STATUS='OLD')
INQUIRE(UNIT=IU, SIZE=NBYTES)
NREC = INT(NBYTES / (8_8*MX*MY), JPIM)
IF( NREC < 1 ) STOP 'cmf_dropin_main: runoff file holds no float64 record of the forcing grid'
NPERREC = MAX(1, INT(DTIN/DT, JPIM))
IF( LRESIDENT )THEN
  ALLOCATE(ZALL(NXIN,NYIN,NREC))
  DO IREC=1, NREC
    READ(IU, POS=1_8 + INT(IREC-1,8)*8_8*MX*MY) ROF
    ZALL(:,:,IREC) = REAL(ROF, JPRB)
  END DO
  CALL CMF_HIP_FORCING_RECORDS(ZALL)
  DEALLOCATE(ZALL)
ENDIF

TSUM = 0
IREC = 0
DO ISTEP=1, NSTEPS
  LNEW = MOD(ISTEP-1, NPERREC) == 0
  IF( LNEW )THEN
    IREC = MOD(IREC, NREC) + 1
  ENDIF
  IF( LNEW .AND. .NOT. LRESIDENT )THEN
    READ(IU, POS=1_8 + INT(IREC-1,8)*8_8*MX*MY) ROF
    IF( LINTERP )THEN
      ZBUFF(:,:,1) = REAL(ROF, JPRB)
    ELSE
      DO ISEQ=1, NSEQALL
        D2RUNOFF(ISEQ,1) = ROF(I1SEQX(ISEQ), I1SEQY(ISEQ))
      END DO
    ENDIF
  ENDIF
  CALL SYSTEM_CLOCK(T0, RATE)
  IF( LINTERP .AND. LNEW )THEN
    IF( LRESIDENT )THEN
      CALL CMF_HIP_FORCING_PUT_RECORD(IREC)
    ELSEIF( LHOSTPUT )THEN
      CALL CMF_FORCING_PUT(ZBUFF)
    ELSE
      CALL CMF_HIP_FORCING_PUT(ZBUFF)
    ENDIF
  ENDIF
  CALL CMF_DRV_ADVANCE(1)
  CALL SYSTEM_CLOCK(T1)
  TSUM = TSUM + (T1 - T0)
  WRITE(*,'(A,1X,I10,1X,ES16.8)') 'STEPTIME', ISTEP, REAL(T1-T0,8)/REAL(RATE,8)
END DO
WRITE(*,'(A,1X,ES16.8,1X,I10,1X,I10)') 'RUNTIME', REAL(TSUM,8)/REAL(RATE,8), NSTEPS, NSEQALL
CLOSE(IU)
CALL CMF_DRV_END
END PROGRAM CMF_DROPIN_MAIN
