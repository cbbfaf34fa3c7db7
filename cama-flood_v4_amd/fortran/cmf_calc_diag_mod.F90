MODULE CMF_CALC_DIAG_MOD
!==========================================================
!* Device-side replacement of the reference module CMF_CALC_DIAG_MOD
!  (src/cmf_calc_diag_mod.F90) for the drop-in build (cmf_dropin_devdiag).
!
!  Output-period diagnostics (diag:222-356) run on the device, next to the
!  substep averages the device physics keeps:
!    CMF_DIAG_AVEMAX_OUTPUT  -- cmf_hip_diag_avemax_output (every outer step)
!    CMF_DIAG_GETAVE_OUTPUT  -- cmf_hip_diag_getave_output, then the averaged
!                               fields are copied into the host D2*_oAVG,
!                               D2*_oMAX and D1PTHFLW_oAVG arrays, which
!                               CMF_OUTPUT_WRITE reads unchanged
!    CMF_DIAG_RESET_OUTPUT   -- cmf_hip_diag_reset_output
!  NADD_out is kept on the host as the reference keeps it. With this module the
!  physics module no longer downloads the substep averages (D2*_aAVG, D2*_aMAX)
!  after every step: nothing else on the host reads them (the tracer, which does,
!  is not supported by the device physics).
!
!  The substep routines (RESET/AVEMAX/GETAVE_ADPSTP, diag:50-215) run inside the
!  device's CMF_PHYSICS_ADVANCE; they keep the reference's names for its
!  interface and stop if called.
!==========================================================
USE ISO_C_BINDING
USE PARKIND1,              ONLY: JPIM, JPRB
IMPLICIT NONE
PRIVATE
PUBLIC :: CMF_DIAG_RESET_ADPSTP, CMF_DIAG_AVEMAX_ADPSTP, CMF_DIAG_GETAVE_ADPSTP
PUBLIC :: CMF_DIAG_RESET_OUTPUT, CMF_DIAG_AVEMAX_OUTPUT, CMF_DIAG_GETAVE_OUTPUT

TYPE, BIND(C) :: CMF_OUTPUT_T
  TYPE(C_PTR) :: RIVOUT, FLDOUT, OUTFLW, RIVVEL, PTHOUT, RUNOFF, DAMINF
  TYPE(C_PTR) :: OUTFLW_MAX, RIVDPH_MAX, STORGE_MAX
  TYPE(C_PTR) :: PTHFLW
END TYPE

INTERFACE
  INTEGER(C_INT) FUNCTION cmf_hip_diag_reset_output(h) BIND(C, NAME='cmf_hip_diag_reset_output')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_diag_avemax_output(h, dt) BIND(C, NAME='cmf_hip_diag_avemax_output')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), VALUE :: dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_diag_getave_output(h, nadd) BIND(C, NAME='cmf_hip_diag_getave_output')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(OUT) :: nadd
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_get_output(h, o) BIND(C, NAME='cmf_hip_get_output')
    IMPORT :: C_INT, C_PTR, CMF_OUTPUT_T
    TYPE(C_PTR), VALUE :: h
    TYPE(CMF_OUTPUT_T), INTENT(IN) :: o     ! the struct holds host addresses; C writes through them
  END FUNCTION
END INTERFACE

CONTAINS
!####################################################################
SUBROUTINE CMF_DIAG_RESET_ADPSTP
CALL CMF_HIP_STOP('CMF_DIAG_RESET_ADPSTP runs inside the device CMF_PHYSICS_ADVANCE')
END SUBROUTINE CMF_DIAG_RESET_ADPSTP
SUBROUTINE CMF_DIAG_AVEMAX_ADPSTP
CALL CMF_HIP_STOP('CMF_DIAG_AVEMAX_ADPSTP runs inside the device CMF_PHYSICS_ADVANCE')
END SUBROUTINE CMF_DIAG_AVEMAX_ADPSTP
SUBROUTINE CMF_DIAG_GETAVE_ADPSTP
CALL CMF_HIP_STOP('CMF_DIAG_GETAVE_ADPSTP runs inside the device CMF_PHYSICS_ADVANCE')
END SUBROUTINE CMF_DIAG_GETAVE_ADPSTP
!####################################################################
SUBROUTINE CMF_DIAG_RESET_OUTPUT
!! diag:222-267: accumulators and NADD_out to zero (the device ones and the host copies)
USE YOS_CMF_INPUT,  ONLY: LOGNAM
USE YOS_CMF_TIME,   ONLY: JYYYYMMDD, JHHMM
USE YOS_CMF_DIAG,   ONLY: NADD_out
USE CMF_CTRL_PHYSICS_MOD, ONLY: CMF_HIP_CHECK, LHIP_OUTDIAG
IMPLICIT NONE
WRITE(LOGNAM,*) "CMF::DIAG_AVERAGE: reset", JYYYYMMDD, JHHMM
LHIP_OUTDIAG = .TRUE.
CALL CMF_HIP_CHECK(cmf_hip_diag_reset_output(HANDLE()), 'diag_reset_output')
CALL ZERO_HOST
NADD_out = 0._JPRB
END SUBROUTINE CMF_DIAG_RESET_OUTPUT
!####################################################################
SUBROUTINE CMF_DIAG_AVEMAX_OUTPUT
!! diag:273-317: o += aAVG * DT, o = MAX(o, aMAX), NADD_out += DT
USE YOS_CMF_INPUT,  ONLY: DT
USE YOS_CMF_DIAG,   ONLY: NADD_out
USE CMF_CTRL_PHYSICS_MOD, ONLY: CMF_HIP_CHECK, LHIP_OUTDIAG
IMPLICIT NONE
LHIP_OUTDIAG = .TRUE.
NADD_out = NADD_out + DT
CALL CMF_HIP_CHECK(cmf_hip_diag_avemax_output(HANDLE(), REAL(DT, C_DOUBLE)), 'diag_avemax_output')
END SUBROUTINE CMF_DIAG_AVEMAX_OUTPUT
!####################################################################
SUBROUTINE CMF_DIAG_GETAVE_OUTPUT
!! diag:323-356: averages / NADD_out, then the output fields to the host arrays
USE YOS_CMF_INPUT,  ONLY: LOGNAM, LPTHOUT, LDAMOUT
USE YOS_CMF_TIME,   ONLY: JYYYYMMDD, JHHMM
USE YOS_CMF_DIAG,   ONLY: NADD_out, D2RIVOUT_oAVG, D2FLDOUT_oAVG, D2OUTFLW_oAVG, D2RIVVEL_oAVG, &
                        & D2PTHOUT_oAVG, D2RUNOFF_oAVG, D2DAMINF_oAVG, D2OUTFLW_oMAX, D2RIVDPH_oMAX, &
                        & D2STORGE_oMAX, D1PTHFLW_oAVG
USE CMF_CTRL_PHYSICS_MOD, ONLY: CMF_HIP_CHECK
IMPLICIT NONE
REAL(C_DOUBLE)     :: ZNADD
TYPE(CMF_OUTPUT_T) :: O
!================================================
WRITE(LOGNAM,*) "CMF::DIAG_AVERAGE: time-average", NADD_out, JYYYYMMDD, JHHMM
CALL CMF_HIP_CHECK(cmf_hip_diag_getave_output(HANDLE(), ZNADD), 'diag_getave_output')
O = CMF_OUTPUT_T(C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, &
  & C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR)
O%RIVOUT = PTR_D2(D2RIVOUT_oAVG);  O%FLDOUT = PTR_D2(D2FLDOUT_oAVG);  O%OUTFLW = PTR_D2(D2OUTFLW_oAVG)
O%RIVVEL = PTR_D2(D2RIVVEL_oAVG);  O%PTHOUT = PTR_D2(D2PTHOUT_oAVG);  O%RUNOFF = PTR_D2(D2RUNOFF_oAVG)
IF( LDAMOUT .AND. ALLOCATED(D2DAMINF_oAVG) ) O%DAMINF = PTR_D2(D2DAMINF_oAVG)
O%OUTFLW_MAX = PTR_D2(D2OUTFLW_oMAX);  O%RIVDPH_MAX = PTR_D2(D2RIVDPH_oMAX)
O%STORGE_MAX = PTR_D2(D2STORGE_oMAX)
IF( LPTHOUT ) O%PTHFLW = PTR_D2(D1PTHFLW_oAVG)
CALL CMF_HIP_CHECK(cmf_hip_get_output(HANDLE(), O), 'get_output')
END SUBROUTINE CMF_DIAG_GETAVE_OUTPUT
!####################################################################
TYPE(C_PTR) FUNCTION HANDLE()
!! the device handle of CMF_PHYSICS_ADVANCE, which opens it on its first call
USE CMF_CTRL_PHYSICS_MOD, ONLY: CMF_HIP_HANDLE
IMPLICIT NONE
HANDLE = CMF_HIP_HANDLE()
IF( .NOT. C_ASSOCIATED(HANDLE) ) &
  & CALL CMF_HIP_STOP('output diagnostics before the first CMF_PHYSICS_ADVANCE')
END FUNCTION HANDLE

SUBROUTINE ZERO_HOST
USE YOS_CMF_DIAG,   ONLY: D2RIVOUT_oAVG, D2FLDOUT_oAVG, D2OUTFLW_oAVG, D2RIVVEL_oAVG, D2PTHOUT_oAVG, &
                        & D2RUNOFF_oAVG, D2DAMINF_oAVG, D2OUTFLW_oMAX, D2RIVDPH_oMAX, D2STORGE_oMAX, &
                        & D1PTHFLW_oAVG
IMPLICIT NONE
D2RIVOUT_oAVG(:,:) = 0._JPRB;  D2FLDOUT_oAVG(:,:) = 0._JPRB;  D2OUTFLW_oAVG(:,:) = 0._JPRB
D2RIVVEL_oAVG(:,:) = 0._JPRB;  D2PTHOUT_oAVG(:,:) = 0._JPRB;  D2RUNOFF_oAVG(:,:) = 0._JPRB
IF( ALLOCATED(D2DAMINF_oAVG) ) D2DAMINF_oAVG(:,:) = 0._JPRB
D2OUTFLW_oMAX(:,:) = 0._JPRB;  D2RIVDPH_oMAX(:,:) = 0._JPRB;  D2STORGE_oMAX(:,:) = 0._JPRB
IF( ALLOCATED(D1PTHFLW_oAVG) ) D1PTHFLW_oAVG(:,:) = 0._JPRB
END SUBROUTINE ZERO_HOST

SUBROUTINE CMF_HIP_STOP(CMSG)
USE YOS_CMF_INPUT,  ONLY: LOGNAM
IMPLICIT NONE
CHARACTER(LEN=*), INTENT(IN) :: CMSG
WRITE(LOGNAM,*) 'CMF_CALC_DIAG_MOD (HIP): ', CMSG
WRITE(*,*) 'CMF_CALC_DIAG_MOD (HIP): ', CMSG
STOP 9
END SUBROUTINE CMF_HIP_STOP

TYPE(C_PTR) FUNCTION PTR_D2(A)
!! address of a contiguous (NSEQMAX,1) or (NPTHOUT,NPTHLEV) array
REAL(KIND=JPRB), TARGET, INTENT(IN) :: A(:,:)
PTR_D2 = C_LOC(A(1,1))
END FUNCTION PTR_D2

END MODULE CMF_CALC_DIAG_MOD
