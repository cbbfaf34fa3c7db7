MODULE CMF_CTRL_PHYSICS_MOD
!==========================================================
!* Drop-in replacement of the reference module CMF_CTRL_PHYSICS_MOD
!  (src/cmf_ctrl_physics_mod.F90) that runs the per-substep physics on an
!  MI355X through the C-ABI of include/cmf_hip.h (libcmf_hip.so).
!
!  Public names and arguments are those of the reference, so CMF_DRV_ADVANCE
!  (src/cmf_drv_advance_mod.F90:39,101) and CMF_DRV_INIT
!  (src/cmf_drv_control_mod.F90:171,243,276) compile unchanged against it:
!    CMF_PHYSICS_ADVANCE  -- one outer step (physics:21-155)
!    CMF_PHYSICS_FLDSTG   -- storage -> stage (physics:288-305)
!  and, for a main program or coupler that calls it in place of CMF_FORCING_PUT
!  (src/cmf_ctrl_forcing_mod.F90:716-750), CMF_HIP_FORCING_PUT: the LINTERP runoff
!  interpolation on the device, so that D2RUNOFF is formed in HBM.
!
!  The model state lives in HBM. The first CMF_PHYSICS_ADVANCE (after the whole
!  initialisation, dams and restart included) uploads the map and the state;
!  later steps upload the runoff forcing. Downloads follow the host's readers:
!  the substep averages every step for the host's CMF_DIAG_AVEMAX_OUTPUT
!  (drv_advance:114; not with this package's device-side CMF_CALC_DIAG_MOD), and
!  the prognostic and diagnostic state only on steps where the host reads it:
!  output steps (the predicate of drv_advance:121: output files, dam and text
!  output) and restart steps (restart_is_write_time of CMF_CTRL_RESTART_MOD, the
!  predicate of CMF_RESTART_WRITE). CMF_HIP_DOWNLOAD=every in the environment
!  downloads the whole state after every step instead (for hosts that read it
!  elsewhere). CMF_PHYSICS_FLDSTG before the first advance (drv_control:243,276)
!  runs on a transient handle.
!  Links into either build of the reference: FP64 (JPRB = 8 bytes) or the default
!  single precision (-DSinglePrec_CMF, JPRB = 4: the JPRB arrays cross the C-ABI
!  as c_float, the P2* storages as c_double). The device path defaults to FP64
!  under the FP64 host and to the mixed-precision path (FP32 flux arithmetic and
!  flux/stage arrays; FP64 storage, storage update and stage inversion) under the
!  single-precision host;
!  CMF_HIP_PRECISION=fp64 or =mixed in the environment chooses it explicitly.
!  Options carried to the device: LKINE, LSLPMIX (with the host's I2MASK),
!  LDAMOUT with LDAMH22 (each dam's own FldVol*1.E-6/upreal) and LDAMYBY, LLEVEE,
!  LPTHOUT, LSEALEV, LUPSINF. The module stops with a message under LTRACE,
!  LSEDOUT and the options the device path does not carry (OPEN_HANDLE).
!==========================================================
USE ISO_C_BINDING
USE PARKIND1,              ONLY: JPIM, JPRB, JPRD
IMPLICIT NONE
PRIVATE
PUBLIC :: CMF_PHYSICS_ADVANCE, CMF_PHYSICS_FLDSTG
!! CMF_FORCING_PUT with the runoff interpolation on the device, for a main program
!! or coupler that calls it in place of the reference's (see CMF_HIP_FORCING_PUT)
PUBLIC :: CMF_HIP_FORCING_PUT
!! a run's forcing fields resident in HBM: copied once, one selected per forcing step
PUBLIC :: CMF_HIP_FORCING_RECORDS, CMF_HIP_FORCING_PUT_RECORD
!! for the device-side CMF_CALC_DIAG_MOD (cmf_calc_diag_mod.F90 of this package):
!! the open handle, the error check, and LHIP_OUTDIAG, which that module sets so
!! that the substep averages (D2*_aAVG, D2*_aMAX) are no longer downloaded after
!! every step (the output-period sums are then formed on the device)
PUBLIC :: CMF_HIP_HANDLE, CMF_HIP_CHECK, LHIP_OUTDIAG
LOGICAL, SAVE :: LHIP_OUTDIAG = .FALSE.
LOGICAL, SAVE :: LEVERY = .FALSE.          ! CMF_HIP_DOWNLOAD=every
LOGICAL, SAVE :: LDEVROF = .FALSE.         ! D2RUNOFF formed on the device (CMF_HIP_FORCING_PUT)
LOGICAL, SAVE :: LINPMAT = .FALSE.         ! input matrix uploaded

TYPE, BIND(C) :: CMF_DIMS_T
  INTEGER(C_INT32_T) :: NSEQALL, NSEQRIV, NLFP, NPTHOUT, NPTHLEV, NDAM
END TYPE
TYPE, BIND(C) :: CMF_OPTIONS_T
  INTEGER(C_INT32_T) :: LADPSTP, LFPLAIN, LFLDOUT, LPTHOUT, LDAMOUT, LLEVEE, PRECISION
  REAL(C_DOUBLE)     :: PGRV, PMANFLD, PDSTMTH, PCADP, PMINSLP
  INTEGER(C_INT32_T) :: JPRB, LDAMH22
END TYPE
TYPE, BIND(C) :: CMF_MAP_T
  TYPE(C_PTR)    :: I1NEXT, GRAREA, ELEVTN, NXTDST, RIVLEN, RIVWTH, RIVHGT, RIVMAN
  TYPE(C_PTR)    :: RIVSTOMAX, RIVELV, DWNELV, FLDHGT, FLDSTOMAX, FLDGRD
  REAL(C_DOUBLE) :: DFRCINC
  TYPE(C_PTR)    :: I2MASK, PTH_UPST, PTH_DOWN, PTH_DST, PTH_ELV, PTH_WTH, PTH_MAN
  TYPE(C_PTR)    :: LEVHGT, LEVFRC, BASHGT, LEVDST, LEVBASSTO, LEVTOPSTO, LEVFILSTO
  TYPE(C_PTR)    :: I1DAM, DAMSEQ, DAMSTAT, DAM_CONVOL, DAM_EMEVOL, DAM_ADJVOL, DAM_QA, DAM_QN, DAM_QF
  TYPE(C_PTR)    :: DAM_NORVOL, DAM_R_VOLUPA
END TYPE
TYPE, BIND(C) :: CMF_STATE_T
  TYPE(C_PTR) :: RIVSTO, FLDSTO, LEVSTO, DAMSTO, DAMINF
  TYPE(C_PTR) :: RIVOUT, FLDOUT, RIVOUT_PRE, FLDOUT_PRE, RIVDPH_PRE, FLDSTO_PRE
  TYPE(C_PTR) :: PTHFLW, PTHFLW_PRE
  TYPE(C_PTR) :: RIVDPH, FLDDPH, FLDFRC, FLDARE, SFCELV, STORGE
  TYPE(C_PTR) :: OUTFLW, RIVVEL, RIVINF, FLDINF, PTHOUT, LEVDPH, PTHFLWSUM
  TYPE(C_PTR) :: RIVOUT_AVG, FLDOUT_AVG, OUTFLW_AVG, RIVVEL_AVG, PTHOUT_AVG
  TYPE(C_PTR) :: RUNOFF_AVG, OUTFLW_MAX, RIVDPH_MAX, STORGE_MAX, DAMINF_AVG
  TYPE(C_PTR) :: PTHFLW_AVG, PTHFLWSUM_AVG
END TYPE

INTERFACE
  INTEGER(C_INT) FUNCTION cmf_hip_create(dims, opts, h) BIND(C, NAME='cmf_hip_create')
    IMPORT :: C_INT, C_PTR, CMF_DIMS_T, CMF_OPTIONS_T
    TYPE(CMF_DIMS_T),    INTENT(IN)  :: dims
    TYPE(CMF_OPTIONS_T), INTENT(IN)  :: opts
    TYPE(C_PTR),         INTENT(OUT) :: h
  END FUNCTION
  SUBROUTINE cmf_hip_destroy(h) BIND(C, NAME='cmf_hip_destroy')
    IMPORT :: C_PTR
    TYPE(C_PTR), VALUE :: h
  END SUBROUTINE
  TYPE(C_PTR) FUNCTION cmf_hip_last_error() BIND(C, NAME='cmf_hip_last_error')
    IMPORT :: C_PTR
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_map(h, mp) BIND(C, NAME='cmf_hip_set_map')
    IMPORT :: C_INT, C_PTR, CMF_MAP_T
    TYPE(C_PTR), VALUE :: h
    TYPE(CMF_MAP_T), INTENT(IN) :: mp
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_state(h, st) BIND(C, NAME='cmf_hip_set_state')
    IMPORT :: C_INT, C_PTR, CMF_STATE_T
    TYPE(C_PTR), VALUE :: h
    TYPE(CMF_STATE_T), INTENT(IN) :: st
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_get_state(h, st) BIND(C, NAME='cmf_hip_get_state')
    IMPORT :: C_INT, C_PTR, CMF_STATE_T
    TYPE(C_PTR), VALUE :: h
    TYPE(CMF_STATE_T), INTENT(IN) :: st     ! the struct holds host addresses; C writes through them
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_runoff(h, r) BIND(C, NAME='cmf_hip_set_runoff')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h, r
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_inpmat(h, inpn, nxin, nyin, inpx, inpy, inpa, drofunit, rmis) &
      & BIND(C, NAME='cmf_hip_set_inpmat')
    IMPORT :: C_INT, C_PTR, C_INT32_T, C_DOUBLE
    TYPE(C_PTR), VALUE :: h, inpx, inpy, inpa
    INTEGER(C_INT32_T), VALUE :: inpn, nxin, nyin
    REAL(C_DOUBLE), VALUE :: drofunit, rmis
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_runoff_grid(h, grid) BIND(C, NAME='cmf_hip_set_runoff_grid')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h, grid
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_runoff_records(h, nrec, grids) BIND(C, NAME='cmf_hip_set_runoff_records')
    IMPORT :: C_INT, C_PTR, C_INT32_T
    TYPE(C_PTR), VALUE :: h, grids
    INTEGER(C_INT32_T), VALUE :: nrec
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_use_runoff_record(h, k) BIND(C, NAME='cmf_hip_use_runoff_record')
    IMPORT :: C_INT, C_PTR, C_INT32_T
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT32_T), VALUE :: k
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_dwnelv(h, d) BIND(C, NAME='cmf_hip_set_dwnelv')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h, d
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_upsinf(h, u) BIND(C, NAME='cmf_hip_set_upsinf')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h, u
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_kinematic(h, lkine) BIND(C, NAME='cmf_hip_set_kinematic')
    IMPORT :: C_INT, C_PTR, C_INT32_T
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT32_T), VALUE :: lkine
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_set_slopemix(h, lslpmix) BIND(C, NAME='cmf_hip_set_slopemix')
    IMPORT :: C_INT, C_PTR, C_INT32_T
    TYPE(C_PTR), VALUE :: h
    INTEGER(C_INT32_T), VALUE :: lslpmix
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_physics_fldstg(h) BIND(C, NAME='cmf_hip_physics_fldstg')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: h
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_get_dt_min(h, dtmin) BIND(C, NAME='cmf_hip_get_dt_min')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(OUT) :: dtmin
  END FUNCTION
  INTEGER(C_INT) FUNCTION cmf_hip_physics_advance(h, dt, nt, glb, cap) BIND(C, NAME='cmf_hip_physics_advance')
    IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
    TYPE(C_PTR), VALUE :: h
    REAL(C_DOUBLE), INTENT(INOUT) :: dt
    INTEGER(C_INT32_T), INTENT(OUT) :: nt
    TYPE(C_PTR), VALUE :: glb
    INTEGER(C_INT32_T), VALUE :: cap
  END FUNCTION
END INTERFACE

TYPE(C_PTR), SAVE            :: HMAIN = C_NULL_PTR
INTEGER(C_INT32_T), PARAMETER :: NGLB = 11, GLBCAP = 4096
REAL(C_DOUBLE), SAVE, TARGET  :: GLB(NGLB, GLBCAP)
!! LDAMH22: FldVol*1.E-6/upreal of every dam at its own index (the module's
!! R_VolUpa holds only one value, at index NDAM: damout:204)
REAL(KIND=JPRB), ALLOCATABLE, SAVE, TARGET :: RVOLUPA(:)

CONTAINS
!####################################################################
SUBROUTINE CMF_PHYSICS_ADVANCE
USE YOS_CMF_INPUT,  ONLY: LOGNAM, DT, LSEALEV, LUPSINF
USE YOS_CMF_PROG,   ONLY: D2RUNOFF, D2UPSINF
USE YOS_CMF_MAP,    ONLY: D2DWNELV
IMPLICIT NONE
REAL(C_DOUBLE)     :: ZDT, ZDTMIN
INTEGER(C_INT32_T) :: NT, IT
INTEGER(C_INT)     :: IRC
!================================================
IF( .NOT. C_ASSOCIATED(HMAIN) )THEN
  CALL OPEN_HANDLE(HMAIN, .TRUE.)
  CALL CHECK(cmf_hip_set_state(HMAIN, STATE_PTRS(.TRUE.)), 'set_state')
ENDIF
!! D2RUNOFF from the host, unless CMF_HIP_FORCING_PUT formed it on the device
IF( .NOT. LDEVROF ) CALL CHECK(cmf_hip_set_runoff(HMAIN, PTR_D(D2RUNOFF)), 'set_runoff')
!! CMF_BOUNDARY_UPDATE (drv_advance:84) moves D2DWNELV every step with LSEALEV
IF( LSEALEV ) CALL CHECK(cmf_hip_set_dwnelv(HMAIN, PTR_D2(D2DWNELV)), 'set_dwnelv')
!! CMF_UPSINF_UPDATE (drv_advance:87) sets D2UPSINF; stonxt:81 adds it
IF( LUPSINF ) CALL CHECK(cmf_hip_set_upsinf(HMAIN, PTR_D(D2UPSINF)), 'set_upsinf')
ZDT = REAL(DT, C_DOUBLE)
IRC = cmf_hip_physics_advance(HMAIN, ZDT, NT, C_LOC(GLB), GLBCAP)
CALL CHECK(IRC, 'physics_advance')
!! the log line of CALC_ADPSTP (physics:208): DT_DEF, DT_MIN, substep DT
IF( NT>=2 )THEN
  CALL CHECK(cmf_hip_get_dt_min(HMAIN, ZDTMIN), 'get_dt_min')
  WRITE(LOGNAM,'(A15,I4,3F10.2)') "ADPSTP: NT=", NT, DT, REAL(ZDTMIN,JPRB), DT*REAL(1./REAL(NT),JPRB)
ENDIF
DO IT=1, MIN(NT, GLBCAP)
  CALL WRITE_WATBAL(IT, DT*REAL(1./REAL(NT),JPRB), GLB(:,IT))
END DO
IF( LEVERY .OR. STATE_DUE() )THEN
  CALL CHECK(cmf_hip_get_state(HMAIN, STATE_PTRS(.FALSE.)), 'get_state')
ELSEIF( .NOT. LHIP_OUTDIAG )THEN
  CALL CHECK(cmf_hip_get_state(HMAIN, STATE_PTRS(.FALSE., .TRUE.)), 'get_state')
ENDIF
END SUBROUTINE CMF_PHYSICS_ADVANCE
!####################################################################
SUBROUTINE CMF_HIP_FORCING_PUT(PBUFF)
!! CMF_FORCING_PUT (src/cmf_ctrl_forcing_mod.F90:716-750) for a main program or
!! coupler that calls it in place of the reference's: with LINTERP, the input field
!! PBUFF(:,:,1) [NXIN,NYIN] goes to HBM and ROFF_INTERP (:757-790) runs there with
!! the input matrix of YOS_CMF_MAP (uploaded on the first call), so D2RUNOFF is
!! formed on the device, bit-identical to the host routine's, and never crosses
!! PCIe; D2RUNOFF in host memory is then not updated (the device holds it). The
!! host's CMF_FORCING_PUT runs instead (and the next advance uploads D2RUNOFF)
!! without LINTERP (nearest-point CONV_RESOL), with LROSPLIT or LWEVAP, and on a
!! single-precision host, whose interpolation rounds in float.
USE YOS_CMF_INPUT,        ONLY: NXIN, NYIN
USE CMF_CTRL_FORCING_MOD, ONLY: CMF_FORCING_PUT
IMPLICIT NONE
REAL(KIND=JPRB), INTENT(IN), TARGET :: PBUFF(:,:,:)
REAL(KIND=JPRB), ALLOCATABLE, TARGET, SAVE :: ZGRID(:,:)
!================================================
IF( .NOT. DEVICE_FORCING() )THEN
  CALL CMF_FORCING_PUT(PBUFF)
  LDEVROF = .FALSE.
  RETURN
ENDIF
CALL DEVICE_FORCING_INIT
IF( IS_CONTIGUOUS(PBUFF) .AND. SIZE(PBUFF,1) == NXIN .AND. SIZE(PBUFF,2) == NYIN )THEN
  CALL CHECK(cmf_hip_set_runoff_grid(HMAIN, PTR_D3(PBUFF)), 'set_runoff_grid')
ELSE
  IF( .NOT. ALLOCATED(ZGRID) ) ALLOCATE(ZGRID(NXIN, NYIN))
  ZGRID(:,:) = PBUFF(1:NXIN, 1:NYIN, 1)
  CALL CHECK(cmf_hip_set_runoff_grid(HMAIN, PTR_D2(ZGRID)), 'set_runoff_grid')
ENDIF
LDEVROF = .TRUE.
END SUBROUTINE CMF_HIP_FORCING_PUT
!####################################################################
SUBROUTINE CMF_HIP_FORCING_RECORDS(PBUFFS)
!! A run's input fields PBUFFS(NXIN,NYIN,NREC), in the order the main program will
!! select them, copied to HBM once (LINTERP on an FP64 host; the module stops
!! otherwise); CMF_HIP_FORCING_PUT_RECORD(IREC) then does CMF_HIP_FORCING_PUT's
!! work for field IREC with nothing crossing PCIe.
USE YOS_CMF_INPUT,        ONLY: NXIN, NYIN
IMPLICIT NONE
REAL(KIND=JPRB), INTENT(IN), TARGET :: PBUFFS(:,:,:)
!================================================
IF( .NOT. DEVICE_FORCING() ) CALL STOPMSG('resident forcing needs LINTERP on an FP64 host')
IF( SIZE(PBUFFS,1) /= NXIN .OR. SIZE(PBUFFS,2) /= NYIN .OR. .NOT. IS_CONTIGUOUS(PBUFFS) ) &
  & CALL STOPMSG('resident forcing: PBUFFS must be a contiguous (NXIN,NYIN,NREC) array')
CALL DEVICE_FORCING_INIT
CALL CHECK(cmf_hip_set_runoff_records(HMAIN, INT(SIZE(PBUFFS,3), C_INT32_T), PTR_D3(PBUFFS)), 'set_runoff_records')
END SUBROUTINE CMF_HIP_FORCING_RECORDS
!####################################################################
SUBROUTINE CMF_HIP_FORCING_PUT_RECORD(IREC)
!! D2RUNOFF from resident field IREC (1-based) of CMF_HIP_FORCING_RECORDS
IMPLICIT NONE
INTEGER(KIND=JPIM), INTENT(IN) :: IREC
!================================================
CALL CHECK(cmf_hip_use_runoff_record(HMAIN, INT(IREC-1, C_INT32_T)), 'use_runoff_record')
LDEVROF = .TRUE.
END SUBROUTINE CMF_HIP_FORCING_PUT_RECORD
!####################################################################
LOGICAL FUNCTION DEVICE_FORCING()
!! the options under which ROFF_INTERP runs on the device: LINTERP, FP64 host, no
!! second or third forcing field (LROSPLIT, LWEVAP)
USE YOS_CMF_INPUT,        ONLY: LROSPLIT, LWEVAP
USE CMF_CTRL_FORCING_MOD, ONLY: LINTERP
IMPLICIT NONE
DEVICE_FORCING = LINTERP .AND. .NOT. LROSPLIT .AND. .NOT. LWEVAP .AND. STORAGE_SIZE(0._JPRB) == 64
END FUNCTION DEVICE_FORCING
!####################################################################
SUBROUTINE DEVICE_FORCING_INIT
!! the handle (with the state after the host's initialisation) and, once, the
!! input matrix of YOS_CMF_MAP
USE YOS_CMF_INPUT,        ONLY: NXIN, NYIN, INPN, RMIS
USE YOS_CMF_MAP,          ONLY: INPX, INPY, INPA
USE YOS_CMF_PROG,         ONLY: D2ROFSUB
USE CMF_CTRL_FORCING_MOD, ONLY: DROFUNIT
IMPLICIT NONE
!================================================
IF( .NOT. C_ASSOCIATED(HMAIN) )THEN
  CALL OPEN_HANDLE(HMAIN, .TRUE.)
  CALL CHECK(cmf_hip_set_state(HMAIN, STATE_PTRS(.TRUE.)), 'set_state')
ENDIF
IF( .NOT. LINPMAT )THEN
  CALL CHECK(cmf_hip_set_inpmat(HMAIN, INT(INPN, C_INT32_T), INT(NXIN, C_INT32_T), INT(NYIN, C_INT32_T), &
    & PTR_I2(INPX), PTR_I2(INPY), PTR_D2(INPA), REAL(DROFUNIT, C_DOUBLE), REAL(RMIS, C_DOUBLE)), 'set_inpmat')
  D2ROFSUB(:,:) = 0._JPRB                    !! forcing_put:730-731 without LROSPLIT
  LINPMAT = .TRUE.
ENDIF
END SUBROUTINE DEVICE_FORCING_INIT
!####################################################################
LOGICAL FUNCTION STATE_DUE()
!! the host reads the state after this step: output time (drv_advance:121) or
!! restart time (CMF_RESTART_WRITE, restart:352-369)
USE YOS_CMF_INPUT,  ONLY: LOUTPUT, IFRQ_OUT
USE YOS_CMF_TIME,   ONLY: JHOUR, JMIN
USE CMF_CTRL_RESTART_MOD, ONLY: restart_is_write_time
IMPLICIT NONE
STATE_DUE = restart_is_write_time()
IF( LOUTPUT )THEN
  IF( MOD(JHOUR,IFRQ_OUT)==0 .AND. JMIN==0 ) STATE_DUE = .TRUE.
ENDIF
END FUNCTION STATE_DUE
!####################################################################
SUBROUTINE CMF_PHYSICS_FLDSTG
IMPLICIT NONE
TYPE(C_PTR) :: HT
!================================================
IF( C_ASSOCIATED(HMAIN) )THEN
  CALL CHECK(cmf_hip_set_state(HMAIN, STATE_PTRS(.TRUE.)), 'set_state')
  CALL CHECK(cmf_hip_physics_fldstg(HMAIN), 'physics_fldstg')
  CALL CHECK(cmf_hip_get_state(HMAIN, STATE_PTRS(.FALSE.)), 'get_state')
ELSE
  !! during initialisation (drv_control:243,276): a transient handle without dam tables
  HT = C_NULL_PTR
  CALL OPEN_HANDLE(HT, .FALSE.)
  CALL CHECK(cmf_hip_set_state(HT, STATE_PTRS(.TRUE.)), 'set_state')
  CALL CHECK(cmf_hip_physics_fldstg(HT), 'physics_fldstg')
  CALL CHECK(cmf_hip_get_state(HT, STATE_PTRS(.FALSE.)), 'get_state')
  CALL cmf_hip_destroy(HT)
ENDIF
END SUBROUTINE CMF_PHYSICS_FLDSTG
!####################################################################
SUBROUTINE OPEN_HANDLE(H, LWITHDAM)
USE YOS_CMF_INPUT,  ONLY: LADPSTP, LFPLAIN, LFLDOUT, LPTHOUT, LDAMOUT, LLEVEE, NLFP, &
                        & PGRV, PMANFLD, PDSTMTH, PCADP, PMINSLP, LSPAMAT, LKINE, LSLPMIX, &
                        & LSTG_ES, LGDWDLY, LROSPLIT, LWEVAP, LOUTINS, LSLOPEMOUTH, LTRACE, LSEDOUT
USE YOS_CMF_MAP,    ONLY: NSEQALL, NSEQRIV, NPTHOUT, NPTHLEV, I1NEXT, I2MASK, DFRCINC, &
                        & D2GRAREA, D2ELEVTN, D2NXTDST, D2RIVLEN, D2RIVWTH, D2RIVHGT, D2RIVMAN, &
                        & D2RIVSTOMAX, D2RIVELV, D2DWNELV, D2FLDHGT, D2FLDSTOMAX, D2FLDGRD, &
                        & PTH_UPST, PTH_DOWN, PTH_DST, PTH_ELV, PTH_WTH, PTH_MAN
USE CMF_CTRL_LEVEE_MOD,  ONLY: D2LEVHGT, D2LEVFRC, D2BASHGT, D2LEVDST, D2LEVBASSTO, D2LEVTOPSTO, D2LEVFILSTO
USE CMF_CTRL_DAMOUT_MOD, ONLY: NDAM, DAMSEQ, DAMSTAT, CONVOL, EMEVOL, ADJVOL, QA, QN, QF, I1DAM, LDAMH22, &
                        & NORVOL, FLDVOL, UPREAL
IMPLICIT NONE
TYPE(C_PTR), INTENT(INOUT) :: H
LOGICAL, INTENT(IN)        :: LWITHDAM
TYPE(CMF_DIMS_T)    :: D
TYPE(CMF_OPTIONS_T) :: O
TYPE(CMF_MAP_T)     :: M
CHARACTER(LEN=16)   :: CPREC
LOGICAL             :: LDAM
!================================================
IF( .NOT. LSPAMAT .OR. LSTG_ES .OR. LGDWDLY .OR. LROSPLIT .OR. LWEVAP &
  & .OR. LOUTINS .OR. (LKINE .AND. LSLOPEMOUTH) ) CALL STOPMSG('option not supported by the HIP physics')
!! the tracer (drv_advance:89-90, :110) and sediment (drv_advance:105) schemes read
!! P2RIVSTO/P2FLDSTO and the substep averages on the host every step, which this
!! module downloads only when the host's output or restart is due
IF( LTRACE .OR. LSEDOUT ) CALL STOPMSG('LTRACE/LSEDOUT are not supported by the HIP physics')
LDAM = LDAMOUT .AND. LWITHDAM
D%NSEQALL = NSEQALL;  D%NSEQRIV = NSEQRIV;  D%NLFP = NLFP
D%NPTHOUT = 0;  D%NPTHLEV = 0;  D%NDAM = 0
IF( LPTHOUT ) THEN; D%NPTHOUT = NPTHOUT; D%NPTHLEV = NPTHLEV; ENDIF
IF( LDAM ) D%NDAM = NDAM
O%LADPSTP = L2I(LADPSTP);  O%LFPLAIN = L2I(LFPLAIN);  O%LFLDOUT = L2I(LFLDOUT)
O%LPTHOUT = L2I(LPTHOUT);  O%LDAMOUT = L2I(LDAM);     O%LLEVEE = L2I(LLEVEE)
O%JPRB = INT(STORAGE_SIZE(0._JPRB) / 8, C_INT32_T)
O%PRECISION = 0
IF( O%JPRB == 4 ) O%PRECISION = 1
CALL GET_ENVIRONMENT_VARIABLE('CMF_HIP_PRECISION', CPREC)
IF( TRIM(CPREC)=='mixed' ) O%PRECISION = 1
IF( TRIM(CPREC)=='fp64' ) O%PRECISION = 0
CALL GET_ENVIRONMENT_VARIABLE('CMF_HIP_DOWNLOAD', CPREC)
IF( TRIM(CPREC)=='every' ) LEVERY = .TRUE.
O%PGRV = PGRV;  O%PMANFLD = PMANFLD;  O%PDSTMTH = PDSTMTH;  O%PCADP = PCADP;  O%PMINSLP = PMINSLP
O%LDAMH22 = 0
IF( LDAM ) O%LDAMH22 = L2I(LDAMH22)
CALL CHECK(cmf_hip_create(D, O, H), 'create')
!! LKINE: kinematic-wave outflow (physics:66-72)
IF( LKINE ) CALL CHECK(cmf_hip_set_kinematic(H, 1_C_INT32_T), 'set_kinematic')
!! LSLPMIX: slope-mixed outflow on the I2MASK the host set up (physics:68-70)
IF( LSLPMIX ) CALL CHECK(cmf_hip_set_slopemix(H, 1_C_INT32_T), 'set_slopemix')

M%I1NEXT = PTR_I(I1NEXT)
M%GRAREA = PTR_D2(D2GRAREA);  M%ELEVTN = PTR_D2(D2ELEVTN);  M%NXTDST = PTR_D2(D2NXTDST)
M%RIVLEN = PTR_D2(D2RIVLEN);  M%RIVWTH = PTR_D2(D2RIVWTH);  M%RIVHGT = PTR_D2(D2RIVHGT)
M%RIVMAN = PTR_D2(D2RIVMAN);  M%RIVSTOMAX = PTR_D2(D2RIVSTOMAX);  M%RIVELV = PTR_D2(D2RIVELV)
M%DWNELV = PTR_D2(D2DWNELV);  M%FLDHGT = PTR_D3(D2FLDHGT);  M%FLDSTOMAX = PTR_D3(D2FLDSTOMAX)
M%FLDGRD = PTR_D3(D2FLDGRD);  M%DFRCINC = DFRCINC;  M%I2MASK = PTR_I2(I2MASK)
M%PTH_UPST = C_NULL_PTR; M%PTH_DOWN = C_NULL_PTR; M%PTH_DST = C_NULL_PTR; M%PTH_ELV = C_NULL_PTR
M%PTH_WTH = C_NULL_PTR;  M%PTH_MAN = C_NULL_PTR
IF( LPTHOUT )THEN
  M%PTH_UPST = PTR_I(PTH_UPST);  M%PTH_DOWN = PTR_I(PTH_DOWN);  M%PTH_DST = PTR_D1(PTH_DST)
  M%PTH_ELV = PTR_D2(PTH_ELV);   M%PTH_WTH = PTR_D2(PTH_WTH);   M%PTH_MAN = PTR_D1(PTH_MAN)
ENDIF
M%LEVHGT = C_NULL_PTR; M%LEVFRC = C_NULL_PTR; M%BASHGT = C_NULL_PTR; M%LEVDST = C_NULL_PTR
M%LEVBASSTO = C_NULL_PTR; M%LEVTOPSTO = C_NULL_PTR; M%LEVFILSTO = C_NULL_PTR
IF( LLEVEE )THEN
  M%LEVHGT = PTR_D2(D2LEVHGT);  M%LEVFRC = PTR_D2(D2LEVFRC);  M%BASHGT = PTR_D2(D2BASHGT)
  M%LEVDST = PTR_D2(D2LEVDST);  M%LEVBASSTO = PTR_D2(D2LEVBASSTO)
  M%LEVTOPSTO = PTR_D2(D2LEVTOPSTO);  M%LEVFILSTO = PTR_D2(D2LEVFILSTO)
ENDIF
M%I1DAM = C_NULL_PTR; M%DAMSEQ = C_NULL_PTR; M%DAMSTAT = C_NULL_PTR; M%DAM_CONVOL = C_NULL_PTR
M%DAM_EMEVOL = C_NULL_PTR; M%DAM_ADJVOL = C_NULL_PTR; M%DAM_QA = C_NULL_PTR; M%DAM_QN = C_NULL_PTR
M%DAM_QF = C_NULL_PTR;  M%DAM_NORVOL = C_NULL_PTR;  M%DAM_R_VOLUPA = C_NULL_PTR
IF( LDAM )THEN
  M%I1DAM = PTR_I(I1DAM);  M%DAMSEQ = PTR_I(DAMSEQ);  M%DAMSTAT = PTR_I(DAMSTAT)
  M%DAM_CONVOL = PTR_D1(CONVOL);  M%DAM_EMEVOL = PTR_D1(EMEVOL);  M%DAM_ADJVOL = PTR_D1(ADJVOL)
  M%DAM_QA = PTR_D1(QA);  M%DAM_QN = PTR_D1(QN);  M%DAM_QF = PTR_D1(QF)
  IF( LDAMH22 )THEN
    !! damout:204 with the dam's own index; a dam not yet built (LDAMYBY, FldVol
    !! set to 0 after :204) is never operated, so its ratio is not read
    IF( .NOT. ALLOCATED(RVOLUPA) ) ALLOCATE(RVOLUPA(NDAM))
    RVOLUPA(:) = FLDVOL(:) * 1.E-6 / UPREAL(:)
    M%DAM_NORVOL = PTR_D1(NORVOL);  M%DAM_R_VOLUPA = PTR_D1(RVOLUPA)
  ENDIF
ENDIF
CALL CHECK(cmf_hip_set_map(H, M), 'set_map')
END SUBROUTINE OPEN_HANDLE
!####################################################################
FUNCTION STATE_PTRS(LUP, LAVGONLY) RESULT(S)
!! Pointers to the host state: LUP=.TRUE. for upload (prognostic state and the
!! stage diagnostics), .FALSE. for download (everything the host may read);
!! LAVGONLY: the substep averages and maxima only.
USE YOS_CMF_INPUT,  ONLY: LPTHOUT, LDAMOUT, LLEVEE
USE YOS_CMF_PROG,   ONLY: P2RIVSTO, P2FLDSTO, P2LEVSTO, P2DAMSTO, P2DAMINF, D2RIVOUT, D2FLDOUT, &
                        & D2RIVOUT_PRE, D2FLDOUT_PRE, D2RIVDPH_PRE, D2FLDSTO_PRE, D1PTHFLW, D1PTHFLW_PRE
USE YOS_CMF_DIAG,   ONLY: D2RIVDPH, D2FLDDPH, D2FLDFRC, D2FLDARE, D2SFCELV, D2STORGE, D2OUTFLW, &
                        & D2RIVVEL, D2RIVINF, D2FLDINF, D2PTHOUT, D2LEVDPH, D1PTHFLWSUM, &
                        & D2RIVOUT_aAVG, D2FLDOUT_aAVG, D2OUTFLW_aAVG, D2RIVVEL_aAVG, D2PTHOUT_aAVG, &
                        & D2RUNOFF_aAVG, D2OUTFLW_aMAX, D2RIVDPH_aMAX, D2STORGE_aMAX, D2DAMINF_aAVG, &
                        & D1PTHFLW_aAVG, D1PTHFLWSUM_aAVG
IMPLICIT NONE
LOGICAL, INTENT(IN) :: LUP
LOGICAL, INTENT(IN), OPTIONAL :: LAVGONLY
TYPE(CMF_STATE_T)   :: S
LOGICAL             :: LAVG
!================================================
LAVG = .FALSE.
IF( PRESENT(LAVGONLY) ) LAVG = LAVGONLY
S = CMF_STATE_T(C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, &
  & C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, &
  & C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, &
  & C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, &
  & C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR)
IF( LAVG ) GOTO 100
S%RIVSTO = PTR_P2(P2RIVSTO);  S%FLDSTO = PTR_P2(P2FLDSTO)
S%RIVOUT = PTR_D2(D2RIVOUT);  S%FLDOUT = PTR_D2(D2FLDOUT)
S%RIVOUT_PRE = PTR_D2(D2RIVOUT_PRE);  S%FLDOUT_PRE = PTR_D2(D2FLDOUT_PRE)
S%RIVDPH_PRE = PTR_D2(D2RIVDPH_PRE);  S%FLDSTO_PRE = PTR_D2(D2FLDSTO_PRE)
S%RIVDPH = PTR_D2(D2RIVDPH);  S%FLDDPH = PTR_D2(D2FLDDPH);  S%FLDFRC = PTR_D2(D2FLDFRC)
S%FLDARE = PTR_D2(D2FLDARE);  S%SFCELV = PTR_D2(D2SFCELV);  S%STORGE = PTR_D2(D2STORGE)
IF( LLEVEE )THEN
  S%LEVSTO = PTR_P2(P2LEVSTO);  S%LEVDPH = PTR_D2(D2LEVDPH)
ENDIF
IF( LDAMOUT .AND. ALLOCATED(P2DAMSTO) )THEN
  S%DAMSTO = PTR_P2(P2DAMSTO);  S%DAMINF = PTR_P2(P2DAMINF)
ENDIF
IF( LPTHOUT )THEN
  S%PTHFLW = PTR_D2(D1PTHFLW);  S%PTHFLW_PRE = PTR_D2(D1PTHFLW_PRE);  S%PTHFLWSUM = PTR_D1(D1PTHFLWSUM)
ENDIF
IF( .NOT. LUP )THEN
  S%OUTFLW = PTR_D2(D2OUTFLW);  S%RIVVEL = PTR_D2(D2RIVVEL);  S%RIVINF = PTR_D2(D2RIVINF)
  S%FLDINF = PTR_D2(D2FLDINF);  S%PTHOUT = PTR_D2(D2PTHOUT)
ENDIF
100 CONTINUE
IF( .NOT. LUP .AND. .NOT. LHIP_OUTDIAG )THEN
  S%RIVOUT_AVG = PTR_D2(D2RIVOUT_aAVG);  S%FLDOUT_AVG = PTR_D2(D2FLDOUT_aAVG)
  S%OUTFLW_AVG = PTR_D2(D2OUTFLW_aAVG);  S%RIVVEL_AVG = PTR_D2(D2RIVVEL_aAVG)
  S%PTHOUT_AVG = PTR_D2(D2PTHOUT_aAVG);  S%RUNOFF_AVG = PTR_D2(D2RUNOFF_aAVG)
  S%OUTFLW_MAX = PTR_D2(D2OUTFLW_aMAX);  S%RIVDPH_MAX = PTR_D2(D2RIVDPH_aMAX)
  S%STORGE_MAX = PTR_D2(D2STORGE_aMAX)
  IF( LDAMOUT .AND. ALLOCATED(D2DAMINF_aAVG) ) S%DAMINF_AVG = PTR_D2(D2DAMINF_aAVG)
  IF( LPTHOUT )THEN
    S%PTHFLW_AVG = PTR_D2(D1PTHFLW_aAVG);  S%PTHFLWSUM_AVG = PTR_D1(D1PTHFLWSUM_aAVG)
  ENDIF
ENDIF
END FUNCTION STATE_PTRS
!####################################################################
SUBROUTINE WRITE_WATBAL(IT, ZDT, G)
!! budget log line of CALC_WATBAL (src/cmf_ctrl_physics_mod.F90:216-248)
USE YOS_CMF_INPUT,  ONLY: LOGNAM
USE YOS_CMF_TIME,   ONLY: KMIN
USE CMF_UTILS_MOD,  ONLY: MIN2DATE, SPLITDATE, SPLITHOUR
IMPLICIT NONE
INTEGER(C_INT32_T), INTENT(IN) :: IT
REAL(KIND=JPRB), INTENT(IN)    :: ZDT
REAL(C_DOUBLE), INTENT(IN)     :: G(NGLB)
INTEGER(KIND=JPIM) :: PKMIN, PYEAR, PMON, PDAY, PHOUR, PMIN, PYYYYMMDD, PHHMM
REAL(KIND=JPRD)    :: DERROR, DERROR2
REAL(KIND=JPRD), PARAMETER :: DORD = 1.D-9
PKMIN = INT( KMIN + IT*ZDT/60._JPRB )
CALL MIN2DATE(PKMIN, PYYYYMMDD, PHHMM)
CALL SPLITDATE(PYYYYMMDD, PYEAR, PMON, PDAY)
CALL SPLITHOUR(PHHMM, PHOUR, PMIN)
DERROR  = - (G(1) - G(2) + G(4) - G(5))
DERROR2 = - (G(6) - G(7))
WRITE(LOGNAM,'(I4.4,4(A1,I2.2),I6,a6,3F12.3,G12.3,2x,2F12.3,a6, 2F12.3,G12.3,3F12.3)') &
  PYEAR, '/', PMON, '/', PDAY, '_', PHOUR, ':', PMIN, IT, ' flx: ', &
  G(1)*DORD, G(2)*DORD, G(3)*DORD, DERROR*DORD, G(4)*DORD, G(5)*DORD, ' stg: ', &
  G(6)*DORD, G(7)*DORD, DERROR2*DORD, G(8)*DORD, G(9)*DORD, G(10)*DORD
END SUBROUTINE WRITE_WATBAL
!####################################################################
SUBROUTINE CHECK(IRC, CWHAT)
USE YOS_CMF_INPUT,  ONLY: LOGNAM
IMPLICIT NONE
INTEGER(C_INT), INTENT(IN)   :: IRC
CHARACTER(LEN=*), INTENT(IN) :: CWHAT
CHARACTER(KIND=C_CHAR), POINTER :: CMSG(:)
CHARACTER(LEN=512) :: CBUF
INTEGER :: I
IF( IRC == 0 ) RETURN
CBUF = ''
CALL C_F_POINTER(cmf_hip_last_error(), CMSG, [512])
DO I=1, 512
  IF( CMSG(I) == C_NULL_CHAR ) EXIT
  CBUF(I:I) = CMSG(I)
END DO
WRITE(LOGNAM,*) 'CMF_CTRL_PHYSICS_MOD (HIP): ', CWHAT, ' failed: ', TRIM(CBUF)
WRITE(*,*) 'CMF_CTRL_PHYSICS_MOD (HIP): ', CWHAT, ' failed: ', TRIM(CBUF)
STOP 9
END SUBROUTINE CHECK

TYPE(C_PTR) FUNCTION CMF_HIP_HANDLE()
!! the handle of the model state in HBM (C_NULL_PTR before the first advance)
CMF_HIP_HANDLE = HMAIN
END FUNCTION CMF_HIP_HANDLE

SUBROUTINE CMF_HIP_CHECK(IRC, CWHAT)
INTEGER(C_INT), INTENT(IN)   :: IRC
CHARACTER(LEN=*), INTENT(IN) :: CWHAT
CALL CHECK(IRC, CWHAT)
END SUBROUTINE CMF_HIP_CHECK

SUBROUTINE STOPMSG(CMSG)
USE YOS_CMF_INPUT,  ONLY: LOGNAM
IMPLICIT NONE
CHARACTER(LEN=*), INTENT(IN) :: CMSG
WRITE(LOGNAM,*) 'CMF_CTRL_PHYSICS_MOD (HIP): ', CMSG
WRITE(*,*) 'CMF_CTRL_PHYSICS_MOD (HIP): ', CMSG
STOP 9
END SUBROUTINE STOPMSG

INTEGER(C_INT32_T) FUNCTION L2I(L)
LOGICAL, INTENT(IN) :: L
L2I = 0
IF( L ) L2I = 1
END FUNCTION L2I

!! Addresses of contiguous allocatable arrays (the map arrays are not TARGET in
!! the reference; an assumed-size TARGET dummy gives their address without copy).
TYPE(C_PTR) FUNCTION PTR_D(A)
REAL(KIND=JPRB), TARGET, INTENT(IN) :: A(*)
PTR_D = C_LOC(A(1))
END FUNCTION PTR_D
TYPE(C_PTR) FUNCTION PTR_D1(A)
REAL(KIND=JPRB), TARGET, INTENT(IN) :: A(:)
PTR_D1 = C_LOC(A(1))
END FUNCTION PTR_D1
TYPE(C_PTR) FUNCTION PTR_D2(A)
REAL(KIND=JPRB), TARGET, INTENT(IN) :: A(:,:)
PTR_D2 = C_LOC(A(1,1))
END FUNCTION PTR_D2
TYPE(C_PTR) FUNCTION PTR_D3(A)
REAL(KIND=JPRB), TARGET, INTENT(IN) :: A(:,:,:)
PTR_D3 = C_LOC(A(1,1,1))
END FUNCTION PTR_D3
TYPE(C_PTR) FUNCTION PTR_P2(A)
REAL(KIND=JPRD), TARGET, INTENT(IN) :: A(:,:)
PTR_P2 = C_LOC(A(1,1))
END FUNCTION PTR_P2
TYPE(C_PTR) FUNCTION PTR_I(A)
INTEGER(KIND=JPIM), TARGET, INTENT(IN) :: A(:)
PTR_I = C_LOC(A(1))
END FUNCTION PTR_I
TYPE(C_PTR) FUNCTION PTR_I2(A)
INTEGER(KIND=JPIM), TARGET, INTENT(IN) :: A(:,:)
PTR_I2 = C_LOC(A(1,1))
END FUNCTION PTR_I2

END MODULE CMF_CTRL_PHYSICS_MOD
