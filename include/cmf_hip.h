/*
 * cmf_hip.h -- C-ABI of the MI355X-native CaMa-Flood per-substep physics.
 *
 * Drop-in boundary: these entry points replace the two public subroutines of the
 * reference module CMF_CTRL_PHYSICS_MOD (src/cmf_ctrl_physics_mod.F90):
 *   CMF_PHYSICS_ADVANCE  (:21-155)   -> cmf_hip_physics_advance
 *   CMF_PHYSICS_FLDSTG   (:288-305)  -> cmf_hip_physics_fldstg
 * plus the data movement a device-resident implementation needs around them
 * (maps once, prognostic state when the host changes it, forcing per input step,
 * diagnostics when the host consumes them). The Fortran binding is
 * cama-flood_v4_amd/fortran/cmf_ctrl_physics_mod.F90.
 *
 * Conventions
 *  - Arrays are the host's contiguous Fortran allocatables in NSEQ vector order.
 *    Index arrays keep the reference's 1-based values (I1NEXT, PTH_UPST, ...).
 *    Real arrays are JPRB = double (the reference's FP64 build) unless noted.
 *    2-D tables keep Fortran column order: FLDSTOMAX(NSEQMAX,1,NLFP) is NLFP
 *    planes of NSEQALL; PTH_ELV(NPTHOUT,NPTHLEV) is NPTHLEV planes of NPTHOUT.
 *  - Every function returns 0 on success, a negative CMF_E* code on failure;
 *    cmf_hip_last_error() returns the message of the calling thread's last failure.
 *  - Calls on one handle must not overlap; one HIP stream per handle, default
 *    priority. No pointer passed in is retained after the call returns.
 */
#ifndef CMF_HIP_H
#define CMF_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define CMF_OK          0
#define CMF_EINVAL     -1   /* bad argument / inconsistent sizes */
#define CMF_EHIP       -2   /* HIP runtime error */
#define CMF_ESTATE     -3   /* call out of order (e.g. advance before set_map) */
#define CMF_EMAP       -4   /* map tables fail validation */
#define CMF_ENOMEM     -5
#define CMF_EIO        -6   /* file could not be opened, read or written */

/* Device arithmetic of the flux/stage variables (JPRB on the device). Storages,
 * the storage update, the stage inversion and the budget sums are FP64 in both. */
#define CMF_PREC_FP64   0   /* reference FP64 build semantics */
#define CMF_PREC_MIXED  1   /* FP32 flux arithmetic and flux/stage arrays, FP64 storage */

typedef struct cmf_hip_ctx* cmf_handle;

/* Sizes: NSEQALL, NSEQRIV, NLFP, NPTHOUT, NPTHLEV, NDAM (src/yos_cmf_map.F90,
 * src/cmf_ctrl_damout_mod.F90:41). npthout/npthlev/ndam may be 0 when unused. */
typedef struct cmf_dims {
    int32_t nseqall, nseqriv, nlfp, npthout, npthlev, ndam;
} cmf_dims;

/* Namelist switches and parameters read by the physics
 * (src/cmf_ctrl_nmlist_mod.F90:77-115, :227-232). LSPAMAT is implied (gather
 * inflow). LKINE, LSLPMIX and LUPSINF are switched on by their own entry points
 * (cmf_hip_set_kinematic, cmf_hip_set_slopemix, cmf_hip_set_upsinf); LSTG_ES, LGDWDLY, LROSPLIT,
 * LWEVAP and LOUTINS must be false. */
typedef struct cmf_options {
    int32_t ladpstp, lfplain, lfldout, lpthout, ldamout, llevee;
    int32_t precision;                 /* CMF_PREC_* */
    double pgrv, pmanfld, pdstmth, pcadp, pminslp;
    /* bytes of the host's JPRB (src/parkind1.F90:34-38): 8 (or 0) for the FP64
     * build, 4 for the reference's default -DSinglePrec_CMF build. With 4 every
     * JPRB array passed to or filled by the handle is float (c_float): the map
     * tables of cmf_map, the D2 and D1 arrays of cmf_state, cmf_output, runoff,
     * D2DWNELV, D2UPSINF, D2MEANSL/D2SEALEV, INPA and the runoff grid, although
     * the pointer members and parameters are declared double* (cast float* to
     * them); the P2* storages, the budget sums, DT, DT_MIN and the ensemble
     * factors stay double. The device arithmetic is chosen by `precision`; the
     * float map tables are checked against the FP64 recurrence to 1e-5. */
    int32_t jprb;
    /* LDAMH22 (src/cmf_ctrl_damout_mod.F90:354-378): the Hanazaki-2022 reservoir
     * rule instead of Yamazaki & Funato; needs dam_norvol and dam_r_volupa */
    int32_t ldamh22;
} cmf_options;

/* Static map (YOS_CMF_MAP, levee and dam module tables). Required: i1next,
 * grarea, elevtn, nxtdst, rivlen, rivwth, rivhgt, rivman, rivstomax, rivelv,
 * dwnelv, fldhgt, fldstomax, fldgrd, i2mask. Paths when lpthout, levee tables
 * when llevee, dam tables when ldamout. */
typedef struct cmf_map {
    const int32_t* i1next;                           /* I1NEXT, maps:420 */
    const double *grarea, *elevtn, *nxtdst, *rivlen, *rivwth, *rivhgt, *rivman;
    const double *rivstomax, *rivelv, *dwnelv;      /* maps:716-736 */
    const double *fldhgt, *fldstomax, *fldgrd;       /* NLFP planes */
    double dfrcinc;                                  /* DFRCINC */
    const int32_t* i2mask;
    /* bifurcation (maps:531-639) */
    const int32_t *pth_upst, *pth_down;
    const double *pth_dst, *pth_elv, *pth_wth, *pth_man;
    /* levee (levee:93-264) */
    const double *levhgt, *levfrc, *bashgt, *levdst, *levbassto, *levtopsto, *levfilsto;
    /* dams (damout:123-308): per dam, 1-based DamSeq, DamStat */
    const int32_t *i1dam, *damseq, *damstat;
    const double *dam_convol, *dam_emevol, *dam_adjvol, *dam_qa, *dam_qn, *dam_qf;
    /* with ldamh22 (damout:202-205): NorVol, and R_VolUpa = FldVol*1.E-6/upreal
     * of each dam at its own index (the reference's loop stores every dam's value
     * at index NDAM, so its other dams read an element never set); dam_adjvol and
     * dam_qa may be NULL */
    const double *dam_norvol, *dam_r_volupa;
} cmf_map;

/* Prognostic and diagnostic arrays (YOS_CMF_PROG, YOS_CMF_DIAG). Any pointer may
 * be NULL: that array is skipped by set/get. Storages (P2*) are FP64. */
typedef struct cmf_state {
    double *rivsto, *fldsto, *levsto, *damsto, *daminf;          /* P2* */
    double *rivout, *fldout, *rivout_pre, *fldout_pre;           /* D2*OUT(_PRE) */
    double *rivdph_pre, *fldsto_pre;
    double *pthflw, *pthflw_pre;                                 /* NPTHLEV planes */
    double *rivdph, *flddph, *fldfrc, *fldare, *sfcelv, *storge; /* D2 diag */
    double *outflw, *rivvel, *rivinf, *fldinf, *pthout, *levdph;
    double *pthflwsum;
    /* substep averages/maxima of the last advance (CMF_DIAG_*_ADPSTP) */
    double *rivout_avg, *fldout_avg, *outflw_avg, *rivvel_avg, *pthout_avg;
    double *runoff_avg, *outflw_max, *rivdph_max, *storge_max, *daminf_avg;
    double *pthflw_avg, *pthflwsum_avg;
} cmf_state;

/* Global water-budget sums of one substep, the P0GLB* of YOS_CMF_DIAG
 * (src/yos_cmf_diag.F90:109-125) in this order. */
#define CMF_NGLB 11
enum { CMF_GLB_STOPRE, CMF_GLB_STONXT, CMF_GLB_STONEW, CMF_GLB_RIVINF, CMF_GLB_RIVOUT,
       CMF_GLB_STOPRE2, CMF_GLB_STONEW2, CMF_GLB_RIVSTO, CMF_GLB_FLDSTO, CMF_GLB_FLDARE,
       CMF_GLB_LEVSTO };

/* Create a handle on the current HIP device. No reference counterpart (the
 * reference keeps this state in module globals). */
int cmf_hip_create(const cmf_dims* dims, const cmf_options* opts, cmf_handle* out);

/* Batched ensembles (no reference counterpart: the reference runs one model per
 * process). ndom regional domains are concatenated in the map passed to
 * cmf_hip_set_map (domain d owns the next dom_cells[d] NSEQ cells; I1NEXT, paths
 * and dams stay inside a domain) and every domain runs `members` realisations
 * (a power of two, 1..64). Instance k = d*members + m has its own adaptive time
 * step. Per-cell arrays of cmf_state and cmf_hip_set_runoff then hold `members`
 * planes of NSEQALL values (member-major); path arrays hold `members` planes of
 * NPTHLEV*NPTHOUT (or NPTHOUT). dims holds the totals over all domains.
 * cmf_hip_create(d, o, h) is cmf_hip_create_ensemble(d, o, NULL, h): one domain,
 * one member. */
typedef struct cmf_ensemble {
    int32_t ndom;
    const int32_t* dom_cells;          /* [ndom] */
    int32_t members;
} cmf_ensemble;
int cmf_hip_create_ensemble(const cmf_dims* dims, const cmf_options* opts, const cmf_ensemble* ens,
                            cmf_handle* out);
void cmf_hip_destroy(cmf_handle h);
const char* cmf_hip_last_error(void);

/* Upload and validate the static map; builds the upstream CSR from I1NEXT in
 * ascending source order (the I1UPST fill order, maps:518-522) and the per-cell
 * path lists (I1P_OUT/I1P_INF, maps:607-636). Replaces the reads of the map
 * module globals by CMF_PHYSICS_*. */
int cmf_hip_set_map(cmf_handle h, const cmf_map* map);

/* Host -> device copy of the non-NULL arrays of st (after CMF_PROG_INIT,
 * CMF_RESTART_INIT, CMF_DAMOUT_INIT, or any host-side change of state). */
int cmf_hip_set_state(cmf_handle h, const cmf_state* st);
/* Ensembles: as cmf_hip_set_state, but each array holds one plane (NSEQALL cell
 * values, NPTHLEV*NPTHOUT or NPTHOUT path values) that every member of every
 * domain receives (a common initial state). No reference counterpart. */
int cmf_hip_set_state_shared(cmf_handle h, const cmf_state* st);
/* Device -> host copy into the non-NULL arrays of st. */
int cmf_hip_get_state(cmf_handle h, cmf_state* st);

/* Runoff forcing D2RUNOFF [m3/s] (written by CMF_FORCING_PUT). */
int cmf_hip_set_runoff(cmf_handle h, const double* runoff);
/* Ensemble forcing: D2RUNOFF of instance k = base[cell] * factor[k] (base holds
 * NSEQALL values, factor ndom*members). No reference counterpart. */
int cmf_hip_set_runoff_scaled(cmf_handle h, const double* base, const double* factor);

/* Runoff interpolation on the device: ROFF_INTERP of CMF_FORCING_PUT with LINTERP
 * (src/cmf_ctrl_forcing_mod.F90:757-790). cmf_hip_set_inpmat uploads the input
 * matrix INPX, INPY (1-based input-grid indices, <= 0 = no entry) and INPA [m2]
 * of YOS_CMF_MAP, each [inpn][NSEQALL] (Fortran INPX(NSEQMAX,INPN) as it lies in
 * memory), with DROFUNIT and RMIS (the float RMIS widened to double). Entries with
 * IXIN > 0 must have IYIN >= 1. cmf_hip_set_runoff_grid uploads one runoff field
 * PBUFFIN(NXIN,NYIN) (x fastest) and interpolates it into D2RUNOFF;
 * cmf_hip_interp_runoff does the same from a field already in device memory (the
 * forcing kept resident). The sum over a cell's entries runs in entry order with
 * the reference's operation order, so D2RUNOFF is bit-identical to the host
 * routine's. Every member of an ensemble receives its cell's value. */
int cmf_hip_set_inpmat(cmf_handle h, int32_t inpn, int32_t nxin, int32_t nyin, const int32_t* inpx,
                       const int32_t* inpy, const double* inpa, double drofunit, double rmis);
int cmf_hip_set_runoff_grid(cmf_handle h, const double* grid);
int cmf_hip_interp_runoff(cmf_handle h, const double* d_grid);

/* A run's forcing resident in HBM. No reference counterpart (the reference reads
 * one field per forcing interval, CMF_FORCING_GET): cmf_hip_set_runoff_records
 * copies nrec input fields PBUFFIN(NXIN,NYIN) (x fastest, one after another, JPRB
 * per the handle) to the device once, replacing any earlier set;
 * cmf_hip_use_runoff_record interpolates record k (0-based) into D2RUNOFF as
 * cmf_hip_set_runoff_grid does, with no host copy. Needs cmf_hip_set_inpmat first.
 * A year of hourly 0.25-degree fields is 73 GB in FP64. */
int cmf_hip_set_runoff_records(cmf_handle h, int32_t nrec, const double* grids);
int cmf_hip_use_runoff_record(cmf_handle h, int32_t k);

/* Output-period diagnostics on the device: CMF_DIAG_RESET_OUTPUT,
 * CMF_DIAG_AVEMAX_OUTPUT and CMF_DIAG_GETAVE_OUTPUT (src/cmf_calc_diag_mod.F90:
 * 222-356), called where the reference's CMF_DRV_ADVANCE calls them
 * (src/cmf_drv_advance_mod.F90:114,123,143): avemax after every outer step with
 * its length DT, getave at output time (returns NADD_out), reset after the
 * output is written. The accumulators (D2*_oAVG, D2*_oMAX, D1PTHFLW_oAVG) stay
 * in device memory, FP64 in both precisions; cmf_hip_get_output copies the
 * non-NULL ones out (NSEQ order; members planes for ensembles), so the per-step
 * download of the substep averages is not needed. GDWRTN, ROFSUB and WEVAPEX
 * averages are zero with the options this library supports. */
typedef struct {
    double *rivout, *fldout, *outflw, *rivvel, *pthout, *runoff, *daminf;   /* D2*_oAVG */
    double *outflw_max, *rivdph_max, *storge_max;                           /* D2*_oMAX */
    double* pthflw;                                                          /* D1PTHFLW_oAVG */
} cmf_output;
int cmf_hip_diag_reset_output(cmf_handle h);
int cmf_hip_diag_avemax_output(cmf_handle h, double dt);
int cmf_hip_diag_getave_output(cmf_handle h, double* nadd_out);
int cmf_hip_get_output(cmf_handle h, const cmf_output* out);

/* Sea-level boundary (src/cmf_ctrl_boundary_mod.F90:267-305). cmf_hip_set_dwnelv
 * uploads D2DWNELV as the host's CMF_BOUNDARY_UPDATE left it (call before every
 * advance when LSEALEV is on: the map upload only carries the initial value).
 * cmf_hip_update_boundary forms D2DWNELV = D2ELEVTN (+ D2MEANSL) (+ D2SEALEV) on
 * the device, NULL switching a term off (LMEANSL / LSEALEV false). */
int cmf_hip_set_dwnelv(cmf_handle h, const double* dwnelv);
/* LUPSINF: upstream inflow D2UPSINF [m3/s] (as CMF_UPSINF_UPDATE leaves it, per
 * cell, members planes for ensembles), added to the river storage in the storage
 * update (src/cmf_calc_stonxt_mod.F90:81). The first call switches it on for
 * the handle; call again whenever D2UPSINF changes. */
int cmf_hip_set_upsinf(cmf_handle h, const double* upsinf);
int cmf_hip_update_boundary(cmf_handle h, const double* meansl, const double* sealev);
/* LKINE (namelist NRUNVER): 1 replaces the local inertial river and floodplain
 * outflow and its reverse-flow limiter by the kinematic wave of
 * CMF_CALC_OUTFLW_KINE (src/cmf_opt_outflw_mod.F90:27-127, selected at
 * src/cmf_ctrl_physics_mod.F90:66-72): bed slope ELEVTN(i)-ELEVTN(next) over
 * NXTDST, at least PMINSLP (mouths PMINSLP: LSLOPEMOUTH off), outflow capped by
 * storage / DT. 0 (the default) restores the local inertial scheme. Dams, paths,
 * the limiter of the inflow and the storage update are unchanged. */
int cmf_hip_set_kinematic(cmf_handle h, int32_t lkine);
/* LSLPMIX (namelist NRUNVER): 1 selects CMF_CALC_OUTFLW_KINEMIX
 * (src/cmf_opt_outflw_mod.F90:118-300, selected at src/cmf_ctrl_physics_mod.F90:68-70
 * when LKINE is off): the kinematic wave of cmf_hip_set_kinematic on the cells
 * with I2MASK != 0 (the slope mask of SET_SLOPEMIX, src/cmf_ctrl_maps_mod.F90:1005-1040,
 * passed in cmf_map.i2mask, plus the dam codes), that routine's local inertial
 * variant elsewhere. LKINE, when also on, takes precedence as in the reference. */
int cmf_hip_set_slopemix(cmf_handle h, int32_t lslpmix);

/* One output record of variable `var` appended to `file`, as CMF_OUTPUT_WRITE
 * writes it in binary mode (src/cmf_ctrl_output_mod.F90:474-664): float32,
 * either an NX*NY map with RMIS off the river cells (loutvec = 0) or the NSEQALL
 * vector (LOUTVEC); pthflw as an NPTHOUT*NPTHLEV record. Names and sources as
 * the reference's: rivout, fldout, outflw/totout, rivvel, pthout, pthflw,
 * runoff/rofsfc, daminf (output-period averages, so after
 * cmf_hip_diag_getave_output), maxflw, maxdph, maxsto (output-period maxima),
 * rivsto, fldsto, rivdph, flddph, fldfrc, fldare, sfcelv, storge/totsto, damsto,
 * levsto, levdph (current state). Variables of a switched-off option are skipped
 * (return 0, nothing written), as the reference skips them; other names are
 * CMF_EINVAL. Single-model handles. */
int cmf_hip_output_write(cmf_handle h, const char* var, const char* file, int32_t nx, int32_t ny,
                         const int32_t* seqx, const int32_t* seqy, int32_t loutvec, double rmis);

/* Binary restart straight from / into device state: WRTE_REST_BIN and
 * READ_REST_BIN of CMF_RESTART_WRITE / CMF_RESTART_INIT
 * (src/cmf_ctrl_restart_mod.F90:389-482, 484-517 and its reader), single-model
 * handles. Records of NX*NY values (float64 with lrestdbl, else float32), x
 * fastest, cells placed by seqx/seqy (1-based I1SEQX/I1SEQY, NSEQ order) and the
 * rest of the grid DMIS (float64) or RMIS (float32): P2RIVSTO, P2FLDSTO, then,
 * unless lstoonly, D2RIVOUT_PRE, D2FLDOUT_PRE, D2RIVDPH_PRE, D2FLDSTO_PRE, then
 * P2DAMSTO (LDAMOUT) and P2LEVSTO (LLEVEE); with LPTHOUT, D1PTHFLW_PRE
 * (NPTHOUT x NPTHLEV) in the file `file`.pth. The file name is the caller's
 * (the reference builds CRESTDIR//CVNREST//YYYYMMDDHH//CSUFBIN). The reader is
 * CMF_RESTART_INIT (restart:95-133): the state arrays the file does not carry are
 * zero (D2RIVOUT, D2FLDOUT; with lstoonly also the four t-1 records and
 * D1PTHFLW_PRE, the .pth file not read), the records are loaded into the device
 * state, and with lstoonly D2FLDSTO_PRE = P2FLDSTO. For a storage-only restart
 * the caller then runs cmf_hip_physics_fldstg and cmf_hip_calc_outpre, as
 * CMF_DRV_INIT does (src/cmf_drv_control_mod.F90:273-278). */
int cmf_hip_restart_write(cmf_handle h, const char* file, int32_t nx, int32_t ny, const int32_t* seqx,
                          const int32_t* seqy, int32_t lrestdbl, int32_t lstoonly, double dmis, double rmis);
int cmf_hip_restart_read(cmf_handle h, const char* file, int32_t nx, int32_t ny, const int32_t* seqx,
                         const int32_t* seqy, int32_t lrestdbl, int32_t lstoonly);

/* CMF_CALC_OUTPRE (src/cmf_opt_outflw_mod.F90:308-421): the previous-step flows
 * D2RIVOUT_PRE, D2FLDOUT_PRE (0 without LFLDOUT) and D1PTHFLW_PRE from the
 * current stage (diffusive-wave discharge of the water surface), with
 * D2SFCELV and D2RIVDPH_PRE = D2RIVDPH; D2RIVOUT is left as it is. Call after
 * cmf_hip_physics_fldstg. */
int cmf_hip_calc_outpre(cmf_handle h);

/* CMF_PHYSICS_FLDSTG: storage -> stage on the device state. */
int cmf_hip_physics_fldstg(cmf_handle h);

/* CMF_PHYSICS_ADVANCE for one outer step of length *dt (DT). On return *dt is
 * restored (physics:145), *nt_out = NT. If glb != NULL it receives the budget sums
 * of up to glb_cap substeps (CMF_NGLB values each, substep-major). */
int cmf_hip_physics_advance(cmf_handle h, double* dt, int32_t* nt_out,
                            double* glb, int32_t glb_cap);

/* Allocation-free variant of cmf_hip_physics_advance. No counterpart in the
 * reference: it is for callers that cannot allocate device memory while they
 * step the model (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). It calls no hipMalloc/hipFree and no allocating
 * library routine: the budget partial sums, the budget rows and the CFL partial
 * minima live in a fixed 6.65 MB device array of the library (one copy per GPU,
 * held until the process ends; its address is looked up when a handle is
 * created). Same semantics and results as cmf_hip_physics_advance (the
 * substeps always run as three phase kernels; no graph replay, no event
 * timing). Limits: single-model handles, NSEQALL <= 16,777,216, at most 65,536
 * update tiles (NSEQALL / bs_update) and NT <= 4096 (a larger NT returns
 * CMF_EINVAL after the outer-step stage, which leaves the storages as they
 * were). Every call on one GPU shares that GPU's array and overwrites what it
 * needs before reading it: calls that overlap on one GPU (different handles on
 * different host threads or streams) race and must not be issued. */
int cmf_hip_physics_advance_noalloc(cmf_handle h, double* dt, int32_t* nt_out, double* glb, int32_t glb_cap);

/* Timing support: run the outer step with at most nt_cap substeps (0 = all).
 * Same as cmf_hip_physics_advance otherwise. No reference counterpart. */
int cmf_hip_physics_advance_capped(cmf_handle h, double* dt, int32_t nt_cap,
                                   int32_t* nt_out, double* glb, int32_t glb_cap);

/* One outer step of length dt of every instance of an ensemble handle (also
 * accepted for a single model). nt_out[k] / dt_out[k] (either may be NULL)
 * receive NT and DT of instance k; nt_cap > 0 caps the substeps (timing). No
 * budget sums are kept for more than one instance. No reference counterpart. */
int cmf_hip_ensemble_advance(cmf_handle h, double dt, int32_t nt_cap, int32_t* nt_out, double* dt_out);

/* DT_MIN of CALC_ADPSTP (src/cmf_ctrl_physics_mod.F90:178-201) of the last
 * advance, one value per instance (dt_min[ndom*members]; 0 without LADPSTP): the
 * third value of the reference's "ADPSTP: NT=" log line (physics:208). */
int cmf_hip_get_dt_min(cmf_handle h, double* dt_min);

/* Total storage of every instance: out[ndom*members] = the sum over the
 * instance's cells of P2RIVSTO + P2FLDSTO (+ P2LEVSTO with levees), the
 * P0GLBRIVSTO + P0GLBFLDSTO (+ P0GLBLEVSTO) of the reference's budget
 * (src/cmf_calc_fldstg_mod.F90:99-108) per instance. Deterministic (fixed
 * summation order for a given handle). No reference counterpart. */
int cmf_hip_instance_storage(cmf_handle h, double* out);

/* With kernel timing enabled: milliseconds from the start of the last advance to
 * the end of the last substep of each instance (ms[ndom*members]). */
int cmf_hip_instance_times(cmf_handle h, double* ms);

/* Launch configuration (no reference counterpart). Block sizes 64/128/256/512
 * per substep phase; fused = 1 runs a substep as one persistent cooperative
 * kernel of grid_fused blocks (0 = all co-resident blocks) of bs_fused threads;
 * xcd = 1 gives each XCD a contiguous range of tiles; graph = 1 replays the
 * substep launches of an outer step from a captured HIP graph (phase kernels
 * only). The configuration changes no state value, only the rounding of the
 * logged budget sums. */
typedef struct cmf_tune {
    int32_t bs_outflow, bs_rate, bs_update, fused, bs_fused, grid_fused, xcd, graph;
} cmf_tune;
int cmf_hip_set_tune(cmf_handle h, const cmf_tune* t);
int cmf_hip_get_tune(cmf_handle h, cmf_tune* t);
/* Time nsub substeps of an outer step of length dt under each candidate
 * configuration, keep the fastest (also written to *best if not NULL) and
 * restore the model state. report (optional) receives 8 values per trial:
 * fused, block size, xcd, fused grid, outflow/rate/update ms, total ms (kernel
 * event times; trials with fused = 2 / 3 are wall-clock times of the best
 * configuration without / with graph replay).
 * Returns the number of trials (>= 0) or a negative CMF_E* code. */
int cmf_hip_autotune(cmf_handle h, double dt, int32_t nsub, cmf_tune* best, double* report,
                     int32_t report_cap);

/* Wait for all work of the handle (no reference counterpart). */
int cmf_hip_synchronize(cmf_handle h);

/* Duration in milliseconds of the dominant substep kernels over the last
 * advance, measured with HIP events on the handle's stream (diagnostics for
 * benchmarking; no reference counterpart). out[0]=outflow, [1]=rate,
 * [2]=update(stage), [3]=stage+CFL, [4]=number of substeps timed. */
int cmf_hip_kernel_times(cmf_handle h, int32_t enable, double* out5);

#ifdef __cplusplus
}
#endif
#endif
