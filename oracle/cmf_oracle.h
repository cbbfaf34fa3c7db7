/*
 * cmf_oracle.h -- CPU restatement of the reference's per-substep physics (FP64 mode).
 *
 * TEST INFRASTRUCTURE ONLY. This oracle is the checker for the HIP physics: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (cama-flood_v4_amd/) never links, loads or calls it.
 *
 * It restates, function by function, the reference Fortran built in FP64 mode
 * (no -DSinglePrec_CMF), with the default namelist options that the device
 * path uses (LSPAMAT, local-inertial outflow, optional LFLDOUT/LPTHOUT/LDAMOUT/
 * LLEVEE/LADPSTP/LKINE/LUPSINF). Every function cites the reference file:line it
 * follows.
 *
 * Pinning: the physics (every function called by cmfo_physics_advance and
 * cmfo_physics_fldstg, the kinematic wave included) is bit-identical to the
 * reference's own run on the golden fixtures (every state array of every outer
 * step, the substep averages and maxima with D2DAMINF_aAVG included, NT, budget
 * sums; tests/test_oracle_golden.py); cmfo_roff_interp and the
 * output diagnostics are bit-identical to the reference routines compiled from
 * their sources (tests/test_forcing.py, tests/test_output_diag.py); the upstream
 * inflow term is pinned by the one-substep storage difference it must produce
 * (tests/test_steady_state.py).
 * Arithmetic keeps the reference's operation order and constant kinds: default-kind
 * literals such as 1.E-5 are the float value widened to double; x**(-1.) is 1/x,
 * x**0.5 is sqrt(x), x**(-0.5) is 1/sqrt(x), x**2 is x*x and x**3 is x*(x*x) (as
 * ROCm flang lowers them); other real powers call pow().
 *
 * Arrays are 0-based in the reference's NSEQ vector order. Layer/level tables are
 * planar: tab[l*n + i]. Pointers not needed by the enabled options may be NULL.
 */
#ifndef CMF_ORACLE_H
#define CMF_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct cmfo_model {
    /* sizes */
    int32_t n, nseqriv, nlfp, npth, npthlev, ndam;
    /* options (namelist NRUNVER) */
    int32_t ladpstp, lfldout, lpthout, ldamout, llevee;
    /* parameters (namelist NPARAM) */
    double pgrv, pmanfld, pdstmth, pcadp, pminslp, dfrcinc;
    /* topology */
    const int32_t* next;       /* downstream cell (0-based) or negative mouth code */
    const int32_t* up_ptr;     /* upstream CSR, ascending source order */
    const int32_t* up_idx;
    /* static map (JPRB, FP64 here) */
    const double *grarea, *elevtn, *nxtdst, *rivlen, *rivwth, *rivhgt, *rivman;
    const double *rivstomax, *rivelv, *dwnelv;
    const double *fldstomax, *fldgrd, *fldhgt;     /* [nlfp][n] */
    const int32_t* i2mask;
    /* bifurcation paths */
    const int32_t *pth_upst, *pth_down;            /* 0-based cell or -1 outside */
    const double *pth_dst, *pth_elv, *pth_wth, *pth_man;   /* elv/wth [npthlev][npth] */
    const int32_t *pout_ptr, *pout_idx, *pinf_ptr, *pinf_idx;  /* per-cell path lists */
    /* levee tables */
    const double *levhgt, *levfrc, *bashgt, *levdst, *levbassto, *levtopsto, *levfilsto;
    /* dams */
    const int32_t *i1dam, *damseq, *damstat;
    const double *dam_convol, *dam_emevol, *dam_adjvol, *dam_qa, *dam_qn, *dam_qf;
    /* prognostic state */
    double *rivsto, *fldsto, *levsto, *damsto, *daminf;
    double *rivout, *fldout, *rivout_pre, *fldout_pre, *rivdph_pre, *fldsto_pre;
    double *pthflw, *pthflw_pre;                   /* [npthlev][npth] */
    double *runoff;
    /* diagnostic state */
    double *rivdph, *flddph, *fldfrc, *fldare, *sfcelv, *storge, *outflw, *rivvel;
    double *rivinf, *fldinf, *pthout, *levdph, *pthflwsum;
    /* substep averages / maxima (CMF_DIAG_*_ADPSTP) */
    double *rivout_avg, *fldout_avg, *outflw_avg, *rivvel_avg, *pthout_avg, *runoff_avg;
    double *outflw_max, *rivdph_max, *storge_max;
    double *pthflw_avg, *pthflwsum_avg;            /* [npthlev][npth], [npth] */
    double nadd;
    /* global sums of the last substep: STOPRE STONXT STONEW RIVINF RIVOUT
     * STOPRE2 STONEW2 RIVSTO FLDSTO FLDARE LEVSTO */
    double glb[11];
    /* work arrays, each of length n */
    double *w_sfcelv_pre, *w_dwnelv_pre, *w_flddph_pre, *w_stoout, *w_rate;
    double *w_rivinf, *w_fldinf, *w_pthout;
    /* LUPSINF: upstream inflow D2UPSINF [m3/s] added to the river in STONXT
     * (src/cmf_calc_stonxt_mod.F90:81); NULL = off */
    const double* upsinf;
    /* LKINE: kinematic-wave outflow (CMF_CALC_OUTFLW_KINE) in place of the local
     * inertial one; 0 = off */
    int32_t lkine;
    /* optional record of the budget sums of every substep (the values
     * CALC_WATBAL logs per substep, physics:216-248): glb of substep it goes to
     * glb_hist[it*11 .. it*11+10] for it < glb_cap; NULL = off */
    double* glb_hist;
    int32_t glb_cap;
    /* LDAMH22: the Hanazaki-2022 dam rule with NorVol and the flood-control volume
     * over the upstream area per dam (damout:202-205, 354-378); 0 = Yamazaki-Funato */
    int32_t ldamh22;
    const double *dam_norvol, *dam_r_volupa;
    /* LSLPMIX: CMF_CALC_OUTFLW_KINEMIX, kinematic wave where I2MASK != 0 */
    int32_t lslpmix;
    /* D2DAMINF_aAVG of CMF_DIAG_*_ADPSTP (diag:130-135, :188-194) with LDAMOUT;
     * NULL = not kept */
    double* daminf_avg;
} cmfo_model;

/* CMF_PHYSICS_FLDSTG (src/cmf_ctrl_physics_mod.F90:288-305) */
void cmfo_physics_fldstg(cmfo_model* m);
/* CMF_PHYSICS_ADVANCE (src/cmf_ctrl_physics_mod.F90:21-155); dt is DT (in/out as
 * in the reference, restored on return). Returns NT. nt_cap>0 limits the number of
 * substeps run (for bounded timing samples); 0 runs all NT. */
int32_t cmfo_physics_advance(cmfo_model* m, double* dt, int32_t nt_cap);

/* single routines, exposed for unit tests */
void cmfo_calc_fldstg_def(cmfo_model* m);
void cmfo_levee_fldstg(cmfo_model* m);
int32_t cmfo_calc_adpstp(cmfo_model* m, double dt_def, double* dt);
void cmfo_calc_outflw(cmfo_model* m, double dt);
void cmfo_calc_outflw_kine(cmfo_model* m, double dt);
/* CMF_CALC_OUTPRE (src/cmf_opt_outflw_mod.F90:308-421): t-1 flows of a
 * storage-only restart from the current stage */
void cmfo_calc_outpre(cmfo_model* m);
void cmfo_calc_pthout(cmfo_model* m, double dt);
void cmfo_levee_opt_pthout(cmfo_model* m, double dt);
void cmfo_damout_calc(cmfo_model* m, double dt);
void cmfo_calc_outflw_kinemix(cmfo_model* m, double dt);
/* the Hanazaki-2022 operation rule of one dam before the limiter (damout:354-378) */
double cmfo_dam_rule_h22(double vol, double inflow, double norvol, double convol, double emevol, double qn,
                         double qf, double r_volupa);
void cmfo_calc_inflow_lspamat(cmfo_model* m, double dt);
void cmfo_damout_watbal(cmfo_model* m, double dt);
void cmfo_calc_vars_pre(cmfo_model* m);
void cmfo_calc_stonxt(cmfo_model* m, double dt);
void cmfo_diag_reset_adpstp(cmfo_model* m);
void cmfo_diag_avemax_adpstp(cmfo_model* m, double dt);
void cmfo_diag_getave_adpstp(cmfo_model* m);

/* Output-period accumulators (D2*_oAVG, D2*_oMAX, D1PTHFLW_oAVG of YOS_CMF_DIAG),
 * or, as the input of cmfo_diag_avemax_output, the substep averages / maxima of
 * one outer step (D2*_aAVG, D2*_aMAX). n cells, npthl = NPTHOUT*NPTHLEV. */
typedef struct cmfo_output {
    double *rivout, *fldout, *outflw, *rivvel, *pthout, *runoff, *daminf;
    double *outflw_max, *rivdph_max, *storge_max;
    double* pthflw;
    double nadd;
} cmfo_output;
/* CMF_DIAG_RESET_OUTPUT (src/cmf_calc_diag_mod.F90:222-267) */
void cmfo_diag_reset_output(cmfo_output* o, int32_t n, int32_t npthl, int32_t ldamout);
/* CMF_DIAG_AVEMAX_OUTPUT (src/cmf_calc_diag_mod.F90:273-317), DT = dt */
void cmfo_diag_avemax_output(cmfo_output* o, const cmfo_output* a, int32_t n, int32_t npthl, int32_t ldamout,
                             double dt);
/* CMF_DIAG_GETAVE_OUTPUT (src/cmf_calc_diag_mod.F90:323-356) */
void cmfo_diag_getave_output(cmfo_output* o, int32_t n, int32_t npthl, int32_t ldamout);

/* ROFF_INTERP (src/cmf_ctrl_forcing_mod.F90:757-790): out[i] = sum over the input
 * matrix entries k of cell i (inpx/inpy 1-based, [inpn][n]) with IXIN > 0 inside
 * the NXIN x NYIN grid and grid value != rmis of grid * INPA / DROFUNIT. */
void cmfo_roff_interp(int32_t n, int32_t inpn, int32_t nxin, int32_t nyin, const int32_t* inpx,
                      const int32_t* inpy, const double* inpa, const double* grid, double drofunit,
                      double rmis, double* out);

#ifdef __cplusplus
}
#endif
#endif
