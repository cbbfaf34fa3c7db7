/*
 * cmf_oracle.c -- CPU restatement of the reference's per-substep physics (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see cmf_oracle.h). Compiled with gcc
 * -ffp-contract=off; every statement keeps the reference's evaluation order.
 * Notation in comments: "1.E-5f" = default-kind literal widened to double.
 */
#include "cmf_oracle.h"
#include <math.h>
#include <string.h>

#define DMAX(a, b) ((a) > (b) ? (a) : (b))
#define DMIN(a, b) ((a) < (b) ? (a) : (b))

static const double F1EM5  = (double)1.e-5f;   /* default-kind 1.E-5  */
static const double F1EM10 = (double)1.e-10f;  /* default-kind 1.E-10 */
static const double F0P01  = (double)0.01f;    /* default-kind 0.01   */
static const double F2O3   = (double)(2.f / 3.f);
static const double F0P1   = (double)0.1f;
static const double P7O3   = -7.0 / 3.0;       /* -7._JPRB/3._JPRB */
static const double P4O3   = -4.0 / 3.0;       /* -4._JPRB/3._JPRB */
static const double P7O3F  = (double)(-7.f / 3.f);   /* default-kind -7./3. */
static const double P4O3F  = (double)(-4.f / 3.f);   /* default-kind -4./3. */

enum { G_STOPRE, G_STONXT, G_STONEW, G_RIVINF, G_RIVOUT,
       G_STOPRE2, G_STONEW2, G_RIVSTO, G_FLDSTO, G_FLDARE, G_LEVSTO };

/* ------------------------------------------------------------------------
 * CMF_CALC_FLDSTG_DEF  src/cmf_calc_fldstg_mod.F90:21-110 */
void cmfo_calc_fldstg_def(cmfo_model* m)
{
    const int n = m->n, nl = m->nlfp;
    double g_pre2 = 0.0, g_new2 = 0.0, g_riv = 0.0, g_fld = 0.0, g_are = 0.0;
    for (int i = 0; i < n; ++i) {
        const double pstoall = m->rivsto[i] + m->fldsto[i];
        const double dstoall = pstoall;
        g_pre2 = g_pre2 + pstoall;
        if (pstoall > m->rivstomax[i]) {
            int I = 1;
            double dsto_fil = m->rivstomax[i];
            double dwth_fil = m->rivwth[i];
            double ddph_fil = 0.0;
            const double dwth_inc = (m->grarea[i] / m->rivlen[i]) * m->dfrcinc;
            while (I <= nl && pstoall > m->fldstomax[(size_t)(I - 1) * n + i]) {     /* :59-65 */
                dsto_fil = m->fldstomax[(size_t)(I - 1) * n + i];
                dwth_fil = dwth_fil + dwth_inc;
                ddph_fil = ddph_fil + m->fldgrd[(size_t)(I - 1) * n + i] * dwth_inc;
                I = I + 1;
            }
            double dwth_add;
            if (I > nl) {                                                           /* :66-70 */
                const double dsto_add = dstoall - dsto_fil;
                dwth_add = 0.0;
                m->flddph[i] = ddph_fil + (dsto_add / dwth_fil) / m->rivlen[i];
            } else {                                                                /* :71-74 */
                const double g = m->fldgrd[(size_t)(I - 1) * n + i];
                const double dsto_add = dstoall - dsto_fil;
                dwth_add = -dwth_fil + sqrt(dwth_fil * dwth_fil + ((2.0 * dsto_add) / m->rivlen[i]) / g);
                m->flddph[i] = ddph_fil + g * dwth_add;
            }
            double rs = m->rivstomax[i] + (m->rivlen[i] * m->rivwth[i]) * m->flddph[i];
            rs = DMIN(rs, pstoall);
            m->rivsto[i] = rs;
            m->rivdph[i] = (rs / m->rivlen[i]) / m->rivwth[i];
            double fs = pstoall - rs;
            m->fldsto[i] = DMAX(fs, 0.0);
            double fr = ((-m->rivwth[i] + dwth_fil) + dwth_add) / (dwth_inc * (double)nl);
            fr = DMAX(fr, 0.0);
            fr = DMIN(fr, 1.0);
            m->fldfrc[i] = fr;
            m->fldare[i] = m->grarea[i] * fr;
        } else {                                                                    /* :87-95 */
            m->rivsto[i] = pstoall;
            double d = (dstoall / m->rivlen[i]) / m->rivwth[i];
            m->rivdph[i] = DMAX(d, 0.0);
            m->fldsto[i] = 0.0;
            m->flddph[i] = 0.0;
            m->fldfrc[i] = 0.0;
            m->fldare[i] = 0.0;
        }
    }
    for (int i = 0; i < n; ++i) {                                                   /* :99-108 */
        m->sfcelv[i] = m->rivelv[i] + m->rivdph[i];
        m->storge[i] = m->rivsto[i] + m->fldsto[i];
        g_new2 = (g_new2 + m->rivsto[i]) + m->fldsto[i];
        g_riv = g_riv + m->rivsto[i];
        g_fld = g_fld + m->fldsto[i];
        g_are = g_are + m->fldare[i];
    }
    m->glb[G_STOPRE2] = g_pre2; m->glb[G_STONEW2] = g_new2;
    m->glb[G_RIVSTO] = g_riv; m->glb[G_FLDSTO] = g_fld; m->glb[G_FLDARE] = g_are;
}

/* ------------------------------------------------------------------------
 * CMF_LEVEE_FLDSTG  src/cmf_ctrl_levee_mod.F90:271-490 */
void cmfo_levee_fldstg(cmfo_model* m)
{
    const int n = m->n, nl = m->nlfp;
#define FSM(l) m->fldstomax[(size_t)((l) - 1) * n + i]
#define FGR(l) m->fldgrd[(size_t)((l) - 1) * n + i]
#define FHG(l) m->fldhgt[(size_t)((l) - 1) * n + i]
    double g_pre2 = 0.0, g_new2 = 0.0, g_riv = 0.0, g_fld = 0.0, g_lev = 0.0, g_are = 0.0;
    for (int i = 0; i < n; ++i) {
        const double pstoall = (m->rivsto[i] + m->fldsto[i]) + m->levsto[i];
        const double dstoall = pstoall;
        g_pre2 = g_pre2 + pstoall;
        const double dwth_inc = (m->grarea[i] / m->rivlen[i]) * m->dfrcinc;
        const double rl = m->rivlen[i], rw = m->rivwth[i];
        if (dstoall > m->rivstomax[i]) {
            if (dstoall < m->levbassto[i]) {                    /* case 1  :312-349 */
                int I = 1;
                double dsto_fil = m->rivstomax[i], dwth_fil = rw, ddph_fil = 0.0, dwth_add;
                while (I <= nl && dstoall > FSM(I)) {
                    dsto_fil = FSM(I);
                    dwth_fil = dwth_fil + dwth_inc;
                    ddph_fil = ddph_fil + FGR(I) * dwth_inc;
                    I = I + 1;
                }
                if (I <= nl) {
                    const double dsto_add = dstoall - dsto_fil;
                    dwth_add = -dwth_fil + sqrt(dwth_fil * dwth_fil + ((2.0 * dsto_add) / rl) / FGR(I));
                    m->flddph[i] = ddph_fil + FGR(I) * dwth_add;
                } else {
                    const double dsto_add = dstoall - dsto_fil;
                    dwth_add = 0.0;
                    m->flddph[i] = ddph_fil + (dsto_add / dwth_fil) / rl;
                }
                m->rivsto[i] = m->rivstomax[i] + (rl * rw) * m->flddph[i];
                m->rivdph[i] = (m->rivsto[i] / rl) / rw;
                double fs = pstoall - m->rivsto[i];
                m->fldsto[i] = DMAX(fs, 0.0);
                double fr = ((-rw + dwth_fil) + dwth_add) / (dwth_inc * (double)nl);
                fr = DMAX(fr, 0.0);
                fr = DMIN(fr, 1.0);
                m->fldfrc[i] = fr;
                m->fldare[i] = m->grarea[i] * fr;
                m->levsto[i] = 0.0;
                m->levdph[i] = 0.0;
            } else if (dstoall < m->levtopsto[i]) {             /* case 2  :353-368 */
                const double dsto_add = dstoall - m->levbassto[i];
                const double dwth_add = m->levdst[i] + rw;
                m->flddph[i] = m->bashgt[i] + (dsto_add / dwth_add) / rl;
                m->rivsto[i] = m->rivstomax[i] + (rl * rw) * m->flddph[i];
                m->rivdph[i] = (m->rivsto[i] / rl) / rw;
                double fs = pstoall - m->rivsto[i];
                m->fldsto[i] = DMAX(fs, 0.0);
                m->fldfrc[i] = m->levfrc[i];
                m->fldare[i] = m->grarea[i] * m->fldfrc[i];
                m->levsto[i] = 0.0;
                m->levdph[i] = 0.0;
            } else if (dstoall < m->levfilsto[i]) {             /* case 3  :372-420 */
                m->flddph[i] = m->levhgt[i];
                m->rivsto[i] = m->rivstomax[i] + (rl * rw) * m->flddph[i];
                m->rivdph[i] = (m->rivsto[i] / rl) / rw;
                double fs = m->levtopsto[i] - m->rivsto[i];
                m->fldsto[i] = DMAX(fs, 0.0);
                double ls = (pstoall - m->rivsto[i]) - m->fldsto[i];
                m->levsto[i] = DMAX(ls, 0.0);
                const int ilev = (int)(m->levfrc[i] * (double)nl) + 1;
                double dsto_fil = m->levtopsto[i], dwth_fil = 0.0, ddph_fil = 0.0;
                int I = ilev;
                while (I <= nl) {
                    const double dsto_add = ((m->levdst[i] + rw) * (m->levhgt[i] - FHG(I))) * rl;
                    if (dstoall < FSM(I) + dsto_add) break;
                    dsto_fil = FSM(I) + dsto_add;
                    dwth_fil = dwth_inc * (double)I - m->levdst[i];
                    ddph_fil = FHG(I) - m->bashgt[i];
                    I = I + 1;
                }
                if (I <= nl) {
                    const double dsto_add = dstoall - dsto_fil;
                    const double dwth_add = -dwth_fil + sqrt(dwth_fil * dwth_fil + ((2.0 * dsto_add) / rl) / FGR(I));
                    const double ddph_add = dwth_add * FGR(I);
                    m->levdph[i] = (m->bashgt[i] + ddph_fil) + ddph_add;
                    double fr = (dwth_fil + m->levdst[i]) / (dwth_inc * (double)nl);
                    fr = DMAX(fr, 0.0);
                    fr = DMIN(fr, 1.0);
                    m->fldfrc[i] = fr;
                    m->fldare[i] = m->grarea[i] * fr;
                } else {
                    const double dsto_add = dstoall - dsto_fil;
                    const double ddph_add = (dsto_add / dwth_fil) / rl;
                    m->levdph[i] = (m->bashgt[i] + ddph_fil) + ddph_add;
                    m->fldfrc[i] = 1.0;
                    m->fldare[i] = m->grarea[i] * m->fldfrc[i];
                }
            } else {                                            /* case 4  :424-461 */
                int I = 1;
                double dsto_fil = m->rivstomax[i], dwth_fil = rw, ddph_fil = 0.0, dwth_add;
                while (I <= nl && dstoall > FSM(I)) {
                    dsto_fil = FSM(I);
                    dwth_fil = dwth_fil + dwth_inc;
                    ddph_fil = ddph_fil + FGR(I) * dwth_inc;
                    I = I + 1;
                }
                if (I <= nl) {
                    const double dsto_add = dstoall - dsto_fil;
                    dwth_add = -dwth_fil + sqrt(dwth_fil * dwth_fil + ((2.0 * dsto_add) / rl) / FGR(I));
                    m->flddph[i] = ddph_fil + FGR(I) * dwth_add;
                } else {
                    const double dsto_add = dstoall - dsto_fil;
                    dwth_add = 0.0;
                    m->flddph[i] = ddph_fil + (dsto_add / dwth_fil) / rl;
                }
                m->fldfrc[i] = ((-rw + dwth_fil) + dwth_add) / (dwth_inc * (double)nl);
                m->fldare[i] = m->grarea[i] * m->fldfrc[i];
                m->rivsto[i] = m->rivstomax[i] + (rl * rw) * m->flddph[i];
                m->rivdph[i] = (m->rivsto[i] / rl) / rw;
                const double dsto_add = ((m->flddph[i] - m->levhgt[i]) * (m->levdst[i] + rw)) * rl;
                double fs = (m->levtopsto[i] + dsto_add) - m->rivsto[i];
                m->fldsto[i] = DMAX(fs, 0.0);
                double ls = (pstoall - m->rivsto[i]) - m->fldsto[i];
                m->levsto[i] = DMAX(ls, 0.0);
                m->levdph[i] = m->flddph[i];
            }
        } else {                                                /* case 0  :463-473 */
            m->rivsto[i] = pstoall;
            double d = (dstoall / rl) / rw;
            m->rivdph[i] = DMAX(d, 0.0);
            m->fldsto[i] = 0.0;
            m->flddph[i] = 0.0;
            m->fldfrc[i] = 0.0;
            m->fldare[i] = 0.0;
            m->levsto[i] = 0.0;
            m->levdph[i] = 0.0;
        }
    }
#undef FSM
#undef FGR
#undef FHG
    for (int i = 0; i < n; ++i) {                                                   /* :477-487 */
        m->sfcelv[i] = m->rivelv[i] + m->rivdph[i];
        m->storge[i] = (m->rivsto[i] + m->fldsto[i]) + m->levsto[i];
        g_new2 = ((g_new2 + m->rivsto[i]) + m->fldsto[i]) + m->levsto[i];
        g_riv = g_riv + m->rivsto[i];
        g_fld = g_fld + m->fldsto[i];
        g_lev = g_lev + m->levsto[i];
        g_are = g_are + m->fldare[i];
    }
    m->glb[G_STOPRE2] = g_pre2; m->glb[G_STONEW2] = g_new2; m->glb[G_RIVSTO] = g_riv;
    m->glb[G_FLDSTO] = g_fld; m->glb[G_FLDARE] = g_are; m->glb[G_LEVSTO] = g_lev;
}

/* CMF_PHYSICS_FLDSTG  src/cmf_ctrl_physics_mod.F90:288-305 (LSTG_ES not supported) */
void cmfo_physics_fldstg(cmfo_model* m)
{
    if (m->llevee) cmfo_levee_fldstg(m);
    else cmfo_calc_fldstg_def(m);
}

/* ------------------------------------------------------------------------
 * CALC_ADPSTP  src/cmf_ctrl_physics_mod.F90:161-210 */
int32_t cmfo_calc_adpstp(cmfo_model* m, double dt_def, double* dt)
{
    double dt_min = dt_def;
    for (int i = 0; i < m->nseqriv; ++i) {
        if (m->i2mask[i] == 0) {
            const double ddph = DMAX(m->rivdph[i], 0.01);
            const double ddst = m->nxtdst[i];
            const double c = (m->pcadp * ddst) * (1.0 / sqrt(m->pgrv * ddph));
            dt_min = DMIN(dt_min, c);
        }
    }
    for (int i = m->nseqriv; i < m->n; ++i) {
        if (m->i2mask[i] == 0) {
            const double ddph = DMAX(m->rivdph[i], 0.01);
            const double ddst = m->pdstmth;
            const double c = (m->pcadp * ddst) * (1.0 / sqrt(m->pgrv * ddph));
            dt_min = DMIN(dt_min, c);
        }
    }
    const int32_t nt = (int32_t)(dt_def * (1.0 / dt_min) - F0P01) + 1;      /* :205 */
    *dt = dt_def * (double)(1.0f / (float)nt);                              /* :206 */
    return nt;
}

/* ------------------------------------------------------------------------
 * CMF_CALC_OUTFLW  src/cmf_calc_outflw_mod.F90:32-185 */
void cmfo_calc_outflw(cmfo_model* m, double dt)
{
    const int n = m->n, nr = m->nseqriv;
    double* sp = m->w_sfcelv_pre;
    double* dw = m->w_dwnelv_pre;     /* holds DWNELV for river cells */
    double* dwp = m->w_rate;          /* holds DWNELV_PRE for river cells */
    double* fdp = m->w_flddph_pre;
    const double gdt = m->pgrv * dt;
    for (int i = 0; i < n; ++i) {                                           /* :43-48 */
        m->sfcelv[i] = m->rivelv[i] + m->rivdph[i];
        sp[i] = m->rivelv[i] + m->rivdph_pre[i];
        double d = m->rivdph_pre[i] - m->rivhgt[i];
        fdp[i] = DMAX(d, 0.0);
    }
    for (int i = 0; i < nr; ++i) {                                          /* :52-57 */
        const int j = m->next[i];
        dw[i] = m->sfcelv[j];
        dwp[i] = sp[j];
    }
    for (int i = 0; i < nr; ++i) {                                          /* :60-80 river */
        const double dsfc = DMAX(m->sfcelv[i], dw[i]);
        const double dslp = (m->sfcelv[i] - dw[i]) * (1.0 / m->nxtdst[i]);
        const double dflw = dsfc - m->rivelv[i];
        double dare = m->rivwth[i] * dflw;
        dare = DMAX(dare, 1.e-10);
        const double dsfc_pr = DMAX(sp[i], dwp[i]);
        const double dflw_pr = dsfc_pr - m->rivelv[i];
        double dflw_im = sqrt(dflw * dflw_pr);
        dflw_im = DMAX(dflw_im, 1.e-6);
        const double dout_pr = m->rivout_pre[i] * (1.0 / m->rivwth[i]);
        const double dout = (m->rivwth[i] * (dout_pr + (gdt * dflw_im) * dslp)) *
            (1.0 / (1.0 + ((gdt * (m->rivman[i] * m->rivman[i])) * fabs(dout_pr)) * pow(dflw_im, P7O3)));
        const double dvel = m->rivout[i] * (1.0 / dare);
        const int mask = (dflw_im > F1EM5) && (dare > F1EM5);
        m->rivout[i] = mask ? dout : 0.0;
        m->rivvel[i] = mask ? dvel : 0.0;
    }
    if (m->lfldout) {                                                       /* :84-116 floodplain */
        const double gdtn2 = gdt * (m->pmanfld * m->pmanfld);
        for (int i = 0; i < nr; ++i) {
            const double dfsto = m->fldsto[i];
            const double dsfc = DMAX(m->sfcelv[i], dw[i]);
            double dslp = (m->sfcelv[i] - dw[i]) * (1.0 / m->nxtdst[i]);
            dslp = DMAX(-0.005, DMIN(0.005, dslp));
            double dflw = dsfc - m->elevtn[i];
            dflw = DMAX(dflw, 0.0);
            double dare = dfsto * (1.0 / m->rivlen[i]);
            dare = DMAX(dare - m->flddph[i] * m->rivwth[i], 0.0);
            const double dsfc_pr = DMAX(sp[i], dwp[i]);
            const double dflw_pr = dsfc_pr - m->elevtn[i];
            double t = dflw * dflw_pr;
            double dflw_im = sqrt(DMAX(t, 0.0));
            dflw_im = DMAX(dflw_im, 1.e-6);
            double dare_pr = m->fldsto_pre[i] * (1.0 / m->rivlen[i]);
            dare_pr = DMAX(dare_pr - fdp[i] * m->rivwth[i], 1.e-6);
            double dare_im = sqrt(dare * dare_pr);
            dare_im = DMAX(dare_im, 1.e-6);
            const double dout_pr = m->fldout_pre[i];
            const double dout = (dout_pr + (gdt * dare_im) * dslp) *
                (1.0 / (1.0 + ((gdtn2 * fabs(dout_pr)) * pow(dflw_im, P4O3)) * (1.0 / dare_im)));
            const int mask = (dflw_im > 1.e-5) && (dare > 1.e-5);
            double f = mask ? dout : 0.0;
            const int mask2 = (f * m->rivout[i] > 0.0);
            m->fldout[i] = mask2 ? f : 0.0;
        }
    }
    for (int i = nr; i < n; ++i) {                                          /* :119-139 mouth */
        const double dslp = (m->sfcelv[i] - m->dwnelv[i]) * (1.0 / m->pdstmth);
        const double dflw = m->rivdph[i];
        double dare = m->rivwth[i] * dflw;
        dare = DMAX(dare, 1.e-10);
        const double dflw_pr = m->rivdph_pre[i];
        double dflw_im = sqrt(dflw * dflw_pr);
        dflw_im = DMAX(dflw_im, 1.e-6);
        const double dout_pr = m->rivout_pre[i] * (1.0 / m->rivwth[i]);
        const double dout = (m->rivwth[i] * (dout_pr + (gdt * dflw_im) * dslp)) *
            (1.0 / (1.0 + ((gdt * (m->rivman[i] * m->rivman[i])) * fabs(dout_pr)) * pow(dflw_im, P7O3)));
        const double dvel = m->rivout[i] * (1.0 / dare);
        const int mask = (dflw_im > F1EM5) && (dare > F1EM5);
        m->rivout[i] = mask ? dout : 0.0;
        m->rivvel[i] = mask ? dvel : 0.0;
    }
    if (m->lfldout) {                                                       /* :143-171 mouth fld */
        const double gdtn2 = gdt * (m->pmanfld * m->pmanfld);
        for (int i = nr; i < n; ++i) {
            const double dfsto = m->fldsto[i];
            double dslp = (m->sfcelv[i] - m->dwnelv[i]) * (1.0 / m->pdstmth);
            dslp = DMAX(-0.005, DMIN(0.005, dslp));
            const double dflw = m->sfcelv[i] - m->elevtn[i];
            double dare = DMAX(dfsto * (1.0 / m->rivlen[i]) - m->flddph[i] * m->rivwth[i], 0.0);
            const double dflw_pr = sp[i] - m->elevtn[i];
            double t = dflw * dflw_pr;
            double dflw_im = sqrt(DMAX(t, 0.0));
            dflw_im = DMAX(dflw_im, 1.e-6);
            double dare_pr = DMAX(m->fldsto_pre[i] * (1.0 / m->rivlen[i]) - fdp[i] * m->rivwth[i], 1.e-6);
            double dare_im = sqrt(dare * dare_pr);
            dare_im = DMAX(dare_im, 1.e-6);
            const double dout_pr = m->fldout_pre[i];
            const double dout = (dout_pr + (gdt * dare_im) * dslp) *
                (1.0 / (1.0 + ((gdtn2 * fabs(dout_pr)) * pow(dflw_im, P4O3)) * (1.0 / dare_im)));
            const int mask = (dflw_im > F1EM5) && (dare > F1EM5);
            double f = mask ? dout : 0.0;
            const int mask2 = (f * m->rivout[i] > 0.0);
            m->fldout[i] = mask2 ? f : 0.0;
        }
    }
    for (int i = 0; i < nr; ++i) {                                          /* :175-182 limiter */
        double dout = ((-m->rivout[i]) - m->fldout[i]) * dt;
        dout = DMAX(dout, F1EM10);
        double rate = (0.05 * m->storge[i]) / dout;
        rate = DMIN(rate, 1.0);
        m->rivout[i] = m->rivout[i] * rate;
        m->fldout[i] = m->fldout[i] * rate;
    }
}

/* ------------------------------------------------------------------------
 * path limiter shared by CMF_CALC_PTHOUT (src/cmf_calc_pthout_mod.F90:77-95) and
 * CMF_LEVEE_OPT_PTHOUT (src/cmf_ctrl_levee_mod.F90:583-601) */
static void pth_sum_and_limit(cmfo_model* m, double dt)
{
    const int np = m->npth, nl = m->npthlev;
    for (int p = 0; p < np; ++p) {
        double s = 0.0;
        for (int l = 0; l < nl; ++l) s = s + m->pthflw[(size_t)l * np + p];
        m->pthflwsum[p] = s;
    }
    for (int p = 0; p < np; ++p) {
        if (m->pthflwsum[p] != 0.0) {
            const int ip = m->pth_upst[p], jp = m->pth_down[p];
            double rate = (0.05 * DMIN(m->storge[ip], m->storge[jp])) / fabs(m->pthflwsum[p] * dt);
            rate = DMIN(rate, 1.0);
            for (int l = 0; l < nl; ++l) m->pthflw[(size_t)l * np + p] = m->pthflw[(size_t)l * np + p] * rate;
            m->pthflwsum[p] = m->pthflwsum[p] * rate;
        }
    }
}

/* CMF_CALC_PTHOUT  src/cmf_calc_pthout_mod.F90:20-97 */
void cmfo_calc_pthout(cmfo_model* m, double dt)
{
    const int n = m->n, np = m->npth, nl = m->npthlev;
    double* sp = m->w_sfcelv_pre;
    const double gdt = m->pgrv * dt;
    for (int i = 0; i < n; ++i) sp[i] = m->rivelv[i] + m->rivdph_pre[i];
    for (size_t k = 0; k < (size_t)np * nl; ++k) m->pthflw[k] = 0.0;
    for (int p = 0; p < np; ++p) {
        const int ip = m->pth_upst[p], jp = m->pth_down[p];
        if (ip < 0 || jp < 0) continue;
        if (m->i2mask[ip] > 0 || m->i2mask[jp] > 0) continue;
        double dslp = (m->sfcelv[ip] - m->sfcelv[jp]) / m->pth_dst[p];
        dslp = DMAX(-0.005, DMIN(0.005, dslp));
        for (int l = 0; l < nl; ++l) {
            const double elv = m->pth_elv[(size_t)l * np + p];
            double dflw = DMAX(m->sfcelv[ip], m->sfcelv[jp]) - elv;
            dflw = DMAX(dflw, 0.0);
            double dflw_pr = DMAX(sp[ip], sp[jp]) - elv;
            dflw_pr = DMAX(dflw_pr, 0.0);
            double dflw_im = sqrt(dflw * dflw_pr);
            dflw_im = DMAX(dflw_im, sqrt(dflw * 0.01));
            double q = 0.0;
            if (dflw_im > 1.e-5) {
                const double w = m->pth_wth[(size_t)l * np + p];
                const double man = m->pth_man[l];
                const double dout_pr = m->pthflw_pre[(size_t)l * np + p] / w;
                q = (w * (dout_pr + (gdt * dflw_im) * dslp)) /
                    (1.0 + ((gdt * (man * man)) * fabs(dout_pr)) * pow(dflw_im, P7O3));
            }
            m->pthflw[(size_t)l * np + p] = q;
        }
    }
    pth_sum_and_limit(m, dt);
}

/* CMF_LEVEE_OPT_PTHOUT  src/cmf_ctrl_levee_mod.F90:497-603 */
void cmfo_levee_opt_pthout(cmfo_model* m, double dt)
{
    const int n = m->n, np = m->npth, nl = m->npthlev;
    double* sp = m->w_sfcelv_pre;
    double* sl = m->w_stoout;           /* D2SFCELV_LEV */
    const double gdt = m->pgrv * dt;
    for (int i = 0; i < n; ++i) {
        sl[i] = m->sfcelv[i];
        if (m->levfrc[i] < 1.0) sl[i] = DMIN(m->elevtn[i] + m->levdph[i], m->sfcelv[i]);
        sp[i] = m->rivelv[i] + m->rivdph_pre[i];
    }
    for (size_t k = 0; k < (size_t)np * nl; ++k) m->pthflw[k] = 0.0;
    for (int p = 0; p < np; ++p) {
        const int ip = m->pth_upst[p], jp = m->pth_down[p];
        if (ip < 0 || jp < 0) continue;
        if (m->i2mask[ip] == 1 || m->i2mask[jp] == 1) continue;
        double dslp = (m->sfcelv[ip] - m->sfcelv[jp]) / m->pth_dst[p];
        dslp = DMAX(-0.005, DMIN(0.005, dslp));
        {   /* level 1: river channel (:540-556) */
            const double elv = m->pth_elv[p];
            double dflw = DMAX(m->sfcelv[ip], m->sfcelv[jp]) - elv;
            dflw = DMAX(dflw, 0.0);
            double dflw_pr = DMAX(sp[ip], sp[jp]) - elv;
            dflw_pr = DMAX(dflw_pr, 0.0);
            double dflw_im = sqrt(dflw * dflw_pr);
            if (dflw_im <= 0.0) dflw_im = dflw;
            double q = 0.0;
            if (dflw_im > 1.e-5) {
                const double w = m->pth_wth[p], man = m->pth_man[0];
                const double dout_pr = m->pthflw_pre[p] / w;
                q = (w * (dout_pr + (gdt * dflw_im) * dslp)) /
                    (1.0 + ((gdt * (man * man)) * fabs(dout_pr)) * pow(dflw_im, P7O3));
            }
            m->pthflw[p] = q;
        }
        if (nl <= 1) continue;
        for (int l = 1; l < nl; ++l) {  /* overland levels (:565-577) */
            const double elv = m->pth_elv[(size_t)l * np + p];
            double dflw = DMAX(sl[ip], sl[jp]) - elv;
            dflw = DMAX(dflw, 0.0);
            const double dflw_im = dflw;
            double q = 0.0;
            if (dflw_im > 1.e-5) {
                const double w = m->pth_wth[(size_t)l * np + p], man = m->pth_man[l];
                const double dout_pr = m->pthflw_pre[(size_t)l * np + p] * (1.0 / w);
                q = (w * (dout_pr + (gdt * dflw_im) * dslp)) /
                    (1.0 + ((gdt * (man * man)) * fabs(dout_pr)) * pow(dflw_im, P7O3));
            }
            m->pthflw[(size_t)l * np + p] = q;
        }
    }
    pth_sum_and_limit(m, dt);
}

/* ------------------------------------------------------------------------
 * CMF_DAMOUT_CALC  src/cmf_ctrl_damout_mod.F90:316-498 (Yamazaki & Funato rule;
 * the NoAtom_CMF serial order of the dam-inflow scatter at :456-464) */
/* CMF_DAMOUT_CALC, Hanazaki-2022 branch (LDAMH22)  src/cmf_ctrl_damout_mod.F90:354-378.
 * The ratio r_volupa is the dam's own FldVol*1.E-6/upreal (damout:204 writes every
 * dam's ratio to R_VolUpa(NDAM); here each dam keeps its own). */
double cmfo_dam_rule_h22(double vol, double inflow, double norvol, double convol, double emevol, double qn,
                         double qf, double r_volupa)
{
    const double F0P2 = (double)0.2f;                                      /* default-kind 0.2 */
    double out;
    if (vol <= norvol) {                                                    /* :356-357 */
        out = qn * (vol / convol);
    } else if (norvol < vol && vol <= convol) {                             /* :359-365 */
        if (qf <= inflow) {
            out = qn * 0.5 + ((vol - norvol) / (convol - norvol)) * (qf - qn);
        } else {
            const double x = (vol - norvol) / (emevol - norvol);
            out = qn * 0.5 + (x * x) * (qf - qn);
        }
    } else if (convol < vol && vol < emevol) {                              /* :367-373 */
        if (qf <= inflow) {
            out = qf + ((DMAX(1.0 - r_volupa / F0P2, 0.0) * (vol - convol)) / (emevol - convol)) * (inflow - qf);
        } else {
            const double x = (vol - norvol) / (emevol - norvol);
            out = qn * 0.5 + (x * x) * (qf - qn);
        }
    } else {                                                                /* :375-376 */
        out = DMAX(inflow, qf);
    }
    return out;
}

void cmfo_damout_calc(cmfo_model* m, double dt)
{
    const int n = m->n;
    for (int i = 0; i < n; ++i) {                                           /* :443-449 */
        if (m->i1dam[i] > 0) {
            m->rivout[i] = 0.0;
            m->fldout[i] = 0.0;
            m->daminf[i] = m->runoff[i];
        }
    }
    for (int i = 0; i < n; ++i) {                                           /* :456-464 */
        if (m->i1dam[i] == 10 || m->i1dam[i] == 11) {
            const int j = m->next[i];
            m->daminf[j] = (m->daminf[j] + m->rivout_pre[i]) + m->fldout_pre[i];
        }
    }
    for (int i = 0; i < m->nseqriv; ++i) {                                  /* :471-493 */
        if (m->i1dam[i] == 10) {
            const int j = m->next[i];
            double dslope = (m->elevtn[i] - m->elevtn[j]) * (1.0 / m->nxtdst[i]);
            dslope = DMAX(dslope, m->pminslp);
            const double dvel = ((1.0 / m->rivman[i]) * sqrt(dslope)) * pow(m->rivdph[i], F2O3);
            const double darea = m->rivwth[i] * m->rivdph[i];
            m->rivvel[i] = dvel;
            double q = darea * dvel;
            m->rivout[i] = DMIN(q, m->rivsto[i] / dt);
            const double dslope_f = DMIN(0.005, dslope);
            const double dvel_f = ((1.0 / m->pmanfld) * sqrt(dslope_f)) * pow(m->flddph[i], F2O3);
            double dare_f = m->fldsto[i] * (1.0 / m->rivlen[i]);
            dare_f = DMAX(dare_f - m->flddph[i] * m->rivwth[i], 0.0);
            double qf = dare_f * dvel_f;
            m->fldout[i] = DMIN(qf, m->fldsto[i] / dt);
        }
    }
    for (int d = 0; d < m->ndam; ++d) {                                     /* :344-419 */
        if (m->damstat[d] <= 0) continue;
        const int is = m->damseq[d];
        const double vol = m->damsto[is];
        const double inf = m->daminf[is];
        const double cv = m->dam_convol[d], av = m->dam_adjvol[d], ev = m->dam_emevol[d];
        const double qn = m->dam_qn[d], qf = m->dam_qf[d], qa = m->dam_qa[d];
        double out = 0.0;
        if (m->ldamh22) {
            out = cmfo_dam_rule_h22(vol, inf, m->dam_norvol[d], cv, ev, qn, qf, m->dam_r_volupa[d]);
        } else if (vol <= cv) {
            out = qn * sqrt(vol / cv);
        } else if (vol > cv && vol <= av) {
            const double x = (vol - cv) / (av - cv);
            out = qn + (x * (x * x)) * (qa - qn);
        } else if (vol > av && vol <= ev) {
            if (inf >= qf) {
                out = qn + ((vol - cv) / (ev - cv)) * (inf - qn);
                const double tmp = qa + pow((vol - av) / (ev - av), F0P1) * (qf - qa);
                out = DMAX(out, tmp);
            } else {
                out = qa + pow((vol - av) / (ev - av), F0P1) * (qf - qa);
            }
        } else if (vol > ev) {
            out = (inf >= qf) ? inf : qf;
        }
        out = DMIN(DMIN(out, vol / dt), (m->rivsto[is] + m->fldsto[is]) / dt);
        out = DMAX(out, 0.0);
        m->rivout[is] = out;
        m->fldout[is] = 0.0;
    }
}

/* CMF_DAMOUT_WATBAL  src/cmf_ctrl_damout_mod.F90:504-542 */
void cmfo_damout_watbal(cmfo_model* m, double dt)
{
    for (int d = 0; d < m->ndam; ++d) {
        if (m->damstat[d] == -9999) continue;
        const int is = m->damseq[d];
        const double inflow = (m->rivinf[is] + m->fldinf[is]) + m->runoff[is];
        const double outflw = m->rivout[is] + m->fldout[is];
        m->damsto[is] = (m->damsto[is] + inflow * dt) - outflw * dt;
    }
}

/* ------------------------------------------------------------------------
 * CMF_CALC_INFLOW_LSPAMAT  src/cmf_calc_outflw_mod.F90:386-546 */
void cmfo_calc_inflow_lspamat(cmfo_model* m, double dt)
{
    const int n = m->n, nr = m->nseqriv, np = m->npth, nl = m->npthlev;
    double* so = m->w_stoout;
    double* rate = m->w_rate;
    for (int i = 0; i < n; ++i) {                                           /* :410-419 */
        double s = DMAX(m->rivout[i], 0.0) + DMAX(m->fldout[i], 0.0);
        for (int k = m->up_ptr[i]; k < m->up_ptr[i + 1]; ++k) {
            const int j = m->up_idx[k];
            s = (s + DMAX(-m->rivout[j], 0.0)) + DMAX(-m->fldout[j], 0.0);
        }
        so[i] = s * dt;
    }
    if (m->lpthout) {                                                       /* :423-441 */
        for (int i = 0; i < n; ++i) {
            for (int k = m->pout_ptr[i]; k < m->pout_ptr[i + 1]; ++k)
                so[i] = so[i] + DMAX(m->pthflwsum[m->pout_idx[k]], 0.0) * dt;
            for (int k = m->pinf_ptr[i]; k < m->pinf_ptr[i + 1]; ++k)
                so[i] = so[i] + DMAX(-m->pthflwsum[m->pinf_idx[k]], 0.0) * dt;
        }
    }
    for (int i = 0; i < n; ++i) {                                           /* :446-451 */
        const double dout = DMAX(so[i], 1.e-10);
        const double dsto = m->rivsto[i] + m->fldsto[i];
        rate[i] = DMIN(dsto / dout, 1.0);
    }
    for (int i = 0; i < nr; ++i) {                                          /* :456-465 */
        const double r = (m->rivout[i] > 0.0) ? rate[i] : rate[m->next[i]];
        m->rivout[i] = m->rivout[i] * r;
        m->fldout[i] = m->fldout[i] * r;
    }
    for (int i = nr; i < n; ++i) {                                          /* :470-473 */
        m->rivout[i] = m->rivout[i] * rate[i];
        m->fldout[i] = m->fldout[i] * rate[i];
    }
    for (int i = 0; i < n; ++i) {                                           /* :481-489 */
        double ri = 0.0, fi = 0.0;
        for (int k = m->up_ptr[i]; k < m->up_ptr[i + 1]; ++k) {
            const int j = m->up_idx[k];
            ri = ri + m->rivout[j];
            fi = fi + m->fldout[j];
        }
        m->rivinf[i] = ri;
        m->fldinf[i] = fi;
        m->pthout[i] = 0.0;
    }
    if (m->lpthout) {                                                       /* :494-535 */
        for (int p = 0; p < np; ++p) {
            const int ip = m->pth_upst[p], jp = m->pth_down[p];
            if (ip < 0 || jp < 0) continue;
            if (m->i2mask[ip] > 0 || m->i2mask[jp] > 0) continue;
            for (int l = 0; l < nl; ++l) {
                double* q = &m->pthflw[(size_t)l * np + p];
                *q = (*q >= 0.0) ? *q * rate[ip] : *q * rate[jp];
            }
            m->pthflwsum[p] = (m->pthflwsum[p] >= 0.0) ? m->pthflwsum[p] * rate[ip]
                                                       : m->pthflwsum[p] * rate[jp];
        }
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int k = m->pout_ptr[i]; k < m->pout_ptr[i + 1]; ++k) s = s + m->pthflwsum[m->pout_idx[k]];
            for (int k = m->pinf_ptr[i]; k < m->pinf_ptr[i + 1]; ++k) s = s - m->pthflwsum[m->pinf_idx[k]];
            m->pthout[i] = s;
        }
    }
}

/* CALC_VARS_PRE  src/cmf_ctrl_physics_mod.F90:254-276 */
void cmfo_calc_vars_pre(cmfo_model* m)
{
    for (int i = 0; i < m->n; ++i) {
        m->rivout_pre[i] = m->rivout[i];
        m->rivdph_pre[i] = m->rivdph[i];
        m->fldout_pre[i] = m->fldout[i];
        m->fldsto_pre[i] = m->fldsto[i];
    }
    if (m->lpthout)
        memcpy(m->pthflw_pre, m->pthflw, sizeof(double) * (size_t)m->npth * m->npthlev);
}

/* CMF_CALC_STONXT  src/cmf_calc_stonxt_mod.F90:20-100 (no groundwater delay,
 * no sub-surface split, no evaporation, no prescribed upstream inflow) */
void cmfo_calc_stonxt(cmfo_model* m, double dt)
{
    double g_pre = 0.0, g_inf = 0.0, g_out = 0.0, g_nxt = 0.0, g_new = 0.0;
    for (int i = 0; i < m->n; ++i) {
        g_pre = (g_pre + m->rivsto[i]) + m->fldsto[i];
        g_inf = (g_inf + m->rivinf[i] * dt) + m->fldinf[i] * dt;
        g_out = ((g_out + m->rivout[i] * dt) + m->fldout[i] * dt) + m->pthout[i] * dt;
        double rs = (m->rivsto[i] + m->rivinf[i] * dt) - m->rivout[i] * dt;
        double fs = m->fldsto[i];
        if (rs < 0.0) { fs = fs + rs; rs = 0.0; }
        fs = ((fs + m->fldinf[i] * dt) - m->fldout[i] * dt) - m->pthout[i] * dt;
        if (fs < 0.0) { rs = DMAX(rs + fs, 0.0); fs = 0.0; }
        g_nxt = (g_nxt + rs) + fs;
        m->outflw[i] = m->rivout[i] + m->fldout[i];
        double drivrof = (m->runoff[i] * (1.0 - m->fldfrc[i])) * dt;
        const double dfldrof = (m->runoff[i] * m->fldfrc[i]) * dt;
        if (m->upsinf) drivrof = drivrof + m->upsinf[i] * dt;               /* stonxt:81 */
        rs = rs + drivrof;
        fs = fs + dfldrof;
        m->rivsto[i] = rs;
        m->fldsto[i] = fs;
        double st = rs + fs;
        if (m->llevee) st = st + m->levsto[i];
        m->storge[i] = st;
        g_new = g_new + st;
    }
    m->glb[G_STOPRE] = g_pre; m->glb[G_STONXT] = g_nxt; m->glb[G_STONEW] = g_new;
    m->glb[G_RIVINF] = g_inf; m->glb[G_RIVOUT] = g_out;
}

/* ------------------------------------------------------------------------
 * CMF_DIAG_{RESET,AVEMAX,GETAVE}_ADPSTP  src/cmf_calc_diag_mod.F90:50-215 */
void cmfo_diag_reset_adpstp(cmfo_model* m)
{
    m->nadd = 0.0;
    for (int i = 0; i < m->n; ++i) {
        m->rivout_avg[i] = 0.0; m->fldout_avg[i] = 0.0; m->outflw_avg[i] = 0.0;
        m->rivvel_avg[i] = 0.0; m->pthout_avg[i] = 0.0; m->runoff_avg[i] = 0.0;
        m->storge_max[i] = 0.0; m->outflw_max[i] = 0.0; m->rivdph_max[i] = 0.0;
    }
    if (m->lpthout) {
        for (size_t k = 0; k < (size_t)m->npth * m->npthlev; ++k) m->pthflw_avg[k] = 0.0;
        for (int p = 0; p < m->npth; ++p) m->pthflwsum_avg[p] = 0.0;
    }
    if (m->ldamout && m->daminf_avg)
        for (int i = 0; i < m->n; ++i) m->daminf_avg[i] = 0.0;
}

void cmfo_diag_avemax_adpstp(cmfo_model* m, double dt)
{
    m->nadd = m->nadd + dt;
    for (int i = 0; i < m->n; ++i) {
        m->rivout_avg[i] = m->rivout_avg[i] + m->rivout[i] * dt;
        m->fldout_avg[i] = m->fldout_avg[i] + m->fldout[i] * dt;
        m->rivvel_avg[i] = m->rivvel_avg[i] + m->rivvel[i] * dt;
        m->outflw_avg[i] = m->outflw_avg[i] + m->outflw[i] * dt;
        m->pthout_avg[i] = (m->pthout_avg[i] + m->pthout[i] * dt) - 0.0 * dt;   /* D2PTHINF = 0 */
        m->runoff_avg[i] = m->runoff_avg[i] + m->runoff[i] * dt;
        m->outflw_max[i] = DMAX(m->outflw_max[i], fabs(m->outflw[i]));
        m->rivdph_max[i] = DMAX(m->rivdph_max[i], m->rivdph[i]);
        m->storge_max[i] = DMAX(m->storge_max[i], m->storge[i]);
    }
    if (m->lpthout)
        for (size_t k = 0; k < (size_t)m->npth * m->npthlev; ++k)
            m->pthflw_avg[k] = m->pthflw_avg[k] + m->pthflw[k] * dt;
    if (m->ldamout && m->daminf_avg)                   /* diag:130-135, every cell */
        for (int i = 0; i < m->n; ++i) m->daminf_avg[i] = m->daminf_avg[i] + m->daminf[i] * dt;
}

void cmfo_diag_getave_adpstp(cmfo_model* m)
{
    const double a = m->nadd;
    for (int i = 0; i < m->n; ++i) {
        m->rivout_avg[i] = m->rivout_avg[i] / a;
        m->fldout_avg[i] = m->fldout_avg[i] / a;
        m->outflw_avg[i] = m->outflw_avg[i] / a;
        m->rivvel_avg[i] = m->rivvel_avg[i] / a;
        m->pthout_avg[i] = m->pthout_avg[i] / a;
        m->runoff_avg[i] = m->runoff_avg[i] / a;
    }
    if (m->ldamout && m->daminf_avg)                   /* diag:188-194 */
        for (int i = 0; i < m->n; ++i) m->daminf_avg[i] = m->daminf_avg[i] / a;
    if (m->lpthout) {
        const int np = m->npth;
        for (int l = 0; l < m->npthlev; ++l)
            for (int p = 0; p < np; ++p) {
                m->pthflw_avg[(size_t)l * np + p] = m->pthflw_avg[(size_t)l * np + p] / a;
                m->pthflwsum_avg[p] = m->pthflwsum_avg[p] + m->pthflw_avg[(size_t)l * np + p];
            }
    }
}

/* ------------------------------------------------------------------------
 * CMF_CALC_OUTFLW_KINE  src/cmf_opt_outflw_mod.F90:27-127 (LKINE): kinematic
 * wave with the bed slope (ELEVTN difference over NXTDST, at least PMINSLP;
 * mouths PMINSLP, the LSLOPEMOUTH = .FALSE. branch :86-91), outflow capped by
 * storage / DT, no reverse-flow limiter. RIVVEL is the kinematic velocity. */
void cmfo_calc_outflw_kine(cmfo_model* m, double dt)
{
    const int n = m->n, nr = m->nseqriv;
    for (int i = 0; i < n; ++i) m->sfcelv[i] = m->rivelv[i] + m->rivdph[i];   /* :44-48 */
    for (int i = 0; i < n; ++i) {
        double dslope;
        if (i < nr) {                                                       /* :53-78 */
            const int j = m->next[i];
            dslope = (m->elevtn[i] - m->elevtn[j]) * (1.0 / m->nxtdst[i]);   /* :57 */
            dslope = DMAX(dslope, m->pminslp);                              /* :58 */
        } else {                                                            /* :85-123 */
            dslope = m->pminslp;
        }
        const double dvel = ((1.0 / m->rivman[i]) * sqrt(dslope)) * pow(m->rivdph[i], F2O3);
        const double darea = m->rivwth[i] * m->rivdph[i];
        m->rivvel[i] = dvel;
        const double q = darea * dvel;
        m->rivout[i] = DMIN(q, m->rivsto[i] / dt);
        if (m->lfldout) {                                                   /* :66-75, :105-114 */
            const double dslope_f = DMIN(0.005, dslope);
            const double dvel_f = ((1.0 / m->pmanfld) * sqrt(dslope_f)) * pow(m->flddph[i], F2O3);
            double dare_f = m->fldsto[i] * (1.0 / m->rivlen[i]);
            dare_f = DMAX(dare_f - m->flddph[i] * m->rivwth[i], 0.0);
            const double qf = dare_f * dvel_f;
            m->fldout[i] = DMIN(qf, m->fldsto[i] / dt);
        }
    }
}

/* ------------------------------------------------------------------------
 * CMF_CALC_OUTFLW_KINEMIX  src/cmf_opt_outflw_mod.F90:118-300 (LSLPMIX): cells
 * with I2MASK = 0 take the local inertial equations, in the variant written
 * there; the others (I2MASK = 1: steep cells of the slope mask, dam-upstream
 * cells; 2: dams) the kinematic wave of CMF_CALC_OUTFLW_KINE. */
static void kinemix_fld(cmfo_model* m, int i, double dt, double dflw_f, double dflw_pre_f, double dslope_f,
                        double flddph_pre)
{
    const double gdt = m->pgrv * dt;
    double dare_f = m->fldsto[i] * (1.0 / m->rivlen[i]);                    /* :181-182, :258-259 */
    dare_f = DMAX(dare_f - m->flddph[i] * m->rivwth[i], 0.0);
    const double t = dflw_f * dflw_pre_f;
    double dflw_imp_f = sqrt(DMAX(t, 0.0));                                 /* :185, :262 */
    dflw_imp_f = DMAX(dflw_imp_f, 1.e-6);
    double dare_pre_f = m->fldsto_pre[i] * (1.0 / m->rivlen[i]);           /* :187-188, :264-265 */
    dare_pre_f = DMAX(dare_pre_f - flddph_pre * m->rivwth[i], 1.e-6);
    double dare_imp_f = sqrt(dare_f * dare_pre_f);                          /* :189, :266 */
    dare_imp_f = DMAX(dare_imp_f, 1.e-6);
    double q = 0.0;
    if (dflw_imp_f > F1EM5 && dare_imp_f > F1EM5) {                        /* :191-197, :268-274 */
        const double dout_pre_f = m->fldout_pre[i];
        q = (dout_pre_f + (gdt * dare_imp_f) * dslope_f) *
            (1.0 / (1.0 + (((gdt * (m->pmanfld * m->pmanfld)) * fabs(dout_pre_f)) * pow(dflw_imp_f, P4O3F)) *
                              (1.0 / dare_imp_f)));
    }
    if (q * m->rivout[i] < 0.0) q = 0.0;                                    /* :199, :276 */
    m->fldout[i] = q;
}

static void kinemix_riv(cmfo_model* m, int i, double dt, double dflw, double dflw_pre, double dslope)
{
    const double gdt = m->pgrv * dt;
    const double darea = m->rivwth[i] * dflw;                               /* :159, :236 */
    double dflw_imp = sqrt(dflw * dflw_pre);                                /* :162, :239 */
    dflw_imp = DMAX(dflw_imp, 1.e-6);
    if (dflw_imp > F1EM5 && darea > F1EM5) {                                /* :164-172, :241-249 */
        const double dout_pre = m->rivout_pre[i] * (1.0 / m->rivwth[i]);
        const double man2 = m->rivman[i] * m->rivman[i];
        m->rivout[i] = (m->rivwth[i] * (dout_pre + (gdt * dflw_imp) * dslope)) *
                       (1.0 / (1.0 + ((gdt * man2) * fabs(dout_pre)) * pow(dflw_imp, P7O3F)));
        m->rivvel[i] = m->rivout[i] * (1.0 / darea);
    } else {
        m->rivout[i] = 0.0;
        m->rivvel[i] = 0.0;
    }
}

static void kinemix_kine(cmfo_model* m, int i, double dt, double dslope)
{
    const double dvel = ((1.0 / m->rivman[i]) * sqrt(dslope)) * pow(m->rivdph[i], F2O3);   /* :206-207 */
    const double darea = m->rivwth[i] * m->rivdph[i];
    m->rivvel[i] = dvel;
    const double q = darea * dvel;
    m->rivout[i] = DMIN(q, m->rivsto[i] / dt);
    if (m->lfldout) {                                                       /* :214-222 */
        const double dslope_f = DMIN(0.005, dslope);
        const double dvel_f = ((1.0 / m->pmanfld) * sqrt(dslope_f)) * pow(m->flddph[i], F2O3);
        double dare_f = m->fldsto[i] * (1.0 / m->rivlen[i]);
        dare_f = DMAX(dare_f - m->flddph[i] * m->rivwth[i], 0.0);
        const double qf = dare_f * dvel_f;
        m->fldout[i] = DMIN(qf, m->fldsto[i] / dt);
    }
}

void cmfo_calc_outflw_kinemix(cmfo_model* m, double dt)
{
    const int n = m->n, nr = m->nseqriv;
    double* sp = m->w_sfcelv_pre;
    double* fdp = m->w_flddph_pre;
    for (int i = 0; i < n; ++i) {                                           /* :141-146 */
        m->sfcelv[i] = m->rivelv[i] + m->rivdph[i];
        sp[i] = m->rivelv[i] + m->rivdph_pre[i];
        const double d = m->rivdph_pre[i] - m->rivhgt[i];
        fdp[i] = DMAX(d, 0.0);
    }
    for (int i = 0; i < nr; ++i) {                                          /* :149-226 */
        const int j = m->next[i];
        if (m->i2mask[i] == 0) {
            const double dsfcmax = DMAX(m->sfcelv[i], m->sfcelv[j]);
            const double dsfcmax_pre = DMAX(sp[i], sp[j]);
            const double dslope = (m->sfcelv[i] - m->sfcelv[j]) * (1.0 / m->nxtdst[i]);
            const double dslope_f = DMAX(-0.005, DMIN(0.005, dslope));
            kinemix_riv(m, i, dt, dsfcmax - m->rivelv[i], dsfcmax_pre - m->rivelv[i], dslope);
            if (m->lfldout)
                kinemix_fld(m, i, dt, DMAX(dsfcmax - m->elevtn[i], 0.0), dsfcmax_pre - m->elevtn[i], dslope_f,
                            fdp[i]);
        } else {
            double dslope = (m->elevtn[i] - m->elevtn[j]) * (1.0 / m->nxtdst[i]);   /* :203-204 */
            dslope = DMAX(dslope, m->pminslp);
            kinemix_kine(m, i, dt, dslope);
        }
    }
    for (int i = nr; i < n; ++i) {                                          /* :229-297 */
        if (m->i2mask[i] == 0) {
            const double dslope = (m->sfcelv[i] - m->dwnelv[i]) * (1.0 / m->pdstmth);
            const double dslope_f = DMAX(-0.005, DMIN(0.005, dslope));
            kinemix_riv(m, i, dt, m->rivdph[i], m->rivdph_pre[i], dslope);
            if (m->lfldout)
                kinemix_fld(m, i, dt, m->sfcelv[i] - m->elevtn[i], sp[i] - m->elevtn[i], dslope_f, fdp[i]);
        } else {
            kinemix_kine(m, i, dt, m->pminslp);                            /* :280-281 */
        }
    }
}

/* ------------------------------------------------------------------------
 * CMF_CALC_OUTPRE  src/cmf_opt_outflw_mod.F90:308-421: previous-step flows for a
 * storage-only restart (called after CMF_PHYSICS_FLDSTG, drv_control:273-278):
 * the diffusive-wave (Manning) discharge of the current water surface, signed by
 * the surface slope, for rivers, mouths (slope to DWNELV over PDSTMTH) and
 * bifurcation levels. Writes D2SFCELV, D2RIVDPH_PRE, D2RIVOUT_PRE, D2FLDOUT_PRE
 * and D1PTHFLW_PRE. A path with an endpoint outside the domain (PTH_UPST or
 * PTH_DOWN = 0) gets 0: the reference indexes D2SFCELV(0) there. */
static double outpre_q(double area, double rman_inv, double dflw, double dslope)
{
    /* AREA * ( MAN**(-1.) * DFLW**(2./3.) * abs(DSLOPE)**(0.5) ), signed by DSLOPE */
    const double q = area * ((rman_inv * pow(dflw, F2O3)) * sqrt(fabs(dslope)));
    return dslope < 0.0 ? -q : q;
}

void cmfo_calc_outpre(cmfo_model* m)
{
    const int n = m->n, nr = m->nseqriv;
    for (int i = 0; i < n; ++i) {                                           /* :324-327 */
        m->sfcelv[i] = m->rivelv[i] + m->rivdph[i];
        m->rivdph_pre[i] = m->rivdph[i];
    }
    for (int i = 0; i < n; ++i) {
        double dslope, dflw, dflw_f;
        if (i < nr) {                                                       /* :331-363 */
            const int j = m->next[i];
            const double dsfcmax = DMAX(m->sfcelv[i], m->sfcelv[j]);
            dslope = (m->sfcelv[i] - m->sfcelv[j]) * (1.0 / m->nxtdst[i]);
            dflw = dsfcmax - m->rivelv[i];
            dflw_f = DMAX(dsfcmax - m->elevtn[i], 0.0);
        } else {                                                            /* :368-393 */
            dslope = (m->sfcelv[i] - m->dwnelv[i]) * (1.0 / m->pdstmth);
            dflw = m->rivdph[i];
            dflw_f = DMAX(m->sfcelv[i] - m->elevtn[i], 0.0);
        }
        const double dslope_f = DMAX(-0.005, DMIN(0.005, dslope));
        const double darea = m->rivwth[i] * dflw;
        m->rivout_pre[i] = darea > F1EM5 ? outpre_q(darea, 1.0 / m->rivman[i], dflw, dslope) : 0.0;
        double dare_f = m->fldsto[i] * (1.0 / m->rivlen[i]);
        dare_f = DMAX(dare_f - m->flddph[i] * m->rivwth[i], 0.0);
        m->fldout_pre[i] = dare_f > F1EM5 ? outpre_q(dare_f, 1.0 / m->pmanfld, dflw_f, dslope_f) : 0.0;
        if (!m->lfldout) m->fldout_pre[i] = 0.0;                            /* :416-418 */
    }
    if (m->lpthout) {                                                       /* :396-413 */
        const int np = m->npth;
        for (int p = 0; p < np; ++p) {
            const int ip = m->pth_upst[p], jp = m->pth_down[p];
            const int ok = ip >= 0 && jp >= 0;
            const double dslope = ok ? (m->sfcelv[ip] - m->sfcelv[jp]) * (1.0 / m->pth_dst[p]) : 0.0;
            for (int l = 0; l < m->npthlev; ++l) {
                double q = 0.0;
                if (ok) {
                    double dflw = DMAX(m->sfcelv[ip], m->sfcelv[jp]) - m->pth_elv[(size_t)l * np + p];
                    dflw = DMAX(dflw, 0.0);
                    if (dflw > F1EM5) {
                        q = (m->pth_wth[(size_t)l * np + p] * dflw) *
                            (((1.0 / m->pth_man[l]) * pow(dflw, F2O3)) * sqrt(fabs(dslope)));
                        if (dslope < 0.0) q = -q;
                    }
                }
                m->pthflw_pre[(size_t)l * np + p] = q;
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * CMF_PHYSICS_ADVANCE  src/cmf_ctrl_physics_mod.F90:21-155 */
int32_t cmfo_physics_advance(cmfo_model* m, double* dt, int32_t nt_cap)
{
    const double dt_def = *dt;
    cmfo_physics_fldstg(m);                                                 /* :54 */
    int32_t nt = 1;
    if (m->ladpstp) nt = cmfo_calc_adpstp(m, dt_def, dt);                   /* :58 */
    cmfo_diag_reset_adpstp(m);                                              /* :61 */
    const int32_t nrun = (nt_cap > 0 && nt_cap < nt) ? nt_cap : nt;
    for (int32_t it = 0; it < nrun; ++it) {                                 /* :64-144 */
        if (m->lkine) cmfo_calc_outflw_kine(m, *dt);                       /* :66-72 */
        else if (m->lslpmix) cmfo_calc_outflw_kinemix(m, *dt);
        else cmfo_calc_outflw(m, *dt);
        if (!m->lfldout)
            for (int i = 0; i < m->n; ++i) { m->fldout[i] = 0.0; m->fldout_pre[i] = 0.0; }
        if (m->ldamout) cmfo_damout_calc(m, *dt);
        if (m->lpthout) {
            if (m->llevee) cmfo_levee_opt_pthout(m, *dt);
            else cmfo_calc_pthout(m, *dt);
        }
        cmfo_calc_inflow_lspamat(m, *dt);
        if (m->ldamout) cmfo_damout_watbal(m, *dt);
        cmfo_calc_vars_pre(m);
        cmfo_calc_stonxt(m, *dt);
        cmfo_physics_fldstg(m);
        cmfo_diag_avemax_adpstp(m, *dt);
        if (m->glb_hist && it < m->glb_cap)                                 /* :216-248 */
            for (int k = 0; k < 11; ++k) m->glb_hist[(size_t)it * 11 + k] = m->glb[k];
    }
    *dt = dt_def;                                                           /* :145 */
    cmfo_diag_getave_adpstp(m);                                             /* :147 */
    return nt;
}

/* ------------------------------------------------------------------------
 * ROFF_INTERP  src/cmf_ctrl_forcing_mod.F90:757-790 (inside CMF_FORCING_PUT) */
void cmfo_roff_interp(int32_t n, int32_t inpn, int32_t nxin, int32_t nyin, const int32_t* inpx,
                      const int32_t* inpy, const double* inpa, const double* grid, double drofunit,
                      double rmis, double* out)
{
    for (int32_t i = 0; i < n; ++i) {                                       /* :770 */
        out[i] = 0.0;                                                       /* :771 */
        for (int32_t k = 0; k < inpn; ++k) {                                /* :772 */
            const int32_t ix = inpx[(size_t)k * n + i];                     /* :773 */
            const int32_t iy = inpy[(size_t)k * n + i];                     /* :774 */
            if (ix > 0) {                                                   /* :775 */
                if (ix > nxin || iy > nyin) continue;                       /* :776-780 */
                const double v = grid[(size_t)(iy - 1) * nxin + (ix - 1)];
                if (v != rmis)                                              /* :782 */
                    out[i] = out[i] + (v * inpa[(size_t)k * n + i]) / drofunit;   /* :783 */
            }
        }
    }
}

/* ------------------------------------------------------------------------
 * CMF_DIAG_RESET_OUTPUT  src/cmf_calc_diag_mod.F90:222-267 */
void cmfo_diag_reset_output(cmfo_output* o, int32_t n, int32_t npthl, int32_t ldamout)
{
    o->nadd = 0.0;                                                          /* :229 */
    for (int32_t i = 0; i < n; ++i) {                                       /* :231-245 */
        o->rivout[i] = 0.0; o->fldout[i] = 0.0; o->outflw[i] = 0.0; o->rivvel[i] = 0.0;
        o->pthout[i] = 0.0; o->runoff[i] = 0.0;
        o->storge_max[i] = 0.0; o->outflw_max[i] = 0.0; o->rivdph_max[i] = 0.0;
    }
    if (ldamout)                                                            /* :248-254 */
        for (int32_t i = 0; i < n; ++i) o->daminf[i] = 0.0;
    for (int32_t k = 0; k < npthl; ++k) o->pthflw[k] = 0.0;                 /* :262-265 */
}

/* ------------------------------------------------------------------------
 * CMF_DIAG_AVEMAX_OUTPUT  src/cmf_calc_diag_mod.F90:273-317 */
void cmfo_diag_avemax_output(cmfo_output* o, const cmfo_output* a, int32_t n, int32_t npthl, int32_t ldamout,
                             double dt)
{
    o->nadd = o->nadd + dt;                                                 /* :277 */
    for (int32_t i = 0; i < n; ++i) {                                       /* :280-299 */
        o->rivout[i] = o->rivout[i] + a->rivout[i] * dt;
        o->fldout[i] = o->fldout[i] + a->fldout[i] * dt;
        o->rivvel[i] = o->rivvel[i] + a->rivvel[i] * dt;
        o->outflw[i] = o->outflw[i] + a->outflw[i] * dt;
        o->pthout[i] = o->pthout[i] + a->pthout[i] * dt;
        o->runoff[i] = o->runoff[i] + a->runoff[i] * dt;
        o->outflw_max[i] = DMAX(o->outflw_max[i], fabs(a->outflw_max[i]));
        o->rivdph_max[i] = DMAX(o->rivdph_max[i], a->rivdph_max[i]);
        o->storge_max[i] = DMAX(o->storge_max[i], a->storge_max[i]);
    }
    if (ldamout)                                                            /* :303-309 */
        for (int32_t i = 0; i < n; ++i) o->daminf[i] = o->daminf[i] + a->daminf[i] * dt;
    for (int32_t k = 0; k < npthl; ++k) o->pthflw[k] = o->pthflw[k] + a->pthflw[k] * dt;   /* :311-315 */
}

/* ------------------------------------------------------------------------
 * CMF_DIAG_GETAVE_OUTPUT  src/cmf_calc_diag_mod.F90:323-356 */
void cmfo_diag_getave_output(cmfo_output* o, int32_t n, int32_t npthl, int32_t ldamout)
{
    const double d = o->nadd;
    for (int32_t i = 0; i < n; ++i) {                                       /* :331-348 */
        o->rivout[i] = o->rivout[i] / d;
        o->fldout[i] = o->fldout[i] / d;
        o->outflw[i] = o->outflw[i] / d;
        o->rivvel[i] = o->rivvel[i] / d;
        o->pthout[i] = o->pthout[i] / d;
        o->runoff[i] = o->runoff[i] / d;
        if (ldamout) o->daminf[i] = o->daminf[i] / d;
    }
    for (int32_t k = 0; k < npthl; ++k) o->pthflw[k] = o->pthflw[k] / d;   /* :351-354 */
}
