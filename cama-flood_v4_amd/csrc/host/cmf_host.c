/*
 * cmf_host.c -- host-side map initialisation (see include/cmf_host.h).
 *
 * Arithmetic follows the reference statement by statement, in FP64, with the
 * same operation order as the Fortran source (Fortran evaluates a*b*c as (a*b)*c)
 * and with x**(-1.) written as 1/x, which is how ROCm flang lowers it. Built with
 * -ffp-contract=off so no multiply-add is fused.
 */
#include "cmf_host.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* ---------------------------------------------------------------------------
 * CALC_1D_SEQ  (src/cmf_ctrl_maps_mod.F90:403-525)
 * Grid scans are IY outer / IX inner, as in the reference DO loops. */
int cmfh_calc_1d_seq(int nx, int ny, const int32_t* nextx, const int32_t* nexty,
                     int32_t* seqx, int32_t* seqy, int32_t* next, int32_t* nseqriv)
{
    if (nx <= 0 || ny <= 0 || !nextx || !nexty || !seqx || !seqy || !next || !nseqriv) return -1;
    const size_t ng = (size_t)nx * ny;
    int32_t* nupst = (int32_t*)calloc(ng, sizeof(int32_t));
    int32_t* upnow = (int32_t*)calloc(ng, sizeof(int32_t));
    int32_t* vec   = (int32_t*)calloc(ng, sizeof(int32_t));   /* I2VECTOR, 1-based */
    if (!nupst || !upnow || !vec) { free(nupst); free(upnow); free(vec); return -2; }

    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = (size_t)iy * nx + ix;
            if (nextx[k] > 0) {
                size_t kk = (size_t)(nexty[k] - 1) * nx + (nextx[k] - 1);
                nupst[kk] += 1;
            }
        }
    int32_t iseq = 0;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = (size_t)iy * nx + ix;
            if (nextx[k] > 0 && nupst[k] == upnow[k]) {
                seqx[iseq] = ix + 1; seqy[iseq] = iy + 1;
                iseq += 1; vec[k] = iseq;
            }
        }
    int32_t iseq1 = 1, iseq2 = iseq, jseq = iseq;
    int again = 1;
    while (again) {
        again = 0;
        jseq = iseq2;
        for (int32_t s = iseq1; s <= iseq2; ++s) {
            int ix = seqx[s - 1] - 1, iy = seqy[s - 1] - 1;
            size_t k = (size_t)iy * nx + ix;
            int jx = nextx[k] - 1, jy = nexty[k] - 1;
            size_t kk = (size_t)jy * nx + jx;
            upnow[kk] += 1;
            if (upnow[kk] == nupst[kk] && nextx[kk] > 0) {
                seqx[jseq] = jx + 1; seqy[jseq] = jy + 1;
                jseq += 1; vec[kk] = jseq;
                again = 1;
            }
        }
        iseq1 = iseq2 + 1;
        iseq2 = jseq;
    }
    *nseqriv = jseq;
    iseq = jseq;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = (size_t)iy * nx + ix;
            if (nextx[k] < 0 && nextx[k] != CMF_NEXT_NONE) {
                seqx[iseq] = ix + 1; seqy[iseq] = iy + 1;
                iseq += 1; vec[k] = iseq;
            }
        }
    const int32_t nall = iseq;
    for (int32_t s = 0; s < nall; ++s) {
        size_t k = (size_t)(seqy[s] - 1) * nx + (seqx[s] - 1);
        if (nextx[k] > 0) {
            size_t kk = (size_t)(nexty[k] - 1) * nx + (nextx[k] - 1);
            next[s] = vec[kk] - 1;
        } else {
            next[s] = nextx[k];
        }
    }
    free(nupst); free(upnow); free(vec);
    return nall;
}

/* Upstream CSR: for each receiving cell, its upstream cells in ascending order.
 * Same order as the I1UPST fill at src/cmf_ctrl_maps_mod.F90:518-522. */
int cmfh_build_upstream_csr(int n, int nseqriv, const int32_t* next,
                            int32_t* up_ptr, int32_t* up_idx)
{
    if (n < 0 || nseqriv < 0 || nseqriv > n || !next || !up_ptr || !up_idx) return -1;
    memset(up_ptr, 0, sizeof(int32_t) * (size_t)(n + 1));
    for (int i = 0; i < nseqriv; ++i) {
        int j = next[i];
        if (j < 0 || j >= n) return -3;
        up_ptr[j + 1] += 1;
    }
    for (int i = 0; i < n; ++i) up_ptr[i + 1] += up_ptr[i];
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!fill) return -2;
    for (int i = 0; i < n; ++i) fill[i] = up_ptr[i];
    for (int i = 0; i < nseqriv; ++i) {          /* ascending source order */
        int j = next[i];
        up_idx[fill[j]++] = i;
    }
    free(fill);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Channel tables (src/cmf_ctrl_maps_mod.F90:716-722) and SET_FLDSTG (:969-999). */
int cmfh_set_fldstg(int n, int nlfp, int lfplain,
                    const double* grarea, const double* rivlen, const double* rivwth,
                    const double* rivhgt, const double* elevtn, const double* fldhgt,
                    double* rivstomax, double* rivelv, double* fldstomax, double* fldgrd,
                    double* dfrcinc_out)
{
    if (n < 0 || nlfp <= 0) return -1;
    for (int i = 0; i < n; ++i) {
        rivstomax[i] = lfplain ? (rivlen[i] * rivwth[i]) * rivhgt[i] : (double)1.e18f;
        rivelv[i] = elevtn[i] - rivhgt[i];
    }
    const double dfrcinc = 1.0 / (double)nlfp;          /* REAL(NLFP)**(-1.) */
    if (dfrcinc_out) *dfrcinc_out = dfrcinc;
    for (int i = 0; i < n; ++i) {
        double dstopre = rivstomax[i];
        double dhgtpre = 0.0;
        double dwthinc = (grarea[i] * (1.0 / rivlen[i])) * dfrcinc;
        for (int l = 0; l < nlfp; ++l) {
            double h = fldhgt[(size_t)l * n + i];
            double dstonow = (rivlen[i] * (rivwth[i] + dwthinc * ((double)(l + 1) - 0.5))) * (h - dhgtpre);
            fldstomax[(size_t)l * n + i] = dstopre + dstonow;
            fldgrd[(size_t)l * n + i] = (h - dhgtpre) * (1.0 / dwthinc);
            dstopre = fldstomax[(size_t)l * n + i];
            dhgtpre = h;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * CMF_LEVEE_INIT section [2]  (src/cmf_ctrl_levee_mod.F90:148-261) */
int cmfh_levee_init(int n, int nlfp,
                    const double* grarea, const double* rivlen, const double* rivwth,
                    const double* rivstomax, const double* fldhgt,
                    double* fldstomax, double* fldgrd,
                    double* levhgt, double* levfrc,
                    double* bashgt, double* levdst, double* levbassto,
                    double* levtopsto, double* levfilsto)
{
    if (n < 0 || nlfp <= 0) return -1;
    const double dfrcinc = 1.0 / (double)nlfp;
    for (int i = 0; i < n; ++i) {                                  /* :155-166 */
        if (levhgt[i] <= 0.0) { levhgt[i] = 0.0; levfrc[i] = 1.0; }
        if (levfrc[i] >= 1.0 || levfrc[i] < 0.0) { levhgt[i] = 0.0; levfrc[i] = 1.0; }
        levfrc[i] = dmax(0.0, dmin(1.0, levfrc[i]));
    }
#define FSM(l) fldstomax[(size_t)(l) * n + i]
#define FGR(l) fldgrd[(size_t)(l) * n + i]
#define FHG(l) fldhgt[(size_t)(l) * n + i]
    for (int i = 0; i < n; ++i) {
        double dsto_fil = rivstomax[i];                              /* :172-181 */
        double dhgtpre = 0.0;
        const double dwth_inc = (grarea[i] * (1.0 / rivlen[i])) * dfrcinc;
        for (int l = 0; l < nlfp; ++l) {
            double dsto_add = (rivlen[i] * (rivwth[i] + dwth_inc * ((double)(l + 1) - 0.5))) * (FHG(l) - dhgtpre);
            FSM(l) = dsto_fil + dsto_add;
            FGR(l) = (FHG(l) - dhgtpre) * (1.0 / dwth_inc);
            dsto_fil = FSM(l);
            dhgtpre = FHG(l);
        }
        if (levhgt[i] <= 0.0) {                                      /* :184-189 */
            bashgt[i] = 1.e18; levdst[i] = 1.e18; levbassto[i] = 1.e18;
            levtopsto[i] = 1.e18; levfilsto[i] = 1.e18;
            continue;
        }
        /* [1] levee base and top storage (:194-227) */
        dsto_fil = rivstomax[i];
        dhgtpre = 0.0;
        double dwth_fil = 0.0;
        levdst[i] = (levfrc[i] * dwth_inc) * (double)nlfp;
        int ilev = (int)(levfrc[i] * (double)nlfp) + 1;             /* 1-based layer */
        if (ilev >= 2) {
            dsto_fil = FSM(ilev - 2);
            dhgtpre = FHG(ilev - 2);
            dwth_fil = dwth_inc * (double)(ilev - 1);
        }
        if (ilev <= nlfp) {
            double dwth_add = levdst[i] - dwth_fil;
            double dhgtnow = dwth_add * FGR(ilev - 1);
            bashgt[i] = dhgtnow + dhgtpre;
            levhgt[i] = dmax(levhgt[i], bashgt[i]);
            double dsto_add = (((dwth_add * 0.5 + dwth_fil) + rivwth[i]) * dhgtnow) * rivlen[i];
            levbassto[i] = dsto_fil + dsto_add;
            double dhgtdif = levhgt[i] - bashgt[i];
            levtopsto[i] = levbassto[i] + ((levdst[i] + rivwth[i]) * dhgtdif) * rivlen[i];
        } else {
            bashgt[i] = dhgtpre;
            levhgt[i] = dmax(levhgt[i], bashgt[i]);
            levbassto[i] = dsto_fil;
            double dhgtdif = levhgt[i] - bashgt[i];
            levtopsto[i] = levbassto[i] + ((levdst[i] + rivwth[i]) * dhgtdif) * rivlen[i];
        }
        /* [2] levee fill storage (:231-258) */
        int I = 1;
        dsto_fil = rivstomax[i];
        dwth_fil = rivwth[i];
        dhgtpre = 0.0;
        while (I <= nlfp && levhgt[i] > FHG(I - 1)) {
            dsto_fil = FSM(I - 1);
            dwth_fil = dwth_fil + dwth_inc;
            dhgtpre = FHG(I - 1);
            I = I + 1;
        }
        if (I <= nlfp) {
            double dhgtnow = levhgt[i] - dhgtpre;
            double dwth_add = dhgtnow / FGR(I - 1);
            double dsto_add = ((dwth_add * 0.5 + dwth_fil) * dhgtnow) * rivlen[i];
            levfilsto[i] = dsto_fil + dsto_add;
        } else {
            double dhgtnow = levhgt[i] - dhgtpre;
            double dsto_add = (dwth_fil * dhgtnow) * rivlen[i];
            levfilsto[i] = dsto_fil + dsto_add;
        }
    }
#undef FSM
#undef FGR
#undef FHG
    return 0;
}

/* ---------------------------------------------------------------------------
 * STORAGE_SEA_SURFACE  (src/cmf_ctrl_vars_mod.F90:129-168) */
int cmfh_storage_sea_surface(int n, int nseqriv, const int32_t* next,
                             const double* dwnelv, const double* rivelv, const double* rivhgt,
                             const double* rivlen, const double* rivwth, const double* rivstomax,
                             double* rivsto, double* rivdph_pre)
{
    if (n < 0 || nseqriv < 0 || nseqriv > n) return -1;
    for (int i = nseqriv; i < n; ++i) {
        double dseaelv = dwnelv[i];
        double ddph = dmax(dseaelv - rivelv[i], 0.0);
        ddph = dmin(ddph, rivhgt[i]);
        double s = (ddph * rivlen[i]) * rivwth[i];
        rivsto[i] = dmin(s, rivstomax[i] * 1.0);
        rivdph_pre[i] = ddph;
    }
    for (int i = nseqriv - 1; i >= 0; --i) {
        int j = next[i];
        double dseaelv = rivelv[j] + rivdph_pre[j];
        double ddph = dmax(dseaelv - rivelv[i], 0.0);
        ddph = dmin(ddph, rivhgt[i]);
        double s = (ddph * rivlen[i]) * rivwth[i];
        rivsto[i] = dmin(s, rivstomax[i] * 1.0);
        rivdph_pre[i] = ddph;
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Priority-flood drainage for synthetic networks (Barnes, Lehman & Mulla 2014). */
typedef struct { float z; uint32_t ord; uint32_t k; } pf_node;

static inline int pf_less(const pf_node* a, const pf_node* b)
{
    return a->z < b->z || (a->z == b->z && a->ord < b->ord);
}
static void pf_push(pf_node* h, size_t* sz, pf_node v)
{
    size_t i = (*sz)++;
    h[i] = v;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (!pf_less(&h[i], &h[p])) break;
        pf_node t = h[i]; h[i] = h[p]; h[p] = t; i = p;
    }
}
static pf_node pf_pop(pf_node* h, size_t* sz)
{
    pf_node top = h[0];
    h[0] = h[--(*sz)];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < *sz && pf_less(&h[l], &h[m])) m = l;
        if (r < *sz && pf_less(&h[r], &h[m])) m = r;
        if (m == i) break;
        pf_node t = h[i]; h[i] = h[m]; h[m] = t; i = m;
    }
    return top;
}

int cmfh_priority_flood(int nx, int ny, const float* elev, const uint8_t* land, float eps,
                        float* filled, int32_t* nextx, int32_t* nexty)
{
    if (nx <= 2 || ny <= 0) return -1;
    const size_t ng = (size_t)nx * ny;
    uint8_t* done = (uint8_t*)calloc(ng, 1);
    pf_node* heap = (pf_node*)malloc(sizeof(pf_node) * ng);
    if (!done || !heap) { free(done); free(heap); return -2; }
    size_t sz = 0;
    uint32_t ord = 0;
    static const int DX[8] = {1, 1, 0, -1, -1, -1, 0, 1};
    static const int DY[8] = {0, 1, 1, 1, 0, -1, -1, -1};
    for (size_t k = 0; k < ng; ++k) {
        nextx[k] = CMF_NEXT_NONE; nexty[k] = CMF_NEXT_NONE;
        filled[k] = elev[k];
    }
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = (size_t)iy * nx + ix;
            if (!land[k]) continue;
            int coast = 0;
            for (int d = 0; d < 8 && !coast; ++d) {
                int jy = iy + DY[d];
                if (jy < 0 || jy >= ny) { coast = 1; break; }
                int jx = (ix + DX[d] + nx) % nx;
                if (!land[(size_t)jy * nx + jx]) coast = 1;
            }
            if (coast) {
                nextx[k] = CMF_NEXT_MOUTH; nexty[k] = CMF_NEXT_MOUTH;
                done[k] = 1;
                pf_node v = {filled[k], ord++, (uint32_t)k};
                pf_push(heap, &sz, v);
            }
        }
    while (sz > 0) {
        pf_node c = pf_pop(heap, &sz);
        int cx = (int)(c.k % nx), cy = (int)(c.k / nx);
        for (int d = 0; d < 8; ++d) {
            int jy = cy + DY[d];
            if (jy < 0 || jy >= ny) continue;
            int jx = (cx + DX[d] + nx) % nx;
            size_t kk = (size_t)jy * nx + jx;
            if (!land[kk] || done[kk]) continue;
            done[kk] = 1;
            float zmin = filled[c.k] + eps;
            if (filled[kk] < zmin) filled[kk] = zmin;
            nextx[kk] = cx + 1; nexty[kk] = cy + 1;
            pf_node v = {filled[kk], ord++, (uint32_t)kk};
            pf_push(heap, &sz, v);
        }
    }
    free(done); free(heap);
    return 0;
}

int cmfh_upstream_area(int nx, int ny, const int32_t* nextx, const int32_t* nexty,
                       const double* cellarea, double* uparea)
{
    const size_t ng = (size_t)nx * ny;
    int32_t* indeg = (int32_t*)calloc(ng, sizeof(int32_t));
    size_t* queue = (size_t*)malloc(sizeof(size_t) * ng);
    if (!indeg || !queue) { free(indeg); free(queue); return -2; }
    for (size_t k = 0; k < ng; ++k) uparea[k] = (nextx[k] != CMF_NEXT_NONE) ? cellarea[k] : 0.0;
    for (size_t k = 0; k < ng; ++k)
        if (nextx[k] > 0) indeg[(size_t)(nexty[k] - 1) * nx + (nextx[k] - 1)] += 1;
    size_t qh = 0, qt = 0;
    for (size_t k = 0; k < ng; ++k)
        if (nextx[k] != CMF_NEXT_NONE && indeg[k] == 0) queue[qt++] = k;
    while (qh < qt) {
        size_t k = queue[qh++];
        if (nextx[k] > 0) {
            size_t kk = (size_t)(nexty[k] - 1) * nx + (nextx[k] - 1);
            uparea[kk] += uparea[k];
            if (--indeg[kk] == 0) queue[qt++] = kk;
        }
    }
    free(indeg); free(queue);
    return 0;
}

int cmfh_accumulate(int n, const int32_t* next, int nrec, const double* val, double* out)
{
    /* Kahn order over the vector network (next < 0 = mouth), so any cell order works */
    int32_t* indeg = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!indeg || !order) { free(indeg); free(order); return -2; }
    for (int i = 0; i < n; ++i) {
        if (next[i] >= n) { free(indeg); free(order); return -1; }
        if (next[i] >= 0) indeg[next[i]] += 1;
    }
    int qh = 0, qt = 0;
    for (int i = 0; i < n; ++i) if (indeg[i] == 0) order[qt++] = i;
    while (qh < qt) {
        const int i = order[qh++];
        if (next[i] >= 0 && --indeg[next[i]] == 0) order[qt++] = next[i];
    }
    if (qt != n) { free(indeg); free(order); return -3; }      /* not a forest */
    for (int r = 0; r < nrec; ++r) {
        const double* v = val + (size_t)r * n;
        double* o = out + (size_t)r * n;
        for (int i = 0; i < n; ++i) o[i] = v[i];
        for (int k = 0; k < n; ++k) {
            const int i = order[k];
            if (next[i] >= 0) o[next[i]] += o[i];
        }
    }
    free(indeg); free(order);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Steady state of the local-inertial outflow under constant flows (no reference
 * counterpart; a physical prior for the spin-up surrogate).
 *
 * With RIVOUT = RIVOUT_PRE and equal current/previous depths, the river update of
 * src/cmf_calc_outflw_mod.F90:60-80 is a fixed point exactly when
 *   Q = W h^(5/3) S^(1/2) / n        (h = flow depth, S = water-surface slope)
 * and the floodplain update (:84-116) when
 *   Qf = A hf^(2/3) min(S, 0.005)^(1/2) / n_fld
 * with A = FLDSTO/RIVLEN - FLDDPH*RIVWTH and hf the depth above ELEVTN. Cells are
 * visited downstream first; each takes the surface elevation at which its outflow
 * equals q[i], found by bisection (the outflow grows monotonically with it), given
 * the downstream cell's elevation (DWNELV at a mouth, slope over PDSTMTH). */
static double steady_storage(const cmfh_channel* c, int i, double d, double* fldsto, double* flddph)
{
    const int n = c->n;
    const double rivlen = c->rivlen[i], rivwth = c->rivwth[i];
    const double rivstomax = c->rivstomax[i];
    if (d <= c->rivhgt[i]) { *fldsto = 0.0; *flddph = 0.0; return d * rivlen * rivwth; }
    const double f = d - c->rivhgt[i];
    const double dwth_inc = (c->grarea[i] / rivlen) * c->dfrcinc;
    double dsto_fil = rivstomax, dwth_fil = rivwth, ddph_fil = 0.0, sto;
    int l = 0;
    for (; l < c->nlfp; ++l) {
        const double grd = c->fldgrd[(size_t)l * n + i];
        const double top = ddph_fil + grd * dwth_inc;
        if (f <= top && grd > 0.0) break;
        dsto_fil = c->fldstomax[(size_t)l * n + i];
        dwth_fil += dwth_inc;
        ddph_fil = top;
    }
    if (l == c->nlfp) {
        sto = dsto_fil + (f - ddph_fil) * dwth_fil * rivlen;
    } else {
        const double grd = c->fldgrd[(size_t)l * n + i];
        const double dwth_add = (f - ddph_fil) / grd;
        sto = dsto_fil + rivlen * grd * (dwth_fil * dwth_add + 0.5 * dwth_add * dwth_add);
    }
    const double rs = rivstomax + rivlen * rivwth * f;
    *fldsto = dmax(sto - rs, 0.0);
    *flddph = f;
    return sto;
}

static double steady_flow(const cmfh_channel* c, int i, double s, double s_down, double* qfld)
{
    const double d = s - c->rivelv[i];
    *qfld = 0.0;
    if (d <= 0.0 || s <= s_down) return 0.0;
    const int mouth = c->next[i] < 0;
    const double slope = (s - s_down) / (mouth ? c->pdstmth : c->nxtdst[i]);
    const double h = d;                          /* s > s_down: flow depth above RIVELV */
    double q = c->rivwth[i] * pow(h, 5.0 / 3.0) * sqrt(slope) / c->rivman[i];
    if (c->lfldout) {
        double fsto, fdph;
        (void)steady_storage(c, i, d, &fsto, &fdph);
        const double a = dmax(fsto / c->rivlen[i] - fdph * c->rivwth[i], 0.0);
        const double hf = dmax(s - c->elevtn[i], 0.0);
        if (a > 1e-5 && hf > 1e-5) *qfld = a * pow(hf, 2.0 / 3.0) * sqrt(dmin(slope, 0.005)) / c->pmanfld;
    }
    return q + *qfld;
}

int cmfh_steady_state(const cmfh_channel* c, const double* q, double* rivsto, double* fldsto, double* sfcelv,
                      double* rivout, double* fldout)
{
    if (!c || !q || !rivsto || !fldsto || !sfcelv || c->n <= 0 || c->nlfp <= 0) return -1;
    const int n = c->n;
    int32_t* indeg = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!indeg || !order) { free(indeg); free(order); return -2; }
    for (int i = 0; i < n; ++i) {
        if (c->next[i] >= n) { free(indeg); free(order); return -1; }
        if (c->next[i] >= 0) indeg[c->next[i]] += 1;
    }
    int qh = 0, qt = 0;
    for (int i = 0; i < n; ++i) if (indeg[i] == 0) order[qt++] = i;
    while (qh < qt) {
        const int i = order[qh++];
        if (c->next[i] >= 0 && --indeg[c->next[i]] == 0) order[qt++] = c->next[i];
    }
    if (qt != n) { free(indeg); free(order); return -3; }
    for (int k = n - 1; k >= 0; --k) {
        const int i = order[k];
        const double s_down = c->next[i] >= 0 ? sfcelv[c->next[i]] : c->dwnelv[i];
        double lo = dmax(c->rivelv[i], s_down), s = lo, qf = 0.0;
        if (q[i] > 0.0) {
            double step = 1.0;
            while (steady_flow(c, i, lo + step, s_down, &qf) < q[i] && step < 1e5) step *= 2.0;
            double hi = lo + step;
            for (int it = 0; it < 200 && hi - lo > 1e-12 * (1.0 + fabs(hi)); ++it) {
                const double mid = 0.5 * (lo + hi);
                if (mid <= lo || mid >= hi) break;
                if (steady_flow(c, i, mid, s_down, &qf) < q[i]) lo = mid; else hi = mid;
            }
            s = 0.5 * (lo + hi);
        }
        sfcelv[i] = s;
        const double qt = steady_flow(c, i, s, s_down, &qf);
        if (rivout) rivout[i] = qt - qf;
        if (fldout) fldout[i] = qf;
        double fsto, fdph;
        const double sto = steady_storage(c, i, s - c->rivelv[i], &fsto, &fdph);
        fldsto[i] = fsto;
        rivsto[i] = sto - fsto;
    }
    free(indeg); free(order);
    return 0;
}

int cmfh_accumulate_gain(int n, const int32_t* next, int nrec, const double* v_re, const double* v_im,
                         const double* g_re, const double* g_im, double* o_re, double* o_im)
{
    int32_t* indeg = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!indeg || !order) { free(indeg); free(order); return -2; }
    for (int i = 0; i < n; ++i) {
        if (next[i] >= n) { free(indeg); free(order); return -1; }
        if (next[i] >= 0) indeg[next[i]] += 1;
    }
    int qh = 0, qt = 0;
    for (int i = 0; i < n; ++i) if (indeg[i] == 0) order[qt++] = i;
    while (qh < qt) {
        const int i = order[qh++];
        if (next[i] >= 0 && --indeg[next[i]] == 0) order[qt++] = next[i];
    }
    if (qt != n) { free(indeg); free(order); return -3; }
    for (int r = 0; r < nrec; ++r) {
        const size_t b = (size_t)r * n;
        for (int i = 0; i < n; ++i) { o_re[b + i] = v_re[b + i]; o_im[b + i] = v_im[b + i]; }
        for (int k = 0; k < n; ++k) {
            const int i = order[k];
            const int j = next[i];
            if (j < 0) continue;
            const double a = o_re[b + i], c = o_im[b + i], gr = g_re[b + i], gi = g_im[b + i];
            o_re[b + j] += gr * a - gi * c;
            o_im[b + j] += gr * c + gi * a;
        }
    }
    free(indeg); free(order);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Linear periodic response of the network around a steady state (no reference
 * counterpart; the physical prior of the spin-up surrogate).
 *
 * Around the steady surface elevations s (from cmfh_steady_state), the outflow of
 * cell i varies as dO_i = a_i ds_i - b_i ds_d(i) (d = downstream cell; a mouth's
 * downstream level is fixed) and its storage as dS_i = A_i ds_i, the three
 * derivatives taken numerically from the steady outflow and storage functions.
 * For one harmonic exp(i w t) of the local runoff r, mass balance
 *   i w A_i h_i = r_i + sum_u (a_u h_u - b_u h_i) - (a_i h_i - b_i h_d(i))
 * is a complex linear system on the river tree, solved exactly by elimination
 * from the sources down (h_u = alpha_u + beta_u h_i for every upstream u of i)
 * and substitution back up. */
static void response_coefs(const cmfh_channel* c, int i, double s, double s_down, double* a, double* b, double* area)
{
    const double d = s - c->rivelv[i];
    const double e = 1e-3;
    double qf, fs, fd;
    const double lo = s - e > s_down ? s - e : s;
    *a = (steady_flow(c, i, s + e, s_down, &qf) - steady_flow(c, i, lo, s_down, &qf)) / (s + e - lo);
    if (c->next[i] >= 0) {
        const double hi = s_down + e < s ? s_down + e : s_down;
        *b = (steady_flow(c, i, s, s_down - e, &qf) - steady_flow(c, i, s, hi, &qf)) / (hi - (s_down - e));
    } else {
        *b = 0.0;
    }
    const double dl = d - e > 0.0 ? d - e : 0.0;
    *area = (steady_storage(c, i, d + e, &fs, &fd) - steady_storage(c, i, dl, &fs, &fd)) / (d + e - dl);
}

int cmfh_periodic_response(const cmfh_channel* c, const double* sfcelv, int nh, const double* omega,
                           const double* r_re, const double* r_im, double* h_re, double* h_im,
                           double* o_re, double* o_im)
{
    if (!c || !sfcelv || nh < 0 || (nh > 0 && (!omega || !r_re || !r_im || !h_re || !h_im)) || c->n <= 0)
        return -1;
    const int n = c->n;
    int32_t* indeg = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    double* co = (double*)malloc(sizeof(double) * 3 * (size_t)n);          /* a, b, A */
    double* acc = (double*)malloc(sizeof(double) * 5 * (size_t)n);         /* sum b, sum a*beta, sum a*alpha */
    double* ab = (double*)malloc(sizeof(double) * 4 * (size_t)n);          /* alpha, beta (complex) */
    if (!indeg || !order || !co || !acc || !ab) {
        free(indeg); free(order); free(co); free(acc); free(ab); return -2;
    }
    for (int i = 0; i < n; ++i) {
        if (c->next[i] >= n) { free(indeg); free(order); free(co); free(acc); free(ab); return -1; }
        if (c->next[i] >= 0) indeg[c->next[i]] += 1;
    }
    int qh = 0, qt = 0;
    for (int i = 0; i < n; ++i) if (indeg[i] == 0) order[qt++] = i;
    while (qh < qt) {
        const int i = order[qh++];
        if (c->next[i] >= 0 && --indeg[c->next[i]] == 0) order[qt++] = c->next[i];
    }
    if (qt != n) { free(indeg); free(order); free(co); free(acc); free(ab); return -3; }
    for (int i = 0; i < n; ++i) {
        const double sd = c->next[i] >= 0 ? sfcelv[c->next[i]] : c->dwnelv[i];
        response_coefs(c, i, sfcelv[i], sd, &co[3 * i], &co[3 * i + 1], &co[3 * i + 2]);
    }
    for (int k = 0; k < nh; ++k) {
        const double w = omega[k];
        const double* rr = r_re + (size_t)k * n;
        const double* ri = r_im + (size_t)k * n;
        double* hr = h_re + (size_t)k * n;
        double* hi = h_im + (size_t)k * n;
        for (size_t t = 0; t < 5 * (size_t)n; ++t) acc[t] = 0.0;
        for (int t = 0; t < n; ++t) {
            const int i = order[t];
            const double a = co[3 * i], b = co[3 * i + 1], A = co[3 * i + 2];
            double* s = acc + 5 * (size_t)i;      /* sum b_u, sum a_u beta_u (re, im), sum a_u alpha_u (re, im) */
            const double dr = a + s[0] - s[1], di = w * A - s[2];
            const double nr = rr[i] + s[3], ni = ri[i] + s[4];
            const double den = dr * dr + di * di;
            const double al_r = (nr * dr + ni * di) / den, al_i = (ni * dr - nr * di) / den;
            const double be_r = b * dr / den, be_i = -b * di / den;
            ab[4 * i] = al_r; ab[4 * i + 1] = al_i; ab[4 * i + 2] = be_r; ab[4 * i + 3] = be_i;
            const int j = c->next[i];
            if (j >= 0) {
                double* sj = acc + 5 * (size_t)j;
                sj[0] += b;
                sj[1] += a * be_r; sj[2] += a * be_i;
                sj[3] += a * al_r; sj[4] += a * al_i;
            }
        }
        for (int t = n - 1; t >= 0; --t) {
            const int i = order[t];
            const int j = c->next[i];
            double vr = ab[4 * i], vi = ab[4 * i + 1];
            if (j >= 0) {
                const double br = ab[4 * i + 2], bi = ab[4 * i + 3];
                vr += br * hr[j] - bi * hi[j];
                vi += br * hi[j] + bi * hr[j];
            }
            hr[i] = vr; hi[i] = vi;
        }
        if (o_re && o_im) {
            /* outflow amplitude a_i h_i - b_i h_down (the downstream level of a mouth is fixed) */
            double* orr = o_re + (size_t)k * n;
            double* oi = o_im + (size_t)k * n;
            for (int i = 0; i < n; ++i) {
                const int j = c->next[i];
                const double a = co[3 * i], b = co[3 * i + 1];
                orr[i] = a * hr[i] - (j >= 0 ? b * hr[j] : 0.0);
                oi[i] = a * hi[i] - (j >= 0 ? b * hi[j] : 0.0);
            }
        }
    }
    free(indeg); free(order); free(co); free(acc); free(ab);
    return 0;
}

int cmfh_storage_at_level(const cmfh_channel* c, const double* sfcelv, double* rivsto, double* fldsto)
{
    if (!c || !sfcelv || !rivsto || !fldsto || c->n <= 0) return -1;
    for (int i = 0; i < c->n; ++i) {
        double d = sfcelv[i] - c->rivelv[i];
        if (d < 0.0) d = 0.0;
        double fs, fd;
        const double sto = steady_storage(c, i, d, &fs, &fd);
        fldsto[i] = fs;
        rivsto[i] = sto - fs;
    }
    return 0;
}
