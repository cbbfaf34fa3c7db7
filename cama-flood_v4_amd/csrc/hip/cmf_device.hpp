// cmf_device.hpp -- device-side data layout and per-cell physics of the MI355X
// CaMa-Flood substep. Shared by the kernels (cmf_kernels.hip) and the host glue.
//
// Layout in HBM (one structure of arrays per handle):
//  * cells: the domains of the handle concatenated, each domain in depth-first
//    river order (or NSEQ order); static arrays are indexed by cell c.
//  * members: a handle runs NMEM = 2^lgm members (runoff realisations) of every
//    domain. Per-member state is member-minor: state slot s = (c << lgm) | m, so
//    the members of a cell are adjacent, static map values are read once per
//    cell for all its members, and the neighbour gathers (j << lgm) | m of a
//    wavefront stay as coalesced as the cells themselves. Instance k = dom*NMEM+m
//    has its own adaptive step (InstStep). A single model is NMEM=1, one domain.
//  * static map: the seven float32 source maps (every reference map is read from
//    a float32 file, so float32 holds them exactly) and the NLFP float32 profile
//    planes; the derived tables (RIVSTOMAX, RIVELV, FLDSTOMAX, FLDGRD) are
//    recomputed in FP64 in registers with the reference's own statements, so the
//    FP64 path reads 4 B instead of 8-16 B per map value and stays bit-identical.
//  * topology: I1NEXT (0-based) and the upstream CSR (ascending sources).
//  * state: storages FP64; flux/stage arrays of type R (double for the FP64
//    path, float for the mixed path); inflow accumulators FP64.
//
// Arithmetic of the FP64 path keeps the reference's statement order (the Fortran
// left-to-right product order, x**(-1.) = 1/x, x**0.5 = sqrt(x), default-kind
// literals widened from float). Files are compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cmf {

// default-kind (float) literals of the reference, widened
constexpr double K_1EM5F  = (double)1.e-5f;
constexpr double K_1EM10F = (double)1.e-10f;
constexpr double K_0P01F  = (double)0.01f;
constexpr double K_2O3F   = (double)(2.f / 3.f);
constexpr double K_0P1F   = (double)0.1f;
constexpr double K_0P2F   = (double)0.2f;
constexpr double K_1E18F  = (double)1.e18f;
constexpr double K_P7O3   = -7.0 / 3.0;     // -7._JPRB/3._JPRB
constexpr double K_P4O3   = -4.0 / 3.0;     // -4._JPRB/3._JPRB
constexpr double K_P7O3F  = (double)(-7.f / 3.f);   // default-real -7./3.
constexpr double K_P4O3F  = (double)(-4.f / 3.f);   // default-real -4./3.
constexpr double K_LEVSENT = 1.e18;         // 1.E18_JPRB sentinel of non-levee cells

constexpr int NGLB = 11;
enum { G_STOPRE, G_STONXT, G_STONEW, G_RIVINF, G_RIVOUT,
       G_STOPRE2, G_STONEW2, G_RIVSTO, G_FLDSTO, G_FLDARE, G_LEVSTO };

constexpr int NLFP_MAX = 16;
constexpr int NMEM_MAX = 64;                   // members per domain (power of two)
constexpr int NDOM_ARG = 32;                   // domain offsets carried in the kernel arguments

template <class T> __device__ __forceinline__ T vmax(T a, T b) { return a > b ? a : b; }
template <class T> __device__ __forceinline__ T vmin(T a, T b) { return a < b ? a : b; }
__device__ __forceinline__ double vpow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float vpow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double vsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float vsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double vabs(double x) { return fabs(x); }
__device__ __forceinline__ float vabs(float x) { return fabsf(x); }

struct DevMap {
    int32_t n, nseqriv, nlfp, npth, npthlev, ndam;
    int32_t lfplain, nlevcell;
    int32_t lgm, nmem;                         // members per domain = nmem = 1 << lgm
    int32_t ns, nsp;                           // state slots: n << lgm, npth << lgm
    double dfrcinc;
    const int32_t* cell_dom;                   // domain of each cell (NULL: one domain)
    int32_t ndom_arg;                          // > 0: domains found from dom_off_arg instead
    int32_t dom_off_arg[NDOM_ARG + 1];         // first cell of each domain (device order)
    const int32_t* pth_dom;                    // domain of each path (NULL: one domain)
    const int32_t* next;                       // 0-based downstream, <0 for mouths
    const int32_t* up_ptr;                     // upstream CSR (n+1)
    const int32_t* up_idx;
    const float *elevtn, *rivhgt, *nxtdst, *rivwth, *rivman, *rivlen, *grarea;
    const float* fldhgt;                       // [nlfp][n]
    const double* dwnelv;                      // boundary elevation (mouths)
    const int32_t* i2mask;
    // bifurcation paths, 0-based endpoints (-1 outside domain)
    const int32_t *pth_upst, *pth_down;
    const double *pth_dst, *pth_elv, *pth_wth, *pth_man;   // elv/wth [npthlev][npth]
    const int32_t *pout_ptr, *pout_idx, *pinf_ptr, *pinf_idx;
    // levee: compact table of levee cells, lev_idx[i] = row or -1
    const int32_t* lev_idx;
    const double* lev_tab;                     // [7][nlevcell]
    // dams
    const int32_t* i1dam;                      // per cell (0,1,10,11)
    const int32_t* dam_of_cell;                // per cell: dam row; -2 other cells with I1DAM > 0; -1
    const int32_t* damstat;
    const double *dam_convol, *dam_emevol, *dam_adjvol, *dam_qa, *dam_qn, *dam_qf;
    // Hanazaki-2022 rule (LDAMH22): NorVol and R_VolUpa of every dam
    int32_t ldamh22;
    const double *dam_norvol, *dam_r_volupa;
};

// Per-instance step of the current outer step (written by the host after the
// CFL bound is known): DT, PGRV*DT, NADD = sum of DT over the NT substeps.
// input matrix of the runoff interpolation (INPX, INPY, INPA of YOS_CMF_MAP), device
// cell order, entry-planar: entry k of cell c at k * n + c
struct RoffMap {
    int inpn, nxin, nyin, n;
    const int* inpx;
    const int* inpy;
    const double* inpa;
    double drofunit, rmis;
};

// output-period accumulators (D2*_oAVG / D2*_oMAX of YOS_CMF_DIAG), FP64 in both
// precisions, one value per state slot; pthflw [npthlev][nsp]
struct OutState {
    double *rivout, *fldout, *outflw, *rivvel, *pthout, *runoff, *daminf;
    double *outflw_max, *rivdph_max, *storge_max;
    double* pthflw;
};

struct InstStep {
    double dt, gdt, nadd;
    int32_t nt;                                // substeps of this instance
    int32_t swapfix;                           // odd number of idle substeps: RIVDPH pair swapped
};

template <class R>
struct DevState {
    double *rivsto, *fldsto, *levsto, *damsto, *daminf;
    R *rivout, *fldout;                        // limited outflow == *_PRE after a substep
    const R* rivout_vel;                       // D2RIVOUT read by OUTFLW for RIVVEL
    R *qr_u, *qf_u;                            // outflow before the inflow limiter
    R *rivdph_pre, *fldsto_pre;
    R* rivdph_next;                            // buffer receiving the new RIVDPH (k_update)
    R *rivdph, *flddph, *fldfrc, *fldare, *sfcelv, *storge, *outflw, *rivvel, *levdph;
    double *rivinf, *fldinf, *pthout;
    R* rate;
    R* runoff;
    R* upsinf;                                 // LUPSINF: D2UPSINF [m3/s] (NULL when off)
    R *pthflw, *pthflw_u, *pthflwsum, *pthflwsum_u;   // [npthlev][nsp], [nsp]
    R *rivout_avg, *fldout_avg, *outflw_avg, *rivvel_avg, *pthout_avg, *runoff_avg;
    R *outflw_max, *rivdph_max, *storge_max, *daminf_avg;
    R *pthflw_avg, *pthflwsum_avg;
    double* partials;                          // [nt][ntiles][NGLB] (single instance only)
    double* glb_hist;                          // [nt_cap][NGLB]
    unsigned long long* cfl;                   // [ninst]: bits of the CFL bound (positive doubles)
    double* cfl_part;                          // single instance: CFL bound of each stage block
    const InstStep* inst;                      // [ninst]
    unsigned* bar;                             // [2]: grid barrier of the fused substep kernel
    int32_t* err;                              // [1]: device error flag (barrier time-out)
};

struct StepArgs {
    double pgrv, pmanfld, pdstmth, pcadp, pminslp;
    int32_t lfldout, lpthout, ldamout, llevee;
    int32_t it;                                // substep index within the outer step
    int32_t prow;                              // row of S.partials this substep's budget sums go to
    int32_t budget;                            // write the per-tile water-budget sums
    int32_t xcd;                               // XCD-contiguous tile order
    int32_t lupsinf;                           // add S.upsinf to the river in STONXT
    int32_t lkine;                             // kinematic-wave outflow (LKINE)
    int32_t lslpmix;                           // slope-mixed outflow (LSLPMIX, without LKINE)
    int32_t daminf_sparse;                     // DAMINF is zero in every cell with I1DAM == 0
};

// Cell / path slot -> static index, member and instance. MB = false is the
// single-instance handle (one domain, one member): slot = cell, instance 0,
// known at compile time.
struct Slot {
    int c, m, inst;
};
template <bool MB>
__device__ __forceinline__ Slot cell_slot(const DevMap& M, int s) {
    Slot x;
    if (!MB) { x.c = s; x.m = 0; x.inst = 0; return x; }
    x.c = s >> M.lgm;
    x.m = s & (M.nmem - 1);
    int d = 0;
    if (M.ndom_arg > 0) {
        // compare against the offsets in the kernel arguments: no dependent load
        for (int k = 1; k < M.ndom_arg; ++k) d += x.c >= M.dom_off_arg[k];
    } else if (M.cell_dom) {
        d = M.cell_dom[x.c];
    }
    x.inst = d * M.nmem + x.m;
    return x;
}
template <bool MB>
__device__ __forceinline__ Slot path_slot(const DevMap& M, int q) {
    Slot x;
    if (!MB) { x.c = q; x.m = 0; x.inst = 0; return x; }
    x.c = q >> M.lgm;
    x.m = q & (M.nmem - 1);
    x.inst = (M.pth_dom ? M.pth_dom[x.c] * M.nmem : 0) + x.m;
    return x;
}
template <bool MB>
__device__ __forceinline__ int sidx(const DevMap& M, int c, int m) { return MB ? ((c << M.lgm) | m) : c; }

// ---------------------------------------------------------------------------
// Floodplain profile table recomputed on the fly: SET_FLDSTG
// (src/cmf_ctrl_maps_mod.F90:984-995) and, identically, CMF_LEVEE_INIT :169-181.
struct Geom {                 // float32 map values of one cell, widened
    double rivlen, rivwth, grarea, rivhgt;
};
__device__ __forceinline__ Geom load_geom(const DevMap& M, int i) {
    Geom g;
    g.rivlen = (double)M.rivlen[i]; g.rivwth = (double)M.rivwth[i];
    g.grarea = (double)M.grarea[i]; g.rivhgt = (double)M.rivhgt[i];
    return g;
}

struct FldTab {
    const float* h;          // &fldhgt[i], stride n
    int n;
    double rivlen, rivwth, dwthinc, stopre, hpre;
    __device__ __forceinline__ void init(const DevMap& M, int i, const Geom& g, double rivstomax) {
        h = M.fldhgt + i; n = M.n;
        rivlen = g.rivlen;
        rivwth = g.rivwth;
        dwthinc = (g.grarea * (1.0 / rivlen)) * M.dfrcinc;
        stopre = rivstomax; hpre = 0.0;
    }
    // next layer l (1-based, consecutive calls): FLDSTOMAX(l), FLDGRD(l), FLDHGT(l)
    __device__ __forceinline__ void layer(int l, double& smax, double& grd, double& hgt) {
        const double hl = (double)h[(size_t)(l - 1) * n];
        const double dstonow = (rivlen * (rivwth + dwthinc * ((double)l - 0.5))) * (hl - hpre);
        smax = stopre + dstonow;
        grd = (hl - hpre) * (1.0 / dwthinc);
        hgt = hl;
        stopre = smax; hpre = hl;
    }
};

__device__ __forceinline__ double rivstomax_of(const DevMap& M, int i) {
    // maps:716-721
    return M.lfplain ? ((double)M.rivlen[i] * (double)M.rivwth[i]) * (double)M.rivhgt[i] : K_1E18F;
}
__device__ __forceinline__ double rivstomax_of(const DevMap& M, const Geom& g) {
    return M.lfplain ? (g.rivlen * g.rivwth) * g.rivhgt : K_1E18F;
}
__device__ __forceinline__ double rivelv_of(const DevMap& M, int i) {
    return (double)M.elevtn[i] - (double)M.rivhgt[i];            // maps:722
}

struct Stage {
    double rivsto, fldsto, levsto;
    double rivdph, flddph, fldfrc, fldare, levdph;
};

// CMF_CALC_FLDSTG_DEF per cell, src/cmf_calc_fldstg_mod.F90:49-95
__device__ __forceinline__ void stage_def(const DevMap& M, int i, const Geom& geo, double rivsto, double fldsto,
                                          Stage& o) {
    const double pstoall = rivsto + fldsto;
    const double dstoall = pstoall;
    const double rivstomax = rivstomax_of(M, geo);
    const double rivlen = geo.rivlen, rivwth = geo.rivwth;
    o.levsto = 0.0; o.levdph = 0.0;
    if (pstoall > rivstomax) {
        const int nl = M.nlfp;
        FldTab T; T.init(M, i, geo, rivstomax);
        const double dwth_inc = (geo.grarea / rivlen) * M.dfrcinc;
        int I = 1;
        double dsto_fil = rivstomax, dwth_fil = rivwth, ddph_fil = 0.0;
        double smax, grd, hgt;
        T.layer(1, smax, grd, hgt);
        while (pstoall > smax) {                      // :59-65
            dsto_fil = smax;
            dwth_fil = dwth_fil + dwth_inc;
            ddph_fil = ddph_fil + grd * dwth_inc;
            I = I + 1;
            if (I > nl) break;
            T.layer(I, smax, grd, hgt);
        }
        double dwth_add;
        if (I > nl) {                                 // :66-70
            const double dsto_add = dstoall - dsto_fil;
            dwth_add = 0.0;
            o.flddph = ddph_fil + (dsto_add / dwth_fil) / rivlen;
        } else {                                      // :71-74
            const double dsto_add = dstoall - dsto_fil;
            dwth_add = -dwth_fil + sqrt(dwth_fil * dwth_fil + ((2.0 * dsto_add) / rivlen) / grd);
            o.flddph = ddph_fil + grd * dwth_add;
        }
        double rs = rivstomax + (rivlen * rivwth) * o.flddph;     // :76-77
        rs = vmin(rs, pstoall);
        o.rivsto = rs;
        o.rivdph = (rs / rivlen) / rivwth;
        const double fs = pstoall - rs;
        o.fldsto = vmax(fs, 0.0);
        double fr = ((-rivwth + dwth_fil) + dwth_add) / (dwth_inc * (double)nl);
        fr = vmax(fr, 0.0);
        fr = vmin(fr, 1.0);
        o.fldfrc = fr;
        o.fldare = geo.grarea * fr;
    } else {                                          // :87-95
        o.rivsto = pstoall;
        const double d = (dstoall / rivlen) / rivwth;
        o.rivdph = vmax(d, 0.0);
        o.fldsto = 0.0; o.flddph = 0.0; o.fldfrc = 0.0; o.fldare = 0.0;
    }
}

// levee parameters of one cell (sentinels of CMF_LEVEE_INIT for cells without one)
struct Lev {
    double levhgt, levfrc, bashgt, levdst, levbassto, levtopsto, levfilsto;
};
__device__ __forceinline__ Lev load_lev_row(const DevMap& M, int li) {
    Lev L{0.0, 1.0, K_LEVSENT, K_LEVSENT, K_LEVSENT, K_LEVSENT, K_LEVSENT};
    const int nlc = M.nlevcell;
    if (li >= 0) {
        L.levhgt = M.lev_tab[li]; L.levfrc = M.lev_tab[nlc + li]; L.bashgt = M.lev_tab[2 * nlc + li];
        L.levdst = M.lev_tab[3 * nlc + li]; L.levbassto = M.lev_tab[4 * nlc + li];
        L.levtopsto = M.lev_tab[5 * nlc + li]; L.levfilsto = M.lev_tab[6 * nlc + li];
    }
    return L;
}
__device__ __forceinline__ Lev load_lev(const DevMap& M, int i) { return load_lev_row(M, M.lev_idx[i]); }

// CMF_LEVEE_FLDSTG per cell, src/cmf_ctrl_levee_mod.F90:302-473. Every case
// computes the same eight local values and the result is written once, so the
// compiler keeps the whole inversion in registers.
__device__ __forceinline__ void stage_levee(const DevMap& M, int i, const Geom& geo, const Lev& L, double rivsto,
                                            double fldsto, double levsto, Stage& o) {
    const int nl = M.nlfp;
    const double levhgt = L.levhgt, levfrc = L.levfrc, bashgt = L.bashgt, levdst = L.levdst;
    const double levbassto = L.levbassto, levtopsto = L.levtopsto, levfilsto = L.levfilsto;
    const double pstoall = (rivsto + fldsto) + levsto;
    const double dstoall = pstoall;
    const double rivstomax = rivstomax_of(M, geo);
    const double rl = geo.rivlen, rw = geo.rivwth, ga = geo.grarea;
    const double dwth_inc = (ga / rl) * M.dfrcinc;
    double r_riv = 0.0, r_fld = 0.0, r_lev = 0.0, r_rdph = 0.0, r_fdph = 0.0, r_frc = 0.0, r_are = 0.0,
           r_ldph = 0.0;
    // case selection (:308-311): 0 in bank; 1 below the levee base; 2 river side
    // below the crown; 3 protected side filling; 4 above the protected-side fill
    int kase;
    if (!(dstoall > rivstomax)) kase = 0;
    else if (dstoall < levbassto) kase = 1;
    else if (dstoall < levtopsto) kase = 2;
    else if (dstoall < levfilsto) kase = 3;
    else kase = 4;
    if (kase == 1 || kase == 4) {
        // case 1 (:312-349) and case 4 (:424-461) share the layer search
        FldTab T; T.init(M, i, geo, rivstomax);
        int I = 1;
        double dsto_fil = rivstomax, dwth_fil = rw, ddph_fil = 0.0, dwth_add;
        double smax, grd, hgt;
        T.layer(1, smax, grd, hgt);
        while (dstoall > smax) {
            dsto_fil = smax;
            dwth_fil = dwth_fil + dwth_inc;
            ddph_fil = ddph_fil + grd * dwth_inc;
            I = I + 1;
            if (I > nl) break;
            T.layer(I, smax, grd, hgt);
        }
        if (I <= nl) {
            const double dsto_add = dstoall - dsto_fil;
            dwth_add = -dwth_fil + sqrt(dwth_fil * dwth_fil + ((2.0 * dsto_add) / rl) / grd);
            r_fdph = ddph_fil + grd * dwth_add;
        } else {
            const double dsto_add = dstoall - dsto_fil;
            dwth_add = 0.0;
            r_fdph = ddph_fil + (dsto_add / dwth_fil) / rl;
        }
        if (kase == 1) {
            r_riv = rivstomax + (rl * rw) * r_fdph;
            r_rdph = (r_riv / rl) / rw;
            const double fs = pstoall - r_riv;
            r_fld = vmax(fs, 0.0);
            double fr = ((-rw + dwth_fil) + dwth_add) / (dwth_inc * (double)nl);
            fr = vmax(fr, 0.0);
            fr = vmin(fr, 1.0);
            r_frc = fr;
            r_are = ga * fr;
        } else {
            r_frc = ((-rw + dwth_fil) + dwth_add) / (dwth_inc * (double)nl);
            r_are = ga * r_frc;
            r_riv = rivstomax + (rl * rw) * r_fdph;
            r_rdph = (r_riv / rl) / rw;
            const double dsto_add = ((r_fdph - levhgt) * (levdst + rw)) * rl;
            const double fs = (levtopsto + dsto_add) - r_riv;
            r_fld = vmax(fs, 0.0);
            const double ls = (pstoall - r_riv) - r_fld;
            r_lev = vmax(ls, 0.0);
            r_ldph = r_fdph;
        }
    } else if (kase == 2) {                           // (:353-368)
        const double dsto_add = dstoall - levbassto;
        const double dwth_add = levdst + rw;
        r_fdph = bashgt + (dsto_add / dwth_add) / rl;
        r_riv = rivstomax + (rl * rw) * r_fdph;
        r_rdph = (r_riv / rl) / rw;
        const double fs = pstoall - r_riv;
        r_fld = vmax(fs, 0.0);
        r_frc = levfrc;
        r_are = ga * r_frc;
    } else if (kase == 3) {                           // (:372-420): river side full
        r_fdph = levhgt;
        r_riv = rivstomax + (rl * rw) * r_fdph;
        r_rdph = (r_riv / rl) / rw;
        const double fs = levtopsto - r_riv;
        r_fld = vmax(fs, 0.0);
        const double ls = (pstoall - r_riv) - r_fld;
        r_lev = vmax(ls, 0.0);
        const int ilev = (int)(levfrc * (double)nl) + 1;
        double dsto_fil = levtopsto, dwth_fil = 0.0, ddph_fil = 0.0;
        FldTab T; T.init(M, i, geo, rivstomax);
        double smax = 0.0, grd = 0.0, hgt = 0.0;
        for (int l = 1; l < ilev && l <= nl; ++l) T.layer(l, smax, grd, hgt);
        int I = ilev;
        while (I <= nl) {
            T.layer(I, smax, grd, hgt);
            const double dsto_add = ((levdst + rw) * (levhgt - hgt)) * rl;
            if (dstoall < smax + dsto_add) break;
            dsto_fil = smax + dsto_add;
            dwth_fil = dwth_inc * (double)I - levdst;
            ddph_fil = hgt - bashgt;
            I = I + 1;
        }
        if (I <= nl) {
            const double dsto_add = dstoall - dsto_fil;
            const double dwth_add = -dwth_fil + sqrt(dwth_fil * dwth_fil + ((2.0 * dsto_add) / rl) / grd);
            const double ddph_add = dwth_add * grd;
            r_ldph = (bashgt + ddph_fil) + ddph_add;
            double fr = (dwth_fil + levdst) / (dwth_inc * (double)nl);
            fr = vmax(fr, 0.0);
            fr = vmin(fr, 1.0);
            r_frc = fr;
            r_are = ga * fr;
        } else {
            const double dsto_add = dstoall - dsto_fil;
            const double ddph_add = (dsto_add / dwth_fil) / rl;
            r_ldph = (bashgt + ddph_fil) + ddph_add;
            r_frc = 1.0;
            r_are = ga * r_frc;
        }
    } else {                                          // case 0 (:463-473)
        r_riv = pstoall;
        const double d = (dstoall / rl) / rw;
        r_rdph = vmax(d, 0.0);
    }
    o.rivsto = r_riv; o.fldsto = r_fld; o.levsto = r_lev;
    o.rivdph = r_rdph; o.flddph = r_fdph; o.fldfrc = r_frc; o.fldare = r_are; o.levdph = r_ldph;
}

}  // namespace cmf
