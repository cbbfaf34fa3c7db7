// cmf_kernels.hip -- HIP kernels of one CMF_PHYSICS_ADVANCE outer step on gfx950.
//
// One outer step = k_stage (FLDSTG + per-instance CFL bound) followed by NT
// substeps of three phases; every phase is a grid-wide dependency of the
// reference's substep (src/cmf_ctrl_physics_mod.F90:64-144):
//   outflow  CMF_CALC_OUTFLW + CMF_DAMOUT_CALC (cells) and CMF_CALC_PTHOUT /
//            CMF_LEVEE_OPT_PTHOUT (paths)
//   rate     CMF_CALC_INFLOW_LSPAMAT parts 1-2: outflow volume and limiter rate
//   update   CMF_CALC_INFLOW_LSPAMAT parts 2-3 (limit, upstream gather),
//            CMF_DAMOUT_WATBAL, CALC_VARS_PRE, CMF_CALC_STONXT,
//            CMF_PHYSICS_FLDSTG and CMF_DIAG_AVEMAX_ADPSTP (cells) and the path
//            limiter/averages (paths)
// A substep runs either as three kernels (k_outflow, k_rate, k_update) or as one
// persistent kernel (k_substep) whose blocks loop over the same tiles and meet
// at a grid barrier between the phases. The tile functions are shared, so both
// forms compute the same values.
//
// Every work item is a state slot s = (cell << lgm) | member; all neighbour
// accesses are gathers (downstream cell, upstream CSR list, path endpoints) of
// the same member; nothing is scattered, so results do not depend on
// scheduling, block size or tile order. An instance (domain, member) whose NT
// is reached idles for the remaining substeps of the outer step.
#include "cmf_device.hpp"
#include "cmf_kernels.hpp"

namespace cmf {

// ---------------------------------------------------------------------------
// Block reductions (fixed tree order: deterministic for a given block size)
// Sum over the block: a fixed butterfly inside each 64-lane wavefront, then the
// BS/64 wave sums in wave order. Result in lds[k * NW] for thread-independent
// reads after the final barrier.
template <int BS, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* [NV][BS/64] */) {
    constexpr int NW = BS / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k * NW + wave] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < NW; ++w) s += lds[threadIdx.x * NW + w];
        lds[threadIdx.x * NW] = s;
    }
    __syncthreads();
}

// Compensated block sum (TwoSum, Knuth): every partial sum is carried as hi + lo
// with lo the exact rounding error of each addition, so the result is the
// correctly rounded sum of the NV inputs up to O(2^-106) relative terms, whatever
// the number of cells. Needs IEEE additions without contraction (the files are
// compiled with -ffp-contract=off). Result in lds[k * NW] as hi + lo, rounded.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
template <int BS, int NV>
__device__ __forceinline__ void block_sum_comp(double (&v)[NV], double* lds /* [2][NV][BS/64] */) {
    constexpr int NW = BS / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double lo[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) lo[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double oh = __shfl_down(v[k], off, 64), ol = __shfl_down(lo[k], off, 64);
            double sh, e;
            two_sum(v[k], oh, sh, e);
            v[k] = sh;
            lo[k] = (lo[k] + ol) + e;
        }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) { lds[k * NW + wave] = v[k]; lds[(NV + k) * NW + wave] = lo[k]; }
    __syncthreads();
    if (threadIdx.x < NV) {
        double hi = 0.0, l = 0.0;
        for (int w = 0; w < NW; ++w) {
            double sh, e;
            two_sum(hi, lds[threadIdx.x * NW + w], sh, e);
            hi = sh;
            l = (l + lds[(NV + threadIdx.x) * NW + w]) + e;
        }
        lds[threadIdx.x * NW] = hi + l;
    }
    __syncthreads();
}

// Tile t of a launch of G blocks: with xcd set and G a multiple of 8, the blocks
// that the dispatcher places on one XCD (b, b+8, b+16, ...) take one contiguous
// eighth of the tiles, so neighbour gathers across tile edges mostly hit the
// same L2.
__device__ __forceinline__ int tile_of(int b, int G, int ntiles, int xcd) {
    if (!xcd || (G & 7)) return b;
    const int per = G >> 3;
    return (b & 7) * per + (b >> 3);
}

struct InstView {
    double dt, gdt, nadd;
    int nt;
    bool active, first, last;
};
__device__ __forceinline__ InstView inst_view(const InstStep* tab, int inst, int it) {
    const InstStep& I = tab[inst];
    InstView v;
    v.dt = I.dt; v.gdt = I.gdt; v.nadd = I.nadd; v.nt = I.nt;
    v.active = it < I.nt;
    v.first = it == 0;
    v.last = it == I.nt - 1;
    return v;
}

// ---------------------------------------------------------------------------
// Kinematic wave, CMF_CALC_OUTFLW_KINE (src/cmf_opt_outflw_mod.F90:27-127):
// slope from the bed elevations over NXTDST, at least PMINSLP (mouths: PMINSLP,
// the LSLOPEMOUTH = .FALSE. branch :86-91); velocity (1/n) S^(1/2) h^(2/3);
// outflow capped by storage / DT; RIVVEL is the kinematic velocity. In R (FP64:
// the reference's statements in order; mixed: FP32 flux arithmetic, the slope
// formed in FP64 and the storage caps divided in FP64, each rounded once).
template <class R>
__device__ __forceinline__ void kine_outflow(const DevMap& M, const DevState<R>& S, const StepArgs& A, int i,
                                             int nxt, int s, double dt, R& q_r, R& q_f, R& vel) {
    const R one = R(1), zero = R(0);
    double dsl = A.pminslp;
    if (nxt >= 0) {                                                   // :57-58
        dsl = ((double)M.elevtn[i] - (double)M.elevtn[nxt]) * (1.0 / (double)M.nxtdst[i]);
        dsl = vmax(dsl, A.pminslp);
    }
    const R dslope = R(dsl);
    const R rd = S.rivdph[s];
    const R dvel = ((one / R(M.rivman[i])) * vsqrt(dslope)) * vpow(rd, R(K_2O3F));   // :59
    const R darea = R(M.rivwth[i]) * rd;                             // :60
    vel = dvel;
    q_r = vmin(darea * dvel, R(S.rivsto[s] / dt));                   // :63-64
    q_f = zero;
    if (A.lfldout) {                                                  // :66-75
        const R dslope_f = vmin(R(0.005), dslope);
        const R fd = S.flddph[s];
        const R dvel_f = ((one / R(A.pmanfld)) * vsqrt(dslope_f)) * vpow(fd, R(K_2O3F));
        R dare_f = R(S.fldsto[s]) * (one / R(M.rivlen[i]));
        dare_f = vmax(dare_f - fd * R(M.rivwth[i]), zero);
        q_f = vmin(dare_f * dvel_f, R(S.fldsto[s] / dt));
    }
}

// Slope-mixed outflow, CMF_CALC_OUTFLW_KINEMIX (src/cmf_opt_outflw_mod.F90:118-300),
// for a cell with I2MASK = 0 (the others take kine_outflow): the local inertial
// equations in that routine's variant. Against CMF_CALC_OUTFLW: no floor on the
// river flow area, RIVVEL from the new outflow, the default-real exponents -7./3.
// and -4./3., the floodplain tested on its semi-implicit area against a
// default-real 1.E-5, the floodplain flow dropped only when it opposes the river
// flow (product < 0), and no reverse-flow limiter. Evaluated in FP64 for either
// state type R, in the reference's statement order.
template <class R, bool MB>
__device__ __forceinline__ void kinemix_outflow(const DevMap& M, const DevState<R>& S, const StepArgs& A, int i,
                                                int nxt, int s, int m, double dt, R& q_r, R& q_f, R& vel) {
    const double gdt = A.pgrv * dt;
    const double rivelv = rivelv_of(M, i), elevtn = (double)M.elevtn[i];
    const double rivwth = (double)M.rivwth[i], rivman = (double)M.rivman[i];
    const double rd = (double)S.rivdph[s], rdp = (double)S.rivdph_pre[s];
    const double sfc = rivelv + rd, sfc_pre = rivelv + rdp;                  // :141-142
    const double fdp = vmax(rdp - (double)M.rivhgt[i], 0.0);                 // :143
    double dslope, dflw, dflw_pre, fa, fap;
    if (nxt >= 0) {                                                          // :149-175
        const int sj = sidx<MB>(M, nxt, m);
        const double rivelv_j = rivelv_of(M, nxt);
        const double sfc_j = rivelv_j + (double)S.rivdph[sj], sfc_j_pre = rivelv_j + (double)S.rivdph_pre[sj];
        const double dsfcmax = vmax(sfc, sfc_j), dsfcmax_pre = vmax(sfc_pre, sfc_j_pre);
        dslope = (sfc - sfc_j) * (1.0 / (double)M.nxtdst[i]);
        dflw = dsfcmax - rivelv;
        dflw_pre = dsfcmax_pre - rivelv;
        fa = vmax(dsfcmax - elevtn, 0.0);                                   // :179
        fap = dsfcmax_pre - elevtn;                                         // :184
    } else {                                                                 // :232-253
        dslope = (sfc - (double)M.dwnelv[i]) * (1.0 / A.pdstmth);
        dflw = rd;
        dflw_pre = rdp;
        fa = sfc - elevtn;                                                  // :256
        fap = sfc_pre - elevtn;                                             // :261
    }
    const double darea = rivwth * dflw;
    double dflw_imp = sqrt(dflw * dflw_pre);
    dflw_imp = vmax(dflw_imp, 1.e-6);
    double qr = 0.0, v = 0.0;
    if (dflw_imp > K_1EM5F && darea > K_1EM5F) {
        const double dout_pre = (double)S.rivout[s] * (1.0 / rivwth);     // D2RIVOUT_PRE
        qr = (rivwth * (dout_pre + (gdt * dflw_imp) * dslope)) *
             (1.0 / (1.0 + ((gdt * (rivman * rivman)) * fabs(dout_pre)) * pow(dflw_imp, K_P7O3F)));
        v = qr * (1.0 / darea);
    }
    double qf = 0.0;
    if (A.lfldout) {                                                         // :178-200, :255-277
        const double dslope_f = vmax(-0.005, vmin(0.005, dslope));
        const double rrl = 1.0 / (double)M.rivlen[i];
        const double fd = (double)S.flddph[s];
        double dare_f = S.fldsto[s] * rrl;
        dare_f = vmax(dare_f - fd * rivwth, 0.0);
        const double t = fa * fap;
        double dflw_imp_f = sqrt(vmax(t, 0.0));
        dflw_imp_f = vmax(dflw_imp_f, 1.e-6);
        double dare_pre_f = (double)S.fldsto_pre[s] * rrl;
        dare_pre_f = vmax(dare_pre_f - fdp * rivwth, 1.e-6);
        double dare_imp_f = sqrt(dare_f * dare_pre_f);
        dare_imp_f = vmax(dare_imp_f, 1.e-6);
        if (dflw_imp_f > K_1EM5F && dare_imp_f > K_1EM5F) {
            const double dout_pre_f = (double)S.fldout[s];                  // D2FLDOUT_PRE
            qf = (dout_pre_f + (gdt * dare_imp_f) * dslope_f) *
                 (1.0 / (1.0 + (((gdt * (A.pmanfld * A.pmanfld)) * fabs(dout_pre_f)) * pow(dflw_imp_f, K_P4O3F)) *
                                   (1.0 / dare_imp_f)));
        }
        if (qf * qr < 0.0) qf = 0.0;
    }
    q_r = R(qr);
    q_f = R(qf);
    vel = R(v);
}

// ---------------------------------------------------------------------------
// phase 1a: river / mouth / floodplain outflow and dam operation, per cell slot
// KINE selects the kinematic wave at compile time: k_outflow is instantiated per
// scheme and the launch picks the one StepArgs::lkine asks for, so the default
// (local inertial) kernel is compiled as if the option did not exist; the
// persistent k_substep carries both and branches once per tile.
// Hanazaki-2022 reservoir operation (LDAMH22), src/cmf_ctrl_damout_mod.F90:354-378,
// before the storage limiter. Zones: normal (<= NorVol), water supply (<= ConVol),
// flood control (< EmeVol), emergency. r_volupa is the dam's own flood-control
// volume over its upstream area (damout:204 stores every dam's ratio at index NDAM;
// here each dam reads its own). Default-real constants 1. and 0.2 widened.
__device__ __forceinline__ double dam_rule_h22(double vol, double inflow, double norvol, double convol,
                                               double emevol, double qn, double qf, double r_volupa) {
    if (vol <= norvol) return qn * (vol / convol);
    const double quad = (vol - norvol) / (emevol - norvol);        // ((V-NorVol)/(EmeVol-NorVol))**2
    const double pre = qn * 0.5 + (quad * quad) * (qf - qn);
    if (vol <= convol) {
        if (qf <= inflow) return qn * 0.5 + ((vol - norvol) / (convol - norvol)) * (qf - qn);
        return pre;
    }
    if (vol < emevol) {
        if (qf <= inflow) {
            const double cut = vmax(1.0 - r_volupa / K_0P2F, 0.0);
            return qf + ((cut * (vol - convol)) / (emevol - convol)) * (inflow - qf);
        }
        return pre;
    }
    return vmax(inflow, qf);
}

template <class R, bool MB, bool KINE>
__device__ __forceinline__ void outflow_cell(const DevMap& M, const DevState<R>& S, const StepArgs& A, int s) {
    const Slot x = cell_slot<MB>(M, s);
    const int i = x.c, m = x.m;
    // every load that does not depend on the downstream index is issued first,
    // together with the instance's step. Ensemble slots of an instance that has
    // finished its substeps leave before any load (whole wavefronts idle once its
    // NT is reached); a single instance is active on every substep.
    const InstView I = inst_view(S.inst, x.inst, A.it);
    if (MB && !I.active) return;
    const int nxt = M.next[i];
    const R rivdph = S.rivdph[s], rivdph_pre = S.rivdph_pre[s];
    const R qr_pre = S.rivout[s];                                   // D2RIVOUT_PRE
    const R qr_old = S.rivout_vel[s];                               // D2RIVOUT (for RIVVEL)
    const float f_elevtn = M.elevtn[i], f_rivhgt = M.rivhgt[i], f_rivwth = M.rivwth[i], f_rivman = M.rivman[i];
    const float f_nxtdst = M.nxtdst[i], f_rivlen = M.rivlen[i];
    const R p_rivvel_avg = S.rivvel_avg[s];
    const R storge = S.storge[s];
    double fldsto_d = 0.0;
    R flddph = R(0), fldsto_pre = R(0), fldout_pre = R(0);
    if (A.lfldout) {
        fldsto_d = S.fldsto[s];
        flddph = S.flddph[s];
        fldsto_pre = S.fldsto_pre[s];
        fldout_pre = S.fldout[s];                                   // D2FLDOUT_PRE
    }
    if (!I.active) return;
    const R gdt = R(I.gdt);
    const R zero = R(0), one = R(1);
    const R flddph_pre = vmax(rivdph_pre - R(f_rivhgt), zero);
    const R rivwth = R(f_rivwth), rivman = R(f_rivman);
    const bool river = nxt >= 0;                                     // mouths carry a negative code
    R q_r, q_f = zero, vel;
    if constexpr (KINE) {
        // LKINE: CMF_CALC_OUTFLW_KINE (opt_outflw:27-127) in place of the local
        // inertial equations and their reverse-flow limiter; LSLPMIX (without
        // LKINE): CMF_CALC_OUTFLW_KINEMIX, the same kinematic wave where I2MASK != 0
        if (A.lslpmix && M.i2mask[i] == 0) kinemix_outflow<R, MB>(M, S, A, i, nxt, s, m, I.dt, q_r, q_f, vel);
        else kine_outflow<R>(M, S, A, i, nxt, s, I.dt, q_r, q_f, vel);
    } else {
        // Water-surface terms of outflw:45-58. With R = double they are evaluated as
        // the reference writes them (SFCELV = RIVELV + RIVDPH, slopes and depths from
        // SFCELV). With R = float the same quantities are formed from depths and the
        // exact FP64 elevation difference, rounded once: a ~250 m surface elevation
        // keeps only ~1.5e-5 m in FP32, which the slope (a difference of a few cm)
        // would otherwise inherit.
        //   hs  = SFCELV - DWNELV          (slope numerator)
        //   hr  = max(SFCELV, DWNELV) - RIVELV,  hrp the same with the _PRE values
        //   (mouths: hr = RIVDPH, hrp = RIVDPH_PRE)
        //   fa  = surface above ELEVTN for the floodplain (outflw:88, :147), fap the _PRE one
        R hs, hr, hrp, fa, fap, rdst;
        if (sizeof(R) == sizeof(double)) {
            const R rivelv = R((double)f_elevtn - (double)f_rivhgt);    // maps:722
            const R elevtn = R(f_elevtn);
            const R sfcelv = rivelv + rivdph;                           // outflw:45-47
            const R sfcelv_pre = rivelv + rivdph_pre;
            if (river) {                                                 // outflw:52-57
                const int j = nxt, sj = sidx<MB>(M, j, m);
                const R rivelv_j = R(rivelv_of(M, j));
                const R dwn = rivelv_j + S.rivdph[sj];
                const R dwn_pre = rivelv_j + S.rivdph_pre[sj];
                const R dsfc = vmax(sfcelv, dwn), dsfc_pr = vmax(sfcelv_pre, dwn_pre);
                hs = sfcelv - dwn;
                hr = dsfc - rivelv;
                hrp = dsfc_pr - rivelv;
                fa = dsfc - elevtn;
                fap = dsfc_pr - elevtn;
                rdst = one / R(f_nxtdst);
            } else {
                const R dwn = R(M.dwnelv[i]);
                hs = sfcelv - dwn;
                hr = rivdph;
                hrp = rivdph_pre;
                fa = sfcelv - elevtn;
                fap = sfcelv_pre - elevtn;
                rdst = one / R(A.pdstmth);
            }
        } else {
            const double e_i = (double)f_elevtn - (double)f_rivhgt;
            if (river) {
                const int j = nxt, sj = sidx<MB>(M, j, m);
                const R dz = R(rivelv_of(M, j) - e_i);                   // RIVELV(j) - RIVELV(i)
                const R dj = S.rivdph[sj], dj_pre = S.rivdph_pre[sj];
                hs = (rivdph - dj) - dz;
                hr = vmax(rivdph, dz + dj);
                hrp = vmax(rivdph_pre, dz + dj_pre);
                rdst = one / R(f_nxtdst);
            } else {
                hs = rivdph - R(M.dwnelv[i] - e_i);
                hr = rivdph;
                hrp = rivdph_pre;
                rdst = one / R(A.pdstmth);
            }
            const R hgt = R(f_rivhgt);                                   // ELEVTN = RIVELV + RIVHGT
            fa = hr - hgt;
            fap = hrp - hgt;
        }
        // ---- river channel: outflw:60-80 (river), :119-139 (mouth)
        {
            const R dslp = hs * rdst;
            const R dflw = hr, dflw_pr = hrp;
            R dare = rivwth * dflw;
            dare = vmax(dare, R(1.e-10));
            R dflw_im = vsqrt(dflw * dflw_pr);
            dflw_im = vmax(dflw_im, R(1.e-6));
            const R dout_pr = qr_pre * (one / rivwth);
            const R dout = (rivwth * (dout_pr + (gdt * dflw_im) * dslp)) *
                           (one / (one + ((gdt * (rivman * rivman)) * vabs(dout_pr)) * vpow(dflw_im, R(K_P7O3))));
            const R dvel = qr_old * (one / dare);
            const bool mask = (dflw_im > R(K_1EM5F)) && (dare > R(K_1EM5F));
            q_r = mask ? dout : zero;
            vel = mask ? dvel : zero;
        }
        // ---- floodplain: outflw:84-116 (river), :143-171 (mouth). Surface heights
        // above ELEVTN = RIVELV + RIVHGT: max(SFCELV, DWNELV) - ELEVTN = hr - RIVHGT.
        if (A.lfldout) {
            const R gdtn2 = gdt * (R(A.pmanfld) * R(A.pmanfld));
            const R dfsto = R(fldsto_d);
            const R rrl = one / R(f_rivlen);
            R dslp = hs * rdst;
            dslp = vmax(R(-0.005), vmin(R(0.005), dslp));
            R dflw, dflw_pr, dare, dare_pr;
            if (river) {
                dflw = vmax(fa, zero);
                dare = dfsto * rrl;
                dare = vmax(dare - flddph * rivwth, zero);
                dflw_pr = fap;
                dare_pr = fldsto_pre * rrl;
                dare_pr = vmax(dare_pr - flddph_pre * rivwth, R(1.e-6));
            } else {
                dflw = fa;
                dare = vmax(dfsto * rrl - flddph * rivwth, zero);
                dflw_pr = fap;
                dare_pr = vmax(fldsto_pre * rrl - flddph_pre * rivwth, R(1.e-6));
            }
            const R t = dflw * dflw_pr;
            R dflw_im = vsqrt(vmax(t, zero));
            dflw_im = vmax(dflw_im, R(1.e-6));
            R dare_im = vsqrt(dare * dare_pr);
            dare_im = vmax(dare_im, R(1.e-6));
            const R dout_pr = fldout_pre;
            const R dout = (dout_pr + (gdt * dare_im) * dslp) *
                           (one / (one + ((gdtn2 * vabs(dout_pr)) * vpow(dflw_im, R(K_P4O3))) * (one / dare_im)));
            const R thr = river ? R(1.e-5) : R(K_1EM5F);
            const bool mask = (dflw_im > thr) && (dare > thr);
            const R f = mask ? dout : zero;
            q_f = (f * q_r > zero) ? f : zero;
        }
        // ---- reverse-flow storage limiter, river cells: outflw:175-182
        if (river) {
            R dout = ((-q_r) - q_f) * R(I.dt);
            dout = vmax(dout, R(K_1EM10F));
            R rate = (R(0.05) * storge) / dout;
            rate = vmin(rate, one);
            q_r = q_r * rate;
            q_f = q_f * rate;
        }
    }
    // ---- reservoir operation: damout:316-498 (Yamazaki & Funato or Hanazaki-2022 rule)
    if (A.ldamout) {
        const int code = M.i1dam[i];
        if (code > 0) {
            q_r = zero; q_f = zero;
            double inf = (double)S.runoff[s];                       // :447
            for (int k = M.up_ptr[i]; k < M.up_ptr[i + 1]; ++k) {   // :456-464, ascending
                const int u = M.up_idx[k];
                const int cu = M.i1dam[u];
                if (cu == 10 || cu == 11) {
                    const int su = sidx<MB>(M, u, m);
                    inf = (inf + (double)S.rivout[su]) + (double)S.fldout[su];
                }
            }
            S.daminf[s] = inf;
            const double dt = I.dt;
            if (code == 10 && river) {                               // :471-493
                const int j = nxt;
                double dslope = ((double)M.elevtn[i] - (double)M.elevtn[j]) * (1.0 / (double)M.nxtdst[i]);
                dslope = vmax(dslope, A.pminslp);
                const double rd = (double)rivdph;
                const double dvel = ((1.0 / (double)M.rivman[i]) * sqrt(dslope)) * pow(rd, K_2O3F);
                const double darea = (double)M.rivwth[i] * rd;
                vel = R(dvel);
                double q = darea * dvel;
                q_r = R(vmin(q, S.rivsto[s] / dt));
                const double dslope_f = vmin(0.005, dslope);
                const double fd = (double)S.flddph[s];
                const double dvel_f = ((1.0 / A.pmanfld) * sqrt(dslope_f)) * pow(fd, K_2O3F);
                double dare_f = S.fldsto[s] * (1.0 / (double)M.rivlen[i]);
                dare_f = vmax(dare_f - fd * (double)M.rivwth[i], 0.0);
                const double qf = dare_f * dvel_f;
                q_f = R(vmin(qf, S.fldsto[s] / dt));
            }
            const int d = M.dam_of_cell[i];
            if (d >= 0 && M.damstat[d] > 0) {                        // :344-419
                const double vol = S.damsto[s];
                const double cv = M.dam_convol[d], av = M.dam_adjvol[d], ev = M.dam_emevol[d];
                const double qn = M.dam_qn[d], qf = M.dam_qf[d], qa = M.dam_qa[d];
                double out = 0.0;
                if (M.ldamh22) {
                    out = dam_rule_h22(vol, inf, M.dam_norvol[d], cv, ev, qn, qf, M.dam_r_volupa[d]);
                } else if (vol <= cv) {
                    out = qn * sqrt(vol / cv);
                } else if (vol > cv && vol <= av) {
                    const double xx = (vol - cv) / (av - cv);
                    out = qn + (xx * (xx * xx)) * (qa - qn);
                } else if (vol > av && vol <= ev) {
                    if (inf >= qf) {
                        out = qn + ((vol - cv) / (ev - cv)) * (inf - qn);
                        const double tmp = qa + pow((vol - av) / (ev - av), K_0P1F) * (qf - qa);
                        out = vmax(out, tmp);
                    } else {
                        out = qa + pow((vol - av) / (ev - av), K_0P1F) * (qf - qa);
                    }
                } else if (vol > ev) {
                    out = (inf >= qf) ? inf : qf;
                }
                out = vmin(vmin(out, vol / dt), (S.rivsto[s] + S.fldsto[s]) / dt);
                out = vmax(out, 0.0);
                q_r = R(out);
                q_f = zero;
            }
        }
    }
    S.qr_u[s] = q_r;
    S.qf_u[s] = q_f;
    if (I.last) S.rivvel[s] = vel;                                  // output-only diagnostic
    // CMF_DIAG_AVEMAX_ADPSTP (diag:113) for RIVVEL, final after OUTFLW/DAMOUT
    R va = (I.first ? zero : p_rivvel_avg) + vel * R(I.dt);
    if (I.last) va = va / R(I.nadd);
    S.rivvel_avg[s] = va;
}

// phase 1b: bifurcation path flow, pthout:20-97 / levee:497-603
template <class R, bool MB>
__device__ __forceinline__ void pthflow_path(const DevMap& M, const DevState<R>& S, const StepArgs& A, int q) {
    const Slot x = path_slot<MB>(M, q);
    const InstView I = inst_view(S.inst, x.inst, A.it);
    if (!I.active) return;
    const int p = x.c, m = x.m;
    const int np = M.npth, nsp = M.nsp, nl = M.npthlev;
    const int ip = M.pth_upst[p], jp = M.pth_down[p];
    bool skip = ip < 0 || jp < 0;
    if (!skip) skip = A.llevee ? (M.i2mask[ip] == 1 || M.i2mask[jp] == 1) : (M.i2mask[ip] > 0 || M.i2mask[jp] > 0);
    if (skip) {
        for (int l = 0; l < nl; ++l) S.pthflw_u[(size_t)l * nsp + q] = R(0);
        S.pthflwsum_u[q] = R(0);
        return;
    }
    const int si_ = sidx<MB>(M, ip, m), sj_ = sidx<MB>(M, jp, m);
    const R zero = R(0), one = R(1);
    const R gdt = R(I.gdt);
    // Surface terms; with R = float in difference form (see outflow_cell): each
    // surface height relative to a level elevation is a depth plus the exact FP64
    // elevation difference rounded once.
    constexpr bool kDiff = sizeof(R) == sizeof(float);
    const double ri_d = rivelv_of(M, ip), rj_d = rivelv_of(M, jp);
    const R di = S.rivdph[si_], dj = S.rivdph[sj_];
    const R dpi = S.rivdph_pre[si_], dpj = S.rivdph_pre[sj_];
    const R ri = R(ri_d), rj = R(rj_d);
    const R si = ri + di, sj = rj + dj;
    const R pi = ri + dpi, pj = rj + dpj;
    R dslp = (kDiff ? (di - dj) + R(ri_d - rj_d) : si - sj) / R(M.pth_dst[p]);
    dslp = vmax(R(-0.005), vmin(R(0.005), dslp));
    R li = si, lj = sj;                                              // D2SFCELV_LEV
    bool lev_i = false, lev_j = false;
    R ldi = zero, ldj = zero;
    if (A.llevee) {
        const int a = M.lev_idx[ip], b = M.lev_idx[jp];
        const int nlc = M.nlevcell;
        lev_i = a >= 0 && M.lev_tab[nlc + a] < 1.0;
        lev_j = b >= 0 && M.lev_tab[nlc + b] < 1.0;
        if (lev_i) { ldi = S.levdph[si_]; li = vmin(R(M.elevtn[ip]) + ldi, si); }
        if (lev_j) { ldj = S.levdph[sj_]; lj = vmin(R(M.elevtn[jp]) + ldj, sj); }
    }
    R sum = zero;
    for (int l = 0; l < nl; ++l) {
        const double elv_d = M.pth_elv[(size_t)l * np + p];
        const R elv = R(elv_d);
        // max(si, sj) - elv, max(pi, pj) - elv and max(li, lj) - elv
        R hs, hp, hl;
        if (kDiff) {
            const R ai = R(ri_d - elv_d), aj = R(rj_d - elv_d);
            hs = vmax(di + ai, dj + aj);
            hp = vmax(dpi + ai, dpj + aj);
            const R lvi = lev_i ? vmin(ldi + R((double)M.elevtn[ip] - elv_d), di + ai) : di + ai;
            const R lvj = lev_j ? vmin(ldj + R((double)M.elevtn[jp] - elv_d), dj + aj) : dj + aj;
            hl = vmax(lvi, lvj);
        } else {
            hs = vmax(si, sj) - elv;
            hp = vmax(pi, pj) - elv;
            hl = vmax(li, lj) - elv;
        }
        R dflw, dflw_im;
        if (!A.llevee || l == 0) {
            dflw = vmax(hs, zero);
            R dflw_pr = vmax(hp, zero);
            dflw_im = vsqrt(dflw * dflw_pr);
            if (A.llevee) { if (dflw_im <= zero) dflw_im = dflw; }
            else dflw_im = vmax(dflw_im, vsqrt(dflw * R(0.01)));
        } else {
            dflw = vmax(hl, zero);
            dflw_im = dflw;
        }
        R qq = zero;
        if (dflw_im > R(1.e-5)) {
            const R w = R(M.pth_wth[(size_t)l * np + p]);
            const R man = R(M.pth_man[l]);
            const R pre = S.pthflw[(size_t)l * nsp + q];            // D1PTHFLW_PRE
            const R dout_pr = (A.llevee && l > 0) ? pre * (one / w) : pre / w;
            qq = (w * (dout_pr + (gdt * dflw_im) * dslp)) /
                 (one + ((gdt * (man * man)) * vabs(dout_pr)) * vpow(dflw_im, R(K_P7O3)));
        }
        S.pthflw_u[(size_t)l * nsp + q] = qq;
        sum = sum + qq;
    }
    if (sum != zero) {                                               // pthout:84-95
        R rate = (R(0.05) * vmin(S.storge[si_], S.storge[sj_])) / vabs(sum * R(I.dt));
        rate = vmin(rate, one);
        for (int l = 0; l < nl; ++l) S.pthflw_u[(size_t)l * nsp + q] = S.pthflw_u[(size_t)l * nsp + q] * rate;
        sum = sum * rate;
    }
    S.pthflwsum_u[q] = sum;
}

// ---------------------------------------------------------------------------
// phase 2: outflow volume and limiter rate, inflow_lspamat:410-451
template <class R, bool MB>
__device__ __forceinline__ void rate_cell(const DevMap& M, const DevState<R>& S, const StepArgs& A, int s) {
    const Slot x = cell_slot<MB>(M, s);
    const int i = x.c, m = x.m;
    const InstView I = inst_view(S.inst, x.inst, A.it);
    if (MB && !I.active) return;
    const R qr = S.qr_u[s], qf = S.qf_u[s];
    const int k0 = M.up_ptr[i], k1 = M.up_ptr[i + 1];
    // RIVSTO + FLDSTO, issued ahead of the gathers. Without levees STORGE holds
    // exactly this sum (written with RIVSTO and FLDSTO by the last stage
    // inversion, and converted to R as the rate is), so one array is read.
    const double sto = A.llevee ? S.rivsto[s] + S.fldsto[s] : (double)S.storge[s];
    if (!I.active) return;
    const double dt = I.dt;
    double so = (double)vmax(qr, R(0)) + (double)vmax(qf, R(0));
    // the first two upstream cells (most confluences) are gathered together, the
    // sum keeps the ascending order of the list
    if (k0 < k1) {
        const bool two = k0 + 1 < k1;
        const int s0 = sidx<MB>(M, M.up_idx[k0], m);
        const int s1 = two ? sidx<MB>(M, M.up_idx[k0 + 1], m) : s0;
        const R q0 = S.qr_u[s0], f0_ = S.qf_u[s0], q1 = S.qr_u[s1], f1_ = S.qf_u[s1];
        so = (so + (double)vmax(-q0, R(0))) + (double)vmax(-f0_, R(0));
        if (two) so = (so + (double)vmax(-q1, R(0))) + (double)vmax(-f1_, R(0));
        for (int k = k0 + 2; k < k1; ++k) {
            const int su = sidx<MB>(M, M.up_idx[k], m);
            so = (so + (double)vmax(-S.qr_u[su], R(0))) + (double)vmax(-S.qf_u[su], R(0));
        }
    }
    so = so * dt;
    if (A.lpthout) {
        for (int k = M.pout_ptr[i]; k < M.pout_ptr[i + 1]; ++k)
            so = so + (double)vmax(S.pthflwsum_u[sidx<MB>(M, M.pout_idx[k], m)], R(0)) * dt;
        for (int k = M.pinf_ptr[i]; k < M.pinf_ptr[i + 1]; ++k)
            so = so + (double)vmax(-S.pthflwsum_u[sidx<MB>(M, M.pinf_idx[k], m)], R(0)) * dt;
    }
    const R dout = vmax(R(so), R(1.e-10));
    const R dsto = R(sto);
    S.rate[s] = vmin(dsto / dout, R(1));
}

// ---------------------------------------------------------------------------
// phase 3: limit, gather inflow, dam balance, storage update, stage, averages
template <class R, bool MB>
__device__ __forceinline__ R limited_path_sum(const DevMap& M, const DevState<R>& S, int p, int m) {
    const R s = S.pthflwsum_u[sidx<MB>(M, p, m)];
    const int ip = M.pth_upst[p], jp = M.pth_down[p];
    if (M.i2mask[ip] > 0 || M.i2mask[jp] > 0) return s;               // inflow_lspamat:501
    return (s >= R(0)) ? s * S.rate[sidx<MB>(M, ip, m)] : s * S.rate[sidx<MB>(M, jp, m)];
}

template <class R, bool MB>
__device__ __forceinline__ void update_cell(const DevMap& M, const DevState<R>& S, const StepArgs& A,
                                            int s, double (&g)[NGLB]) {
    const Slot x = cell_slot<MB>(M, s);
    const int i = x.c, m = x.m;
    const InstView I = inst_view(S.inst, x.inst, A.it);
    if (MB && !I.active) return;
    const R zero = R(0);
    // every load that does not depend on a gathered index is issued first, so it
    // is in flight while the upstream and downstream gathers resolve
    const R qr = S.qr_u[s], qf = S.qf_u[s];
    const R ri = S.rate[s];
    const int nxt = M.next[i];
    const int k0 = M.up_ptr[i], k1 = M.up_ptr[i + 1];
    const R runoff = S.runoff[s];
    const R upsinf = A.lupsinf ? S.upsinf[s] : R(0);
    const double rs0 = S.rivsto[s], fs0 = S.fldsto[s];
    const R fldfrc_old = S.fldfrc[s];
    const Geom geo = load_geom(M, i);
    const float f_elevtn = M.elevtn[i];
    // LEVSTO and LEVDPH of a cell without a levee are zero after every stage
    // inversion (the one at the outer-step start included: it moves any LEVSTO of
    // such a cell into the river and floodplain), so only levee cells read and
    // write them here
    const int li = A.llevee ? M.lev_idx[i] : -1;
    const double levsto = li >= 0 ? S.levsto[s] : 0.0;
    Lev lev{};
    if (A.llevee) lev = load_lev_row(M, li);
    const int dam = A.ldamout ? M.dam_of_cell[i] : -1;
    // the mixed path forms OUTFLW's average from the RIVOUT and FLDOUT averages on
    // the last substep instead of accumulating it (it differs by FP32 rounding)
    constexpr bool kOutSum = sizeof(R) == sizeof(float);
    // The running averages and maxima are read where they are used, not up front:
    // fewer values live across the stage inversion, 5 waves per SIMD instead of 4
    // for the FP64 update (measured 5% faster than hoisting them).
    // DAMINF changes only in cells with I1DAM > 0 (damout:442-465); elsewhere it
    // keeps its initial value, zero unless the host set another (daminf_sparse
    // says it did not), so such cells neither read it nor accumulate its average
    const bool dam_inf = A.ldamout && (dam != -1 || !A.daminf_sparse);
    double daminf = 0.0, damsto = 0.0;
    if (dam_inf) daminf = S.daminf[s];
    if (!I.active) return;
    const double dt = I.dt;
    const bool f0 = I.first;
    // limiter, inflow_lspamat:456-473
    const R r = (nxt >= 0 && !(qr > zero)) ? S.rate[sidx<MB>(M, nxt, m)] : ri;
    const R qrl = qr * r, qfl = qf * r;
    // upstream gather, inflow_lspamat:481-489
    double rivinf = 0.0, fldinf = 0.0;
    if (k0 < k1) {
        // the first two upstream cells are gathered together (ascending order kept)
        const bool two = k0 + 1 < k1;
        const int s0 = sidx<MB>(M, M.up_idx[k0], m);
        const int s1 = two ? sidx<MB>(M, M.up_idx[k0 + 1], m) : s0;
        const R q0 = S.qr_u[s0], fu0 = S.qf_u[s0], rr0 = S.rate[s0];
        const R q1 = S.qr_u[s1], fu1 = S.qf_u[s1], rr1 = S.rate[s1];
        const R ru0 = (q0 > zero) ? rr0 : ri;
        rivinf = rivinf + (double)(q0 * ru0);
        fldinf = fldinf + (double)(fu0 * ru0);
        if (two) {
            const R ru1 = (q1 > zero) ? rr1 : ri;
            rivinf = rivinf + (double)(q1 * ru1);
            fldinf = fldinf + (double)(fu1 * ru1);
        }
        for (int k = k0 + 2; k < k1; ++k) {
            const int su = sidx<MB>(M, M.up_idx[k], m);
            const R qu = S.qr_u[su], fu = S.qf_u[su];
            const R ru = (qu > zero) ? S.rate[su] : ri;
            rivinf = rivinf + (double)(qu * ru);
            fldinf = fldinf + (double)(fu * ru);
        }
    }
    double pthout = 0.0;
    if (A.lpthout) {                                                  // :519-535
        for (int k = M.pout_ptr[i]; k < M.pout_ptr[i + 1]; ++k)
            pthout = pthout + (double)limited_path_sum<R, MB>(M, S, M.pout_idx[k], m);
        for (int k = M.pinf_ptr[i]; k < M.pinf_ptr[i + 1]; ++k)
            pthout = pthout - (double)limited_path_sum<R, MB>(M, S, M.pinf_idx[k], m);
    }
    // RIVINF/FLDINF/PTHOUT enter the storage update as the FP64 sums of the
    // limited fluxes (inflow_lspamat:538-544 rounds them to JPRB; with R = double
    // that is the identity). With R = float the FP64 sums are exact sums of the
    // FP32 outflows that the upstream cells subtract, so the update conserves mass
    // to FP64 rounding.
    const double rivinf_d = (double)R(rivinf), fldinf_d = (double)R(fldinf), pthout_d = (double)R(pthout);
    const double rinf = sizeof(R) == sizeof(double) ? rivinf_d : rivinf;
    const double finf = sizeof(R) == sizeof(double) ? fldinf_d : fldinf;
    const double pout = sizeof(R) == sizeof(double) ? pthout_d : pthout;
    const R pthout_r = R(pthout);
    if (I.last) { S.rivinf[s] = rivinf; S.fldinf[s] = fldinf; S.pthout[s] = pthout; }
    // CMF_DAMOUT_WATBAL, damout:521-536
    if (dam >= 0 && M.damstat[dam] != -9999) {
        damsto = S.damsto[s];
        const double inflow = (rinf + finf) + (double)runoff;
        const double outflw = (double)qrl + (double)qfl;
        S.damsto[s] = (damsto + inflow * dt) - outflw * dt;
    }
    // CALC_VARS_PRE, physics:254-276. RIVOUT_PRE/FLDOUT_PRE share the RIVOUT/FLDOUT
    // arrays; RIVDPH/RIVDPH_PRE are a buffer pair whose roles the host swaps after
    // this kernel, so the new depth goes to the buffer of the old RIVDPH_PRE.
    double rs = rs0, fs = fs0;
    S.rivout[s] = qrl;
    S.fldout[s] = qfl;
    S.fldsto_pre[s] = R(fs);
    // CMF_CALC_STONXT, stonxt:54-98
    g[G_STOPRE] = rs + fs;
    g[G_RIVINF] = rinf * dt + finf * dt;
    g[G_RIVOUT] = ((double)qrl * dt + (double)qfl * dt) + pout * dt;
    rs = (rs + rinf * dt) - (double)qrl * dt;
    if (rs < 0.0) { fs = fs + rs; rs = 0.0; }
    fs = ((fs + finf * dt) - (double)qfl * dt) - pout * dt;
    if (fs < 0.0) { rs = vmax(rs + fs, 0.0); fs = 0.0; }
    g[G_STONXT] = rs + fs;
    const R outflw = qrl + qfl;
    R drivrof = (runoff * (R(1) - fldfrc_old)) * R(dt);
    const R dfldrof = (runoff * fldfrc_old) * R(dt);
    if (A.lupsinf) drivrof = drivrof + upsinf * R(dt);             // stonxt:81
    rs = rs + (double)drivrof;
    fs = fs + (double)dfldrof;
    double st = rs + fs;
    if (A.llevee) st = st + levsto;
    g[G_STONEW] = (double)R(st);
    // CMF_PHYSICS_FLDSTG
    Stage o;
    if (A.llevee) stage_levee(M, i, geo, lev, rs, fs, levsto, o);
    else stage_def(M, i, geo, rs, fs, o);
    S.rivsto[s] = o.rivsto;
    S.fldsto[s] = o.fldsto;
    const R rivdph = R(o.rivdph);
    S.rivdph_next[s] = rivdph;
    S.flddph[s] = R(o.flddph);
    S.fldfrc[s] = R(o.fldfrc);
    if (I.last) {                                                   // output-only diagnostics
        S.fldare[s] = R(o.fldare);
        S.sfcelv[s] = R((double)f_elevtn - geo.rivhgt) + rivdph;
        S.outflw[s] = outflw;
    }
    double stg = o.rivsto + o.fldsto;
    if (A.llevee) {
        if (li >= 0) {
            S.levsto[s] = o.levsto;
            S.levdph[s] = R(o.levdph);
        }
        stg = stg + o.levsto;
    }
    const R storge = R(stg);
    S.storge[s] = storge;
    g[G_STOPRE2] = A.llevee ? (rs + fs) + levsto : rs + fs;
    g[G_STONEW2] = stg;
    g[G_RIVSTO] = o.rivsto;
    g[G_FLDSTO] = o.fldsto;
    g[G_FLDARE] = o.fldare;
    g[G_LEVSTO] = A.llevee ? o.levsto : 0.0;
    // CMF_DIAG_AVEMAX_ADPSTP, diag:108-124, and GETAVE :179-186 on the last substep
    const R p_ravg = S.rivout_avg[s], p_favg = S.fldout_avg[s];
    const R p_oavg = kOutSum ? zero : S.outflw_avg[s];
    const R p_omax = S.outflw_max[s], p_dmax = S.rivdph_max[s], p_smax = S.storge_max[s];
    const R p_pavg = A.lpthout ? S.pthout_avg[s] : zero;
    const R p_davg = dam_inf ? S.daminf_avg[s] : zero;
    const R rdt = R(dt);
    R a1 = (f0 ? zero : p_ravg) + qrl * rdt;
    R a2 = (f0 ? zero : p_favg) + qfl * rdt;
    R a3 = (f0 ? zero : p_oavg) + outflw * rdt;
    if (I.last) {
        const R nadd = R(I.nadd);
        a1 = a1 / nadd; a2 = a2 / nadd; a3 = kOutSum ? a1 + a2 : a3 / nadd;
        // RUNOFF is constant over the outer step: its NT additions are replayed in
        // registers, in the same order, instead of a read-modify-write per substep
        R a5 = zero;
        for (int k = 0; k < I.nt; ++k) a5 = a5 + runoff * rdt;
        S.runoff_avg[s] = a5 / nadd;
    }
    S.rivout_avg[s] = a1; S.fldout_avg[s] = a2;
    if (!kOutSum || I.last) S.outflw_avg[s] = a3;
    if (A.lpthout) {
        R a4 = (f0 ? zero : p_pavg) + pthout_r * rdt;
        if (I.last) a4 = a4 / R(I.nadd);
        S.pthout_avg[s] = a4;
    } else if (f0) {
        S.pthout_avg[s] = zero;                                     // PTHOUT = 0 without paths
    }
    S.outflw_max[s] = vmax(f0 ? zero : p_omax, vabs(outflw));
    S.rivdph_max[s] = vmax(f0 ? zero : p_dmax, rivdph);
    S.storge_max[s] = vmax(f0 ? zero : p_smax, storge);
    if (dam_inf) {
        R da = (f0 ? zero : p_davg) + R(daminf * dt);
        if (I.last) da = da / R(I.nadd);
        S.daminf_avg[s] = da;
    } else if (A.ldamout && f0) {
        S.daminf_avg[s] = zero;                                     // 0 + 0*DT, unchanged after
    }
}

template <class R, bool MB>
__device__ __forceinline__ void update_path(const DevMap& M, const DevState<R>& S, const StepArgs& A, int q) {
    const Slot x = path_slot<MB>(M, q);
    const InstView I = inst_view(S.inst, x.inst, A.it);
    if (!I.active) return;
    const int p = x.c, m = x.m;
    const int nsp = M.nsp, nl = M.npthlev;
    const int ip = M.pth_upst[p], jp = M.pth_down[p];
    const bool lim = !(ip < 0 || jp < 0) && !(M.i2mask[ip] > 0 || M.i2mask[jp] > 0);
    const R zero = R(0);
    const R rdt = R(I.dt);
    const R rate_i = lim ? S.rate[sidx<MB>(M, ip, m)] : zero;
    const R rate_j = lim ? S.rate[sidx<MB>(M, jp, m)] : zero;
    R sumavg = zero;
    for (int l = 0; l < nl; ++l) {
        R qq = S.pthflw_u[(size_t)l * nsp + q];
        if (lim) qq = (qq >= zero) ? qq * rate_i : qq * rate_j;       // inflow_lspamat:503-509
        S.pthflw[(size_t)l * nsp + q] = qq;                          // D1PTHFLW == D1PTHFLW_PRE
        R a = (I.first ? zero : S.pthflw_avg[(size_t)l * nsp + q]) + qq * rdt;
        if (I.last) {
            a = a / R(I.nadd);
            sumavg = sumavg + a;                                     // diag:206-213
        }
        S.pthflw_avg[(size_t)l * nsp + q] = a;
    }
    R sm = S.pthflwsum_u[q];
    if (lim) sm = (sm >= zero) ? sm * rate_i : sm * rate_j;          // :511-515
    S.pthflwsum[q] = sm;
    if (I.last) S.pthflwsum_avg[q] = sumavg;
}

// ---------------------------------------------------------------------------
// tiles: a tile is BS consecutive slots; cell tiles first, then path tiles
template <class R, int BS, bool MB, bool KINE>
__device__ __forceinline__ void outflow_tile(const DevMap& M, const DevState<R>& S, const StepArgs& A, int t,
                                             int ntc) {
    if (t < ntc) {
        const int s = t * BS + threadIdx.x;
        if (s < M.ns) outflow_cell<R, MB, KINE>(M, S, A, s);
    } else {
        const int q = (t - ntc) * BS + threadIdx.x;
        if (q < M.nsp) pthflow_path<R, MB>(M, S, A, q);
    }
}

template <class R, int BS, bool MB>
__device__ __forceinline__ void rate_tile(const DevMap& M, const DevState<R>& S, const StepArgs& A, int t) {
    const int s = t * BS + threadIdx.x;
    if (s < M.ns) rate_cell<R, MB>(M, S, A, s);
}

template <class R, int BS, bool MB>
__device__ __forceinline__ void update_tile(const DevMap& M, const DevState<R>& S, const StepArgs& A, int t, int ntc,
                                            double* lds) {
    if (t < ntc) {
        const int s = t * BS + threadIdx.x;
        double g[NGLB];
#pragma unroll
        for (int k = 0; k < NGLB; ++k) g[k] = 0.0;
        if (s < M.ns) update_cell<R, MB>(M, S, A, s, g);
        // budget sums are kept for single-instance handles only (MB = false), so
        // the ensemble kernels carry no budget registers
        if (!MB && A.budget) {
            block_sum_comp<BS, NGLB>(g, lds);
            // per-substep slot: reduced once for all substeps at the end of the outer step
            if (threadIdx.x < NGLB)
                S.partials[((size_t)A.prow * ntc + t) * NGLB + threadIdx.x] = lds[threadIdx.x * (BS / 64)];
            __syncthreads();
        }
    } else {
        const int q = (t - ntc) * BS + threadIdx.x;
        if (q < M.nsp) update_path<R, MB>(M, S, A, q);
    }
}

// The mixed-precision ensemble outflow of the default scheme is held to 64 VGPRs
// (8 waves per SIMD) and spills nothing; the single-model forms and the
// kinematic / slope-mixed ensemble form would spill under that bound and keep the
// compiler's choice.
template <class R, int BS, bool MB, bool KINE>
__global__ __launch_bounds__(BS, (MB && sizeof(R) == 4 && !KINE) ? 8 : 1) void k_outflow(DevMap M, DevState<R> S, StepArgs A, int ntc, int nt_all) {
    const int t = tile_of(blockIdx.x, gridDim.x, nt_all, A.xcd);
    if (t < nt_all) outflow_tile<R, BS, MB, KINE>(M, S, A, t, ntc);
}

template <class R, int BS, bool MB>
__global__ __launch_bounds__(BS) void k_rate(DevMap M, DevState<R> S, StepArgs A, int ntc) {
    const int t = tile_of(blockIdx.x, gridDim.x, ntc, A.xcd);
    if (t < ntc) rate_tile<R, BS, MB>(M, S, A, t);
}

// at most 96 VGPRs (5 waves per SIMD): the FP64 single-model update needs 97
// without the bound and spills nothing to memory with it. The mixed-precision
// ensemble update fits 80 (6 waves; 82 without the bound), again without spills;
// the FP64 ensemble update would spill at 6.
template <class R, int BS, bool MB>
__global__ __launch_bounds__(BS, (MB && sizeof(R) == 4) ? 6 : 5) void k_update(DevMap M, DevState<R> S, StepArgs A, int ntc, int nt_all) {
    __shared__ double lds[2 * NGLB * (BS / 64)];
    const int t = tile_of(blockIdx.x, gridDim.x, nt_all, A.xcd);
    if (t < nt_all) update_tile<R, BS, MB>(M, S, A, t, ntc, lds);
}

// ---------------------------------------------------------------------------
// Persistent substep: blocks loop over the tiles of each phase and meet at a
// grid barrier between phases. Launched cooperatively (all blocks resident).
// The barrier wait is bounded: on time-out it raises S.err[0] and lets the
// kernel finish; the host turns the flag into an error.
constexpr long long BARRIER_SPIN_LIMIT = 1ll << 22;

__device__ __forceinline__ void grid_barrier(unsigned* bar, int32_t* err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();                                           // publish this block's writes
        const unsigned arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > BARRIER_SPIN_LIMIT) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// tiles [lo, hi) of block b in a persistent grid of G blocks
__device__ __forceinline__ void tile_range(int b, int G, int ntiles, int xcd, int& first, int& step, int& end) {
    if (xcd && (G & 7) == 0) {
        const int x = b & 7, per = G >> 3;
        const int lo = (int)(((long long)ntiles * x) >> 3), hi = (int)(((long long)ntiles * (x + 1)) >> 3);
        first = lo + (b >> 3); step = per; end = hi;
    } else {
        first = b; step = G; end = ntiles;
    }
}

template <class R, int BS, bool MB>
__global__ __launch_bounds__(BS) void k_substep(DevMap M, DevState<R> S, StepArgs A, int ntc_o, int nt_o,
                                                int ntc_r, int ntc_u, int nt_u) {
    __shared__ double lds[2 * NGLB * (BS / 64)];
    int t0, dt, te;
    tile_range(blockIdx.x, gridDim.x, nt_o, A.xcd, t0, dt, te);
    for (int t = t0; t < te; t += dt) {
        if (A.lkine || A.lslpmix) outflow_tile<R, BS, MB, true>(M, S, A, t, ntc_o);
        else outflow_tile<R, BS, MB, false>(M, S, A, t, ntc_o);
    }
    grid_barrier(S.bar, S.err);
    tile_range(blockIdx.x, gridDim.x, ntc_r, A.xcd, t0, dt, te);
    for (int t = t0; t < te; t += dt) rate_tile<R, BS, MB>(M, S, A, t);
    grid_barrier(S.bar, S.err);
    tile_range(blockIdx.x, gridDim.x, nt_u, A.xcd, t0, dt, te);
    for (int t = t0; t < te; t += dt) update_tile<R, BS, MB>(M, S, A, t, ntc_u, lds);
}

// deterministic second-level reduction of the per-tile budget sums, compensated:
// block b reduces substep b of the outer step
constexpr int RBS = 256;
__global__ __launch_bounds__(RBS) void k_reduce_budget(const double* partials, int nb, double* out) {
    __shared__ double lds[2 * NGLB * (RBS / 64)];
    const double* src = partials + (size_t)blockIdx.x * nb * NGLB;
    double v[NGLB], lo[NGLB];
#pragma unroll
    for (int k = 0; k < NGLB; ++k) { v[k] = 0.0; lo[k] = 0.0; }
    for (int b = threadIdx.x; b < nb; b += RBS)
#pragma unroll
        for (int k = 0; k < NGLB; ++k) {
            double sh, e;
            two_sum(v[k], src[(size_t)b * NGLB + k], sh, e);
            v[k] = sh;
            lo[k] = lo[k] + e;
        }
#pragma unroll
    for (int k = 0; k < NGLB; ++k) v[k] = v[k] + lo[k];
    block_sum_comp<RBS, NGLB>(v, lds);
    if (threadIdx.x < NGLB) out[(size_t)blockIdx.x * NGLB + threadIdx.x] = lds[threadIdx.x * (RBS / 64)];
}

// ---------------------------------------------------------------------------
// outer-step start: CMF_PHYSICS_FLDSTG and the CFL bound of CALC_ADPSTP
__global__ void k_cfl_init(unsigned long long* cfl, int ninst, double dt_def) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ninst) cfl[k] = (unsigned long long)__double_as_longlong(dt_def);
}

constexpr double CFL_NONE = 1.0e300;            // no bound (masked cell, slot outside the map)

__device__ __forceinline__ unsigned long long cfl_bits(double c) {
    return (unsigned long long)__double_as_longlong(c);
}

// CMF_PHYSICS_FLDSTG of one slot; returns the cell's CFL bound of CALC_ADPSTP
// (physics:178-197) when with_cfl, else CFL_NONE; inst receives the slot's instance
template <class R, bool MB>
__device__ __forceinline__ double stage_slot(const DevMap& M, const DevState<R>& S, const StepArgs& A, int s,
                                             int with_cfl, int& inst) {
    const Slot x = cell_slot<MB>(M, s);
    const int i = x.c;
    inst = x.inst;
    Stage o;
    const double levsto = A.llevee ? S.levsto[s] : 0.0;
    const Geom geo = load_geom(M, i);
    if (A.llevee) stage_levee(M, i, geo, load_lev(M, i), S.rivsto[s], S.fldsto[s], levsto, o);
    else stage_def(M, i, geo, S.rivsto[s], S.fldsto[s], o);
    S.rivsto[s] = o.rivsto;
    S.fldsto[s] = o.fldsto;
    const R rivdph = R(o.rivdph);
    S.rivdph[s] = rivdph;
    S.flddph[s] = R(o.flddph);
    S.fldfrc[s] = R(o.fldfrc);
    S.fldare[s] = R(o.fldare);
    S.sfcelv[s] = R(rivelv_of(M, i)) + rivdph;
    double stg = o.rivsto + o.fldsto;
    if (A.llevee) {
        S.levsto[s] = o.levsto;
        S.levdph[s] = R(o.levdph);
        stg = stg + o.levsto;
    }
    S.storge[s] = R(stg);
    if (with_cfl && M.i2mask[i] == 0) {
        const double ddph = vmax((double)rivdph, 0.01);
        const double ddst = (M.next[i] >= 0) ? (double)M.nxtdst[i] : A.pdstmth;
        return (A.pcadp * ddst) * (1.0 / sqrt(A.pgrv * ddph));
    }
    return CFL_NONE;
}

// Single instance: one block per tile of BS slots; the block minimum of the CFL
// bound goes to cfl_part[block] (no atomics) and k_cfl_reduce takes the minimum
// of the blocks.
template <class R, int BS>
__global__ __launch_bounds__(BS) void k_stage(DevMap M, DevState<R> S, StepArgs A, int with_cfl) {
    __shared__ double lds[BS / 64];
    const int s = blockIdx.x * BS + threadIdx.x;
    double c = CFL_NONE;
    int inst;
    if (s < M.ns) c = stage_slot<R, false>(M, S, A, s, with_cfl, inst);
    if (!with_cfl) return;
    for (int off = 32; off > 0; off >>= 1) c = vmin(c, __shfl_xor(c, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = lds[0];
        for (int w = 1; w < BS / 64; ++w) b = vmin(b, lds[w]);
        S.cfl_part[blockIdx.x] = b;
    }
}

// DT_MIN = min(DT_DEF, bounds of all blocks) (physics:178, the minimum is exact
// in any order), written as the bits of the positive double
__global__ __launch_bounds__(1024) void k_cfl_reduce(const double* part, int nb, double dt_def,
                                                     unsigned long long* cfl) {
    __shared__ double lds[16];
    double c = dt_def;
    for (int b = threadIdx.x; b < nb; b += 1024) c = vmin(c, part[b]);
    for (int off = 32; off > 0; off >>= 1) c = vmin(c, __shfl_xor(c, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; ++w) c = vmin(c, lds[w]);
        cfl[0] = cfl_bits(vmin(c, lds[0]));
    }
}

// Several instances: block b runs the ch tiles b*ch .. b*ch+ch-1. A lane always
// holds the same member (BS is a multiple of NMEM), so it keeps the running
// minimum of its member over the cells of one domain and flushes it with an
// atomic minimum when the domain changes. At the end the lanes of one member
// are combined across the wavefront and across the block's wavefronts, so a
// block issues about NMEM atomics for ch*BS slots. Positive doubles order like
// their bit patterns, so the atomic minimum over the bits is the numeric one.
template <class R, int BS>
__global__ __launch_bounds__(BS) void k_stage_mb(DevMap M, DevState<R> S, StepArgs A, int with_cfl, int ch) {
    constexpr int NW = BS / 64;
    __shared__ double lds_c[NW * 64];
    __shared__ int lds_d[NW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = threadIdx.x & (M.nmem - 1);
    double c = CFL_NONE;
    int dcur = -1;
    const long long t0 = (long long)blockIdx.x * ch;
    for (int k = 0; k < ch; ++k) {
        const long long s64 = (t0 + k) * BS + threadIdx.x;
        if (s64 >= M.ns) break;
        int inst;
        const double cc = stage_slot<R, true>(M, S, A, (int)s64, with_cfl, inst);
        if (cc < CFL_NONE) {
            const int d = inst >> M.lgm;                              // inst = d * NMEM + m
            if (d != dcur) {
                if (dcur >= 0 && c < CFL_NONE) atomicMin(&S.cfl[dcur * M.nmem + m], cfl_bits(c));
                c = CFL_NONE;
                dcur = d;
            }
            c = vmin(c, cc);
        }
    }
    if (!with_cfl) return;
    const bool has = c < CFL_NONE;
    const unsigned long long bal = __ballot(has);
    if (bal == 0ull) {
        if (lane == 0) lds_d[wave] = -1;
    } else {
        const int d0 = __shfl(dcur, __ffsll((long long)bal) - 1, 64);
        if (__all(!has || dcur == d0)) {
            for (int off = M.nmem; off < 64; off <<= 1) c = vmin(c, __shfl_xor(c, off, 64));
            if (lane < M.nmem) lds_c[wave * 64 + lane] = c;
            if (lane == 0) lds_d[wave] = d0;
        } else {
            if (has) atomicMin(&S.cfl[dcur * M.nmem + m], cfl_bits(c));
            if (lane == 0) lds_d[wave] = -1;
        }
    }
    __syncthreads();
    if (wave == 0 && lane < M.nmem) {
        int dacc = -1;
        double cacc = CFL_NONE;
        for (int w = 0; w < NW; ++w) {
            const int d = lds_d[w];
            if (d < 0) continue;
            const double v = lds_c[w * 64 + lane];
            if (d != dacc) {
                if (dacc >= 0 && cacc < CFL_NONE) atomicMin(&S.cfl[dacc * M.nmem + lane], cfl_bits(cacc));
                dacc = d;
                cacc = v;
            } else {
                cacc = vmin(cacc, v);
            }
        }
        if (dacc >= 0 && cacc < CFL_NONE) atomicMin(&S.cfl[dacc * M.nmem + lane], cfl_bits(cacc));
    }
}

// Total storage of every instance (P2RIVSTO + P2FLDSTO (+ P2LEVSTO)), in two
// deterministic passes: block (j, k) sums chunk j of the cells of instance k's
// domain (each thread its strided subset in order, then the fixed block tree)
// into part[k * nch + j]; k_inst_storage_final adds an instance's nch chunks in
// order.
constexpr int ISBS = 256;
__global__ __launch_bounds__(ISBS) void k_inst_storage_part(const double* rivsto, const double* fldsto,
                                                            const double* levsto, const int32_t* dom_off, int lgm,
                                                            int nch, double* part) {
    __shared__ double lds[ISBS / 64];
    const int j = blockIdx.x, k = blockIdx.y;
    const int d = k >> lgm, m = k & ((1 << lgm) - 1);
    const long long c0 = dom_off[d], nc = (long long)dom_off[d + 1] - c0;
    const long long lo = c0 + nc * j / nch, hi = c0 + nc * (j + 1) / nch;
    double v[1] = {0.0};
    for (long long c = lo + threadIdx.x; c < hi; c += ISBS) {
        const size_t sl = ((size_t)c << lgm) | (size_t)m;
        double x = rivsto[sl] + fldsto[sl];
        if (levsto) x = x + levsto[sl];
        v[0] += x;
    }
    block_sum<ISBS, 1>(v, lds);
    if (threadIdx.x == 0) part[(size_t)k * nch + j] = lds[0];
}

__global__ void k_inst_storage_final(const double* part, int nch, int ninst, double* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    double s = 0.0;
    for (int j = 0; j < nch; ++j) s += part[(size_t)k * nch + j];
    out[k] = s;
}

// instances that idled an odd number of substeps hold RIVDPH in the buffer the
// host now calls RIVDPH_PRE: swap the pair back for their slots
template <class R, bool MB>
__global__ void k_swapfix(DevMap M, DevState<R> S) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M.ns) return;
    const Slot x = cell_slot<MB>(M, s);
    if (!S.inst[x.inst].swapfix) return;
    const R a = S.rivdph[s];
    S.rivdph[s] = S.rivdph_pre[s];
    S.rivdph_pre[s] = a;
}

// CMF_CALC_OUTPRE (src/cmf_opt_outflw_mod.F90:308-421): previous-step flows of a
// storage-only restart from the current stage. Per cell slot: SFCELV = RIVELV +
// RIVDPH, RIVDPH_PRE = RIVDPH, and the diffusive-wave (Manning) discharge
// AREA * ((1/n) * DFLW^(2/3) * |S|^(1/2)) signed by the slope S, river channel
// (slope to the downstream surface over NXTDST; mouths: to DWNELV over PDSTMTH)
// and floodplain (slope clamped to +-0.005), written as RIVOUT_PRE / FLDOUT_PRE.
// Evaluated in FP64 in the reference's statement order and rounded once to R.
__device__ __forceinline__ double outpre_q(double area, double rman_inv, double dflw, double dslope) {
    const double q = area * ((rman_inv * pow(dflw, K_2O3F)) * sqrt(fabs(dslope)));
    return dslope < 0.0 ? -q : q;
}

template <class R, bool MB>
__global__ void k_outpre(DevMap M, DevState<R> S, int lfldout, double pmanfld, double pdstmth) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M.ns) return;
    const Slot x = cell_slot<MB>(M, s);
    const int i = x.c;
    const int nxt = M.next[i];
    const double rivelv = rivelv_of(M, i);
    const double rd = (double)S.rivdph[s];
    const double sfc = rivelv + rd;                                      // :324-327
    double dslope, dflw, dflw_f;
    if (nxt >= 0) {                                                      // :331-363
        const double sfc_j = rivelv_of(M, nxt) + (double)S.rivdph[sidx<MB>(M, nxt, x.m)];
        const double dsfcmax = vmax(sfc, sfc_j);
        dslope = (sfc - sfc_j) * (1.0 / (double)M.nxtdst[i]);
        dflw = dsfcmax - rivelv;
        dflw_f = vmax(dsfcmax - (double)M.elevtn[i], 0.0);
    } else {                                                             // :368-393
        dslope = (sfc - M.dwnelv[i]) * (1.0 / pdstmth);
        dflw = rd;
        dflw_f = vmax(sfc - (double)M.elevtn[i], 0.0);
    }
    const double dslope_f = vmax(-0.005, vmin(0.005, dslope));
    const double rivwth = (double)M.rivwth[i];
    const double darea = rivwth * dflw;
    const double qr = darea > K_1EM5F ? outpre_q(darea, 1.0 / (double)M.rivman[i], dflw, dslope) : 0.0;
    double dare_f = S.fldsto[s] * (1.0 / (double)M.rivlen[i]);
    dare_f = vmax(dare_f - (double)S.flddph[s] * rivwth, 0.0);
    double qf = dare_f > K_1EM5F ? outpre_q(dare_f, 1.0 / pmanfld, dflw_f, dslope_f) : 0.0;
    if (!lfldout) qf = 0.0;                                              // :416-418
    S.sfcelv[s] = R(sfc);
    S.rivdph_pre[s] = S.rivdph[s];
    S.rivout[s] = R(qr);                                                 // D2RIVOUT_PRE
    S.fldout[s] = R(qf);                                                 // D2FLDOUT_PRE
}

// the bifurcation levels of CMF_CALC_OUTPRE (:396-413), per path slot; a path
// with an endpoint outside the domain gets 0
template <class R, bool MB>
__global__ void k_outpre_path(DevMap M, DevState<R> S) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= M.nsp) return;
    const Slot x = path_slot<MB>(M, q);
    const int p = x.c;
    const int ip = M.pth_upst[p], jp = M.pth_down[p];
    const bool ok = ip >= 0 && jp >= 0;
    double si = 0.0, sj = 0.0, dslope = 0.0;
    if (ok) {
        si = rivelv_of(M, ip) + (double)S.rivdph[sidx<MB>(M, ip, x.m)];
        sj = rivelv_of(M, jp) + (double)S.rivdph[sidx<MB>(M, jp, x.m)];
        dslope = (si - sj) * (1.0 / M.pth_dst[p]);
    }
    const int np = M.npth;
    for (int l = 0; l < M.npthlev; ++l) {
        double qq = 0.0;
        if (ok) {
            double dflw = vmax(si, sj) - M.pth_elv[(size_t)l * np + p];
            dflw = vmax(dflw, 0.0);
            if (dflw > K_1EM5F) {
                qq = (M.pth_wth[(size_t)l * np + p] * dflw) * (((1.0 / M.pth_man[l]) * pow(dflw, K_2O3F)) * sqrt(fabs(dslope)));
                if (dslope < 0.0) qq = -qq;
            }
        }
        S.pthflw[(size_t)l * M.nsp + q] = R(qq);                         // D1PTHFLW_PRE
    }
}

// Host <-> device transfer of per-slot cell arrays: the host array (member-major
// planes of NSEQALL values in NSEQ order, element type H) is copied to the device
// as it is and permuted into the member-minor slots in device cell order here
// (perm[p] = NSEQ index of device cell p; NULL = identity), converted to R with
// the same rounding as a host-side cast. shared: one plane for every member.
template <class R, class H>
__global__ void k_cells_from_host(R* dst, const H* src, const int32_t* perm, int n, int lgm, int ns, int shared) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int p = s >> lgm, m = s & ((1 << lgm) - 1);
    const size_t hi = (shared ? (size_t)0 : (size_t)m * (size_t)n) + (size_t)(perm ? perm[p] : p);
    dst[s] = (R)src[hi];
}

template <class R, class H>
__global__ void k_cells_to_host(H* dst, const R* src, const int32_t* perm, int n, int lgm, int ns) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int p = s >> lgm, m = s & ((1 << lgm) - 1);
    dst[(size_t)m * (size_t)n + (size_t)(perm ? perm[p] : p)] = (H)src[s];
}

// runoff[s] = base[c] * factor[inst] (multiplicative member perturbation)
template <class R, bool MB>
__global__ void k_scale_runoff(DevMap M, DevState<R> S, const double* base, const double* factor) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M.ns) return;
    const Slot x = cell_slot<MB>(M, s);
    S.runoff[s] = R(base[x.c] * factor[x.inst]);
}

// ROFF_INTERP (src/cmf_ctrl_forcing_mod.F90:757-790): runoff of a cell = sum over
// its input-matrix entries, in entry order, of grid value * INPA / DROFUNIT, for
// entries inside the grid (IXIN > 0, IXIN <= NXIN, IYIN <= NYIN) whose value is
// not RMIS. Every member slot of a cell receives the cell's value.
template <class R, bool MB>
__global__ void k_roff_interp(DevMap M, DevState<R> S, RoffMap Q, const double* grid) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M.ns) return;
    const int c = cell_slot<MB>(M, s).c;
    double acc = 0.0;
    for (int k = 0; k < Q.inpn; ++k) {
        const size_t e = (size_t)k * Q.n + c;
        const int ix = Q.inpx[e], iy = Q.inpy[e];
        if (ix > 0 && ix <= Q.nxin && iy >= 1 && iy <= Q.nyin) {
            const double v = grid[(size_t)(iy - 1) * Q.nxin + (ix - 1)];
            if (v != Q.rmis) acc = acc + (v * Q.inpa[e]) / Q.drofunit;
        }
    }
    S.runoff[s] = R(acc);
}

// CMF_DIAG_AVEMAX_OUTPUT (src/cmf_calc_diag_mod.F90:273-317): after an outer step
// of length dt, o += aAVG * dt for the averages and o = max(o, aMAX) for the
// maxima (|aMAX| for OUTFLW), per slot; paths per level and path slot.
template <class R>
__global__ void k_out_avemax(DevMap M, DevState<R> S, OutState O, double dt, int ldamout, int npth_slots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < M.ns) {
        O.rivout[s] = O.rivout[s] + (double)S.rivout_avg[s] * dt;
        O.fldout[s] = O.fldout[s] + (double)S.fldout_avg[s] * dt;
        O.rivvel[s] = O.rivvel[s] + (double)S.rivvel_avg[s] * dt;
        O.outflw[s] = O.outflw[s] + (double)S.outflw_avg[s] * dt;
        O.pthout[s] = O.pthout[s] + (double)S.pthout_avg[s] * dt;
        O.runoff[s] = O.runoff[s] + (double)S.runoff_avg[s] * dt;
        O.outflw_max[s] = fmax(O.outflw_max[s], fabs((double)S.outflw_max[s]));
        O.rivdph_max[s] = fmax(O.rivdph_max[s], (double)S.rivdph_max[s]);
        O.storge_max[s] = fmax(O.storge_max[s], (double)S.storge_max[s]);
        if (ldamout) O.daminf[s] = O.daminf[s] + (double)S.daminf_avg[s] * dt;
    }
    if (s < npth_slots) O.pthflw[s] = O.pthflw[s] + (double)S.pthflw_avg[s] * dt;
}

// CMF_DIAG_GETAVE_OUTPUT (src/cmf_calc_diag_mod.F90:323-356): averages / NADD_out
__global__ void k_out_getave(OutState O, int ns, int npth_slots, double nadd, int ldamout) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) {
        O.rivout[s] = O.rivout[s] / nadd;
        O.fldout[s] = O.fldout[s] / nadd;
        O.outflw[s] = O.outflw[s] / nadd;
        O.rivvel[s] = O.rivvel[s] / nadd;
        O.pthout[s] = O.pthout[s] / nadd;
        O.runoff[s] = O.runoff[s] / nadd;
        if (ldamout) O.daminf[s] = O.daminf[s] / nadd;
    }
    if (s < npth_slots) O.pthflw[s] = O.pthflw[s] / nadd;
}

// CMF_DIAG_RESET_OUTPUT (src/cmf_calc_diag_mod.F90:222-267): zero every accumulator
// (a kernel, not a memset, so that a captured graph may hold it)
__global__ void k_out_reset(OutState O, int ns, int npth_slots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) {
        O.rivout[s] = 0.0; O.fldout[s] = 0.0; O.outflw[s] = 0.0; O.rivvel[s] = 0.0;
        O.pthout[s] = 0.0; O.runoff[s] = 0.0; O.daminf[s] = 0.0;
        O.outflw_max[s] = 0.0; O.rivdph_max[s] = 0.0; O.storge_max[s] = 0.0;
    }
    if (s < npth_slots) O.pthflw[s] = 0.0;
}

// CMF_BOUNDARY_UPDATE (src/cmf_ctrl_boundary_mod.F90:267-305): D2DWNELV = D2ELEVTN
// (+ D2MEANSL with LMEANSL) (+ D2SEALEV with LSEALEV), in that order; NULL = off
__global__ void k_boundary(int n, const float* elevtn, const double* meansl, const double* sealev,
                           double* dwnelv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = (double)elevtn[i];
    if (meansl) d = d + meansl[i];
    if (sealev) d = d + sealev[i];
    dwnelv[i] = d;
}

// ---------------------------------------------------------------------------
// launch wrappers: MB (several instances) is chosen from the map, the block
// size from the launch configuration
static inline int cdiv(long long a, int b) { return (int)((a + b - 1) / b); }
static inline int pad8(int g, int xcd) { return xcd ? (g + 7) & ~7 : g; }
static inline bool multi(const DevMap& M) { return M.nmem > 1 || M.cell_dom != nullptr; }
constexpr int STAGE_MB_BLOCKS = 4096;

template <class R>
void launch_stage(const DevMap& M, const DevState<R>& S, const StepArgs& A, int with_cfl, int ninst, double dt_def,
                  hipStream_t st) {
    const int ntiles = cdiv(M.ns, 256);
    if (multi(M)) {
        if (with_cfl) hipLaunchKernelGGL(k_cfl_init, dim3(cdiv(ninst, 64)), dim3(64), 0, st, S.cfl, ninst, dt_def);
        // about 4096 blocks, each over ch consecutive tiles
        const int ch = ntiles > STAGE_MB_BLOCKS ? cdiv(ntiles, STAGE_MB_BLOCKS) : 1;
        hipLaunchKernelGGL((k_stage_mb<R, 256>), dim3(cdiv(ntiles, ch)), dim3(256), 0, st, M, S, A, with_cfl, ch);
    } else {
        hipLaunchKernelGGL((k_stage<R, 256>), dim3(ntiles), dim3(256), 0, st, M, S, A, with_cfl);
        if (with_cfl) hipLaunchKernelGGL(k_cfl_reduce, dim3(1), dim3(1024), 0, st, S.cfl_part, ntiles, dt_def, S.cfl);
    }
}

int tiles_cells(const DevMap& M, int bs) { return cdiv(M.ns, bs); }
int tiles_paths(const DevMap& M, const StepArgs& A, int bs) { return A.lpthout ? cdiv(M.nsp, bs) : 0; }

template <class R, int BS, bool MB>
static void launch_phases_bs(const DevMap& M, const DevState<R>& S, const StepArgs& A, int phase, hipStream_t st) {
    const int ntc = tiles_cells(M, BS), ntp = tiles_paths(M, A, BS);
    if (phase == 0)
        if (A.lkine || A.lslpmix)
            hipLaunchKernelGGL((k_outflow<R, BS, MB, true>), dim3(pad8(ntc + ntp, A.xcd)), dim3(BS), 0, st, M, S,
                               A, ntc, ntc + ntp);
        else
            hipLaunchKernelGGL((k_outflow<R, BS, MB, false>), dim3(pad8(ntc + ntp, A.xcd)), dim3(BS), 0, st, M, S,
                               A, ntc, ntc + ntp);
    else if (phase == 1)
        hipLaunchKernelGGL((k_rate<R, BS, MB>), dim3(pad8(ntc, A.xcd)), dim3(BS), 0, st, M, S, A, ntc);
    else
        hipLaunchKernelGGL((k_update<R, BS, MB>), dim3(pad8(ntc + ntp, A.xcd)), dim3(BS), 0, st, M, S, A, ntc,
                           ntc + ntp);
}

template <class R, bool MB>
static void launch_phase_mb(const DevMap& M, const DevState<R>& S, const StepArgs& A, int phase, int bs,
                            hipStream_t st) {
    switch (bs) {
        case 64: launch_phases_bs<R, 64, MB>(M, S, A, phase, st); break;
        case 128: launch_phases_bs<R, 128, MB>(M, S, A, phase, st); break;
        case 512: launch_phases_bs<R, 512, MB>(M, S, A, phase, st); break;
        default: launch_phases_bs<R, 256, MB>(M, S, A, phase, st); break;
    }
}

template <class R>
void launch_phase(const DevMap& M, const DevState<R>& S, const StepArgs& A, int phase, int bs, hipStream_t st) {
    if (multi(M)) launch_phase_mb<R, true>(M, S, A, phase, bs, st);
    else launch_phase_mb<R, false>(M, S, A, phase, bs, st);
}

template <class R, int BS, bool MB>
static const void* substep_fn() { return reinterpret_cast<const void*>(&k_substep<R, BS, MB>); }

template <class R, bool MB>
static const void* substep_kernel_mb(int bs) {
    switch (bs) {
        case 64: return substep_fn<R, 64, MB>();
        case 128: return substep_fn<R, 128, MB>();
        case 512: return substep_fn<R, 512, MB>();
        default: return substep_fn<R, 256, MB>();
    }
}

template <class R>
const void* substep_kernel(const DevMap& M, int bs) {
    return multi(M) ? substep_kernel_mb<R, true>(bs) : substep_kernel_mb<R, false>(bs);
}

template <class R>
hipError_t launch_substep(const DevMap& M, const DevState<R>& S, const StepArgs& A, int bs, int grid,
                          hipStream_t st) {
    const int ntc = tiles_cells(M, bs), ntp = tiles_paths(M, A, bs);
    int ntc_o = ntc, nt_o = ntc + ntp, ntc_r = ntc, ntc_u = ntc, nt_u = ntc + ntp;
    DevMap m = M;
    DevState<R> s = S;
    StepArgs a = A;
    void* args[] = {&m, &s, &a, &ntc_o, &nt_o, &ntc_r, &ntc_u, &nt_u};
    return hipLaunchCooperativeKernel(substep_kernel<R>(M, bs), dim3(grid), dim3(bs), args, 0, st);
}

template <class R>
void launch_swapfix(const DevMap& M, const DevState<R>& S, hipStream_t st) {
    hipLaunchKernelGGL((k_swapfix<R, true>), dim3(cdiv(M.ns, 256)), dim3(256), 0, st, M, S);
}

template <class R, class H>
void launch_cells_from_host(R* dst, const H* src, const int32_t* perm, int n, int lgm, int ns, int shared,
                            hipStream_t st) {
    hipLaunchKernelGGL((k_cells_from_host<R, H>), dim3(cdiv(ns, 256)), dim3(256), 0, st, dst, src, perm, n, lgm, ns,
                       shared);
}
template <class R, class H>
void launch_cells_to_host(H* dst, const R* src, const int32_t* perm, int n, int lgm, int ns, hipStream_t st) {
    hipLaunchKernelGGL((k_cells_to_host<R, H>), dim3(cdiv(ns, 256)), dim3(256), 0, st, dst, src, perm, n, lgm, ns);
}
#define CMF_INSTANTIATE_XFER(R, H)                                                                        \
    template void launch_cells_from_host<R, H>(R*, const H*, const int32_t*, int, int, int, int, hipStream_t); \
    template void launch_cells_to_host<R, H>(H*, const R*, const int32_t*, int, int, int, hipStream_t);
CMF_INSTANTIATE_XFER(double, double)
CMF_INSTANTIATE_XFER(double, float)
CMF_INSTANTIATE_XFER(float, double)
CMF_INSTANTIATE_XFER(float, float)

template <class R>
void launch_outpre(const DevMap& M, const DevState<R>& S, int lfldout, int lpthout, double pmanfld, double pdstmth,
                   hipStream_t st) {
    if (multi(M)) {
        hipLaunchKernelGGL((k_outpre<R, true>), dim3(cdiv(M.ns, 256)), dim3(256), 0, st, M, S, lfldout, pmanfld, pdstmth);
        if (lpthout && M.nsp > 0) hipLaunchKernelGGL((k_outpre_path<R, true>), dim3(cdiv(M.nsp, 256)), dim3(256), 0, st, M, S);
    } else {
        hipLaunchKernelGGL((k_outpre<R, false>), dim3(cdiv(M.ns, 256)), dim3(256), 0, st, M, S, lfldout, pmanfld, pdstmth);
        if (lpthout && M.nsp > 0) hipLaunchKernelGGL((k_outpre_path<R, false>), dim3(cdiv(M.nsp, 256)), dim3(256), 0, st, M, S);
    }
}

template <class R>
void launch_scale_runoff(const DevMap& M, const DevState<R>& S, const double* base, const double* factor,
                         hipStream_t st) {
    if (multi(M)) hipLaunchKernelGGL((k_scale_runoff<R, true>), dim3(cdiv(M.ns, 256)), dim3(256), 0, st, M, S, base, factor);
    else hipLaunchKernelGGL((k_scale_runoff<R, false>), dim3(cdiv(M.ns, 256)), dim3(256), 0, st, M, S, base, factor);
}

template <class R>
void launch_roff_interp(const DevMap& M, const DevState<R>& S, const RoffMap& Q, const double* grid,
                        hipStream_t st) {
    if (multi(M)) hipLaunchKernelGGL((k_roff_interp<R, true>), dim3(cdiv(M.ns, 256)), dim3(256), 0, st, M, S, Q, grid);
    else hipLaunchKernelGGL((k_roff_interp<R, false>), dim3(cdiv(M.ns, 256)), dim3(256), 0, st, M, S, Q, grid);
}

template <class R>
void launch_out_avemax(const DevMap& M, const DevState<R>& S, const OutState& O, double dt, int ldamout,
                       int npth_slots, hipStream_t st) {
    const int n = M.ns > npth_slots ? M.ns : npth_slots;
    hipLaunchKernelGGL((k_out_avemax<R>), dim3(cdiv(n, 256)), dim3(256), 0, st, M, S, O, dt, ldamout, npth_slots);
}

void launch_out_getave(const OutState& O, int ns, int npth_slots, double nadd, int ldamout, hipStream_t st) {
    const int n = ns > npth_slots ? ns : npth_slots;
    hipLaunchKernelGGL(k_out_getave, dim3(cdiv(n, 256)), dim3(256), 0, st, O, ns, npth_slots, nadd, ldamout);
}

void launch_out_reset(const OutState& O, int ns, int npth_slots, hipStream_t st) {
    const int n = ns > npth_slots ? ns : npth_slots;
    hipLaunchKernelGGL(k_out_reset, dim3(cdiv(n, 256)), dim3(256), 0, st, O, ns, npth_slots);
}

void launch_boundary(int n, const float* elevtn, const double* meansl, const double* sealev, double* dwnelv,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_boundary, dim3(cdiv(n, 256)), dim3(256), 0, st, n, elevtn, meansl, sealev, dwnelv);
}

// 1 in *flag when a cell with I1DAM == 0 (dam_of_cell == -1) holds a nonzero
// DAMINF (slots member-minor, cell = slot >> lgm)
__global__ void k_daminf_nonzero(const int32_t* dam_of_cell, const double* daminf, int ns, int lgm, int32_t* flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns && dam_of_cell[s >> lgm] == -1 && daminf[s] != 0.0) *flag = 1;
}

void launch_daminf_nonzero(const int32_t* dam_of_cell, const double* daminf, int ns, int lgm, int32_t* flag,
                           hipStream_t st) {
    if (ns > 0) hipLaunchKernelGGL(k_daminf_nonzero, dim3(cdiv(ns, 256)), dim3(256), 0, st, dam_of_cell, daminf, ns,
                                   lgm, flag);
}

void launch_inst_storage(const double* rivsto, const double* fldsto, const double* levsto, const int32_t* dom_off,
                         int lgm, int ninst, int nch, double* part, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_inst_storage_part, dim3(nch, ninst), dim3(ISBS), 0, st, rivsto, fldsto, levsto, dom_off, lgm,
                       nch, part);
    hipLaunchKernelGGL(k_inst_storage_final, dim3(cdiv(ninst, 256)), dim3(256), 0, st, part, nch, ninst, out);
}

void launch_budget(const double* partials, int ntiles, int nrun, double* glb_hist, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_budget, dim3(nrun), dim3(RBS), 0, st, partials, ntiles, glb_hist);
}

#define CMF_INSTANTIATE(R)                                                                                     \
    template void launch_stage<R>(const DevMap&, const DevState<R>&, const StepArgs&, int, int, double,        \
                                  hipStream_t);                                                                \
    template void launch_phase<R>(const DevMap&, const DevState<R>&, const StepArgs&, int, int, hipStream_t); \
    template const void* substep_kernel<R>(const DevMap&, int);                                                \
    template hipError_t launch_substep<R>(const DevMap&, const DevState<R>&, const StepArgs&, int, int,        \
                                          hipStream_t);                                                        \
    template void launch_swapfix<R>(const DevMap&, const DevState<R>&, hipStream_t);                           \
    template void launch_outpre<R>(const DevMap&, const DevState<R>&, int, int, double, double, hipStream_t);  \
    template void launch_scale_runoff<R>(const DevMap&, const DevState<R>&, const double*, const double*,     \
                                         hipStream_t);                                                          \
    template void launch_roff_interp<R>(const DevMap&, const DevState<R>&, const RoffMap&, const double*,      \
                                        hipStream_t);                                                           \
    template void launch_out_avemax<R>(const DevMap&, const DevState<R>&, const OutState&, double, int, int,   \
                                       hipStream_t);
CMF_INSTANTIATE(double)
CMF_INSTANTIATE(float)

}  // namespace cmf
