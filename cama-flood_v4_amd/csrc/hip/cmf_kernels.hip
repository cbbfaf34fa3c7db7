// cmf_kernels.hip -- HIP kernels of one CMF_PHYSICS_ADVANCE outer step on gfx950.
//
// One outer step = k_stage_cfl (+ k_reduce_cfl) followed by NT substeps of three
// phases; every phase is a grid-wide dependency of the reference's substep
// (src/cmf_ctrl_physics_mod.F90:64-144):
//   k_outflow   CMF_CALC_OUTFLW + CMF_DAMOUT_CALC (cells) and CMF_CALC_PTHOUT /
//               CMF_LEVEE_OPT_PTHOUT (paths) -- one launch, two index ranges
//   k_rate      CMF_CALC_INFLOW_LSPAMAT parts 1-2: outflow volume and limiter rate
//   k_update    CMF_CALC_INFLOW_LSPAMAT parts 2-3 (limit, upstream gather),
//               CMF_DAMOUT_WATBAL, CALC_VARS_PRE, CMF_CALC_STONXT,
//               CMF_PHYSICS_FLDSTG and CMF_DIAG_AVEMAX_ADPSTP (cells) and the path
//               limiter/averages (paths) -- one launch
//   k_reduce_budget  deterministic fixed-order reduction of the water-budget sums
// All neighbour accesses are gathers (downstream cell, upstream CSR list, path
// endpoints); nothing is scattered, so results do not depend on scheduling.
#include "cmf_device.hpp"
#include "cmf_kernels.hpp"

namespace cmf {

// ---------------------------------------------------------------------------
// Block reductions (fixed tree order: deterministic)
// Sum over the block: a fixed butterfly inside each 64-lane wavefront, then the
// BLOCK/64 wave sums in wave order. Result in lds[k * NWAVE] for thread-independent
// reads after the final barrier.
constexpr int NWAVE = BLOCK / 64;
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* [NV][NWAVE] */) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k * NWAVE + wave] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < NWAVE; ++w) s += lds[threadIdx.x * NWAVE + w];
        lds[threadIdx.x * NWAVE] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// phase 1a: river / mouth / floodplain outflow and dam operation, per cell
template <class R>
__device__ __forceinline__ void outflow_cell(const DevMap& M, const DevState<R>& S, const StepArgs& A, int i) {
    const R gdt = R(A.gdt);
    const R zero = R(0), one = R(1);
    const R rivelv = R(rivelv_of(M, i));
    const R rivdph = S.rivdph[i], rivdph_pre = S.rivdph_pre[i];
    const R sfcelv = rivelv + rivdph;                               // outflw:45-47
    const R sfcelv_pre = rivelv + rivdph_pre;
    const R flddph_pre = vmax(rivdph_pre - R(M.rivhgt[i]), zero);
    const R rivwth = R(M.rivwth[i]), rivman = R(M.rivman[i]);
    const R qr_pre = S.rivout[i];                                   // D2RIVOUT_PRE
    const R qr_old = S.rivout_vel[i];                               // D2RIVOUT (for RIVVEL)
    const int nxt = M.next[i];
    const bool river = nxt >= 0;                                     // mouths carry a negative code
    R dwn, dwn_pre, rdst;
    if (river) {                                                     // outflw:52-57
        const int j = nxt;
        const R rivelv_j = R(rivelv_of(M, j));
        dwn = rivelv_j + S.rivdph[j];
        dwn_pre = rivelv_j + S.rivdph_pre[j];
        rdst = one / R(M.nxtdst[i]);
    } else {
        dwn = R(M.dwnelv[i]);
        dwn_pre = dwn;
        rdst = one / R(A.pdstmth);
    }
    // ---- river channel: outflw:60-80 (river), :119-139 (mouth)
    R q_r, q_f = zero, vel;
    {
        R dslp, dflw, dflw_pr;
        if (river) {
            const R dsfc = vmax(sfcelv, dwn);
            dslp = (sfcelv - dwn) * rdst;
            dflw = dsfc - rivelv;
            const R dsfc_pr = vmax(sfcelv_pre, dwn_pre);
            dflw_pr = dsfc_pr - rivelv;
        } else {
            dslp = (sfcelv - dwn) * rdst;
            dflw = rivdph;
            dflw_pr = rivdph_pre;
        }
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
    // ---- floodplain: outflw:84-116 (river), :143-171 (mouth)
    if (A.lfldout) {
        const R gdtn2 = gdt * (R(A.pmanfld) * R(A.pmanfld));
        const R dfsto = R(S.fldsto[i]);
        const R elevtn = R(M.elevtn[i]);
        const R rrl = one / R(M.rivlen[i]);
        const R flddph = S.flddph[i];
        R dslp = (sfcelv - dwn) * rdst;
        dslp = vmax(R(-0.005), vmin(R(0.005), dslp));
        R dflw, dflw_pr, dare, dare_pr;
        if (river) {
            const R dsfc = vmax(sfcelv, dwn);
            dflw = vmax(dsfc - elevtn, zero);
            dare = dfsto * rrl;
            dare = vmax(dare - flddph * rivwth, zero);
            const R dsfc_pr = vmax(sfcelv_pre, dwn_pre);
            dflw_pr = dsfc_pr - elevtn;
            dare_pr = S.fldsto_pre[i] * rrl;
            dare_pr = vmax(dare_pr - flddph_pre * rivwth, R(1.e-6));
        } else {
            dflw = sfcelv - elevtn;
            dare = vmax(dfsto * rrl - flddph * rivwth, zero);
            dflw_pr = sfcelv_pre - elevtn;
            dare_pr = vmax(S.fldsto_pre[i] * rrl - flddph_pre * rivwth, R(1.e-6));
        }
        const R t = dflw * dflw_pr;
        R dflw_im = vsqrt(vmax(t, zero));
        dflw_im = vmax(dflw_im, R(1.e-6));
        R dare_im = vsqrt(dare * dare_pr);
        dare_im = vmax(dare_im, R(1.e-6));
        const R dout_pr = S.fldout[i];                              // D2FLDOUT_PRE
        const R dout = (dout_pr + (gdt * dare_im) * dslp) *
                       (one / (one + ((gdtn2 * vabs(dout_pr)) * vpow(dflw_im, R(K_P4O3))) * (one / dare_im)));
        const R thr = river ? R(1.e-5) : R(K_1EM5F);
        const bool mask = (dflw_im > thr) && (dare > thr);
        const R f = mask ? dout : zero;
        q_f = (f * q_r > zero) ? f : zero;
    }
    // ---- reverse-flow storage limiter, river cells: outflw:175-182
    if (river) {
        R dout = ((-q_r) - q_f) * R(A.dt);
        dout = vmax(dout, R(K_1EM10F));
        R rate = (R(0.05) * S.storge[i]) / dout;
        rate = vmin(rate, one);
        q_r = q_r * rate;
        q_f = q_f * rate;
    }
    // ---- reservoir operation: damout:316-498 (Yamazaki & Funato rule)
    if (A.ldamout) {
        const int code = M.i1dam[i];
        if (code > 0) {
            q_r = zero; q_f = zero;
            double inf = (double)S.runoff[i];                       // :447
            for (int k = M.up_ptr[i]; k < M.up_ptr[i + 1]; ++k) {   // :456-464, ascending
                const int u = M.up_idx[k];
                const int cu = M.i1dam[u];
                if (cu == 10 || cu == 11) inf = (inf + (double)S.rivout[u]) + (double)S.fldout[u];
            }
            S.daminf[i] = inf;
            const double dt = A.dt;
            if (code == 10 && river) {                               // :471-493
                const int j = nxt;
                double dslope = ((double)M.elevtn[i] - (double)M.elevtn[j]) * (1.0 / (double)M.nxtdst[i]);
                dslope = vmax(dslope, A.pminslp);
                const double rd = (double)rivdph;
                const double dvel = ((1.0 / (double)M.rivman[i]) * sqrt(dslope)) * pow(rd, K_2O3F);
                const double darea = (double)M.rivwth[i] * rd;
                vel = R(dvel);
                double q = darea * dvel;
                q_r = R(vmin(q, S.rivsto[i] / dt));
                const double dslope_f = vmin(0.005, dslope);
                const double fd = (double)S.flddph[i];
                const double dvel_f = ((1.0 / A.pmanfld) * sqrt(dslope_f)) * pow(fd, K_2O3F);
                double dare_f = S.fldsto[i] * (1.0 / (double)M.rivlen[i]);
                dare_f = vmax(dare_f - fd * (double)M.rivwth[i], 0.0);
                const double qf = dare_f * dvel_f;
                q_f = R(vmin(qf, S.fldsto[i] / dt));
            }
            const int d = M.dam_of_cell[i];
            if (d >= 0 && M.damstat[d] > 0) {                        // :344-419
                const double vol = S.damsto[i];
                const double cv = M.dam_convol[d], av = M.dam_adjvol[d], ev = M.dam_emevol[d];
                const double qn = M.dam_qn[d], qf = M.dam_qf[d], qa = M.dam_qa[d];
                double out = 0.0;
                if (vol <= cv) {
                    out = qn * sqrt(vol / cv);
                } else if (vol > cv && vol <= av) {
                    const double x = (vol - cv) / (av - cv);
                    out = qn + (x * (x * x)) * (qa - qn);
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
                out = vmin(vmin(out, vol / dt), (S.rivsto[i] + S.fldsto[i]) / dt);
                out = vmax(out, 0.0);
                q_r = R(out);
                q_f = zero;
            }
        }
    }
    S.qr_u[i] = q_r;
    S.qf_u[i] = q_f;
    if (A.last) S.rivvel[i] = vel;                                  // output-only diagnostic
    // CMF_DIAG_AVEMAX_ADPSTP (diag:113) for RIVVEL, final after OUTFLW/DAMOUT
    R va = (A.first ? zero : S.rivvel_avg[i]) + vel * R(A.dt);
    if (A.last) va = va / R(A.nadd);
    S.rivvel_avg[i] = va;
}

// phase 1b: bifurcation path flow, pthout:20-97 / levee:497-603
template <class R>
__device__ __forceinline__ void pthflow_path(const DevMap& M, const DevState<R>& S, const StepArgs& A, int p) {
    const int np = M.npth, nl = M.npthlev;
    const int ip = M.pth_upst[p], jp = M.pth_down[p];
    bool skip = ip < 0 || jp < 0;
    if (!skip) skip = A.llevee ? (M.i2mask[ip] == 1 || M.i2mask[jp] == 1) : (M.i2mask[ip] > 0 || M.i2mask[jp] > 0);
    if (skip) {
        for (int l = 0; l < nl; ++l) S.pthflw_u[(size_t)l * np + p] = R(0);
        S.pthflwsum_u[p] = R(0);
        return;
    }
    const R zero = R(0), one = R(1);
    const R gdt = R(A.gdt);
    const R ri = R(rivelv_of(M, ip)), rj = R(rivelv_of(M, jp));
    const R si = ri + S.rivdph[ip], sj = rj + S.rivdph[jp];
    const R pi = ri + S.rivdph_pre[ip], pj = rj + S.rivdph_pre[jp];
    R dslp = (si - sj) / R(M.pth_dst[p]);
    dslp = vmax(R(-0.005), vmin(R(0.005), dslp));
    R li = si, lj = sj;                                              // D2SFCELV_LEV
    if (A.llevee) {
        const int a = M.lev_idx[ip], b = M.lev_idx[jp];
        const int nlc = M.nlevcell;
        if (a >= 0 && M.lev_tab[nlc + a] < 1.0) li = vmin(R(M.elevtn[ip]) + S.levdph[ip], si);
        if (b >= 0 && M.lev_tab[nlc + b] < 1.0) lj = vmin(R(M.elevtn[jp]) + S.levdph[jp], sj);
    }
    R sum = zero;
    for (int l = 0; l < nl; ++l) {
        const R elv = R(M.pth_elv[(size_t)l * np + p]);
        R dflw, dflw_im;
        if (!A.llevee || l == 0) {
            dflw = vmax(vmax(si, sj) - elv, zero);
            R dflw_pr = vmax(vmax(pi, pj) - elv, zero);
            dflw_im = vsqrt(dflw * dflw_pr);
            if (A.llevee) { if (dflw_im <= zero) dflw_im = dflw; }
            else dflw_im = vmax(dflw_im, vsqrt(dflw * R(0.01)));
        } else {
            dflw = vmax(vmax(li, lj) - elv, zero);
            dflw_im = dflw;
        }
        R q = zero;
        if (dflw_im > R(1.e-5)) {
            const R w = R(M.pth_wth[(size_t)l * np + p]);
            const R man = R(M.pth_man[l]);
            const R pre = S.pthflw[(size_t)l * np + p];             // D1PTHFLW_PRE
            const R dout_pr = (A.llevee && l > 0) ? pre * (one / w) : pre / w;
            q = (w * (dout_pr + (gdt * dflw_im) * dslp)) /
                (one + ((gdt * (man * man)) * vabs(dout_pr)) * vpow(dflw_im, R(K_P7O3)));
        }
        S.pthflw_u[(size_t)l * np + p] = q;
        sum = sum + q;
    }
    if (sum != zero) {                                               // pthout:84-95
        R rate = (R(0.05) * vmin(S.storge[ip], S.storge[jp])) / vabs(sum * R(A.dt));
        rate = vmin(rate, one);
        for (int l = 0; l < nl; ++l) S.pthflw_u[(size_t)l * np + p] = S.pthflw_u[(size_t)l * np + p] * rate;
        sum = sum * rate;
    }
    S.pthflwsum_u[p] = sum;
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_outflow(DevMap M, DevState<R> S, StepArgs A, int nb_cells) {
    if ((int)blockIdx.x < nb_cells) {
        const int i = blockIdx.x * BLOCK + threadIdx.x;
        if (i < M.n) outflow_cell<R>(M, S, A, i);
    } else {
        const int p = (blockIdx.x - nb_cells) * BLOCK + threadIdx.x;
        if (p < M.npth) pthflow_path<R>(M, S, A, p);
    }
}

// ---------------------------------------------------------------------------
// phase 2: outflow volume and limiter rate, inflow_lspamat:410-451
template <class R>
__global__ __launch_bounds__(BLOCK) void k_rate(DevMap M, DevState<R> S, StepArgs A) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= M.n) return;
    const double dt = A.dt;
    double so = (double)vmax(S.qr_u[i], R(0)) + (double)vmax(S.qf_u[i], R(0));
    for (int k = M.up_ptr[i]; k < M.up_ptr[i + 1]; ++k) {
        const int u = M.up_idx[k];
        so = (so + (double)vmax(-S.qr_u[u], R(0))) + (double)vmax(-S.qf_u[u], R(0));
    }
    so = so * dt;
    if (A.lpthout) {
        for (int k = M.pout_ptr[i]; k < M.pout_ptr[i + 1]; ++k)
            so = so + (double)vmax(S.pthflwsum_u[M.pout_idx[k]], R(0)) * dt;
        for (int k = M.pinf_ptr[i]; k < M.pinf_ptr[i + 1]; ++k)
            so = so + (double)vmax(-S.pthflwsum_u[M.pinf_idx[k]], R(0)) * dt;
    }
    const R dout = vmax(R(so), R(1.e-10));
    const R dsto = R(S.rivsto[i] + S.fldsto[i]);
    S.rate[i] = vmin(dsto / dout, R(1));
}

// ---------------------------------------------------------------------------
// phase 3: limit, gather inflow, dam balance, storage update, stage, averages
template <class R>
__device__ __forceinline__ R limited_path_sum(const DevMap& M, const DevState<R>& S, int p) {
    const R s = S.pthflwsum_u[p];
    const int ip = M.pth_upst[p], jp = M.pth_down[p];
    if (M.i2mask[ip] > 0 || M.i2mask[jp] > 0) return s;               // inflow_lspamat:501
    return (s >= R(0)) ? s * S.rate[ip] : s * S.rate[jp];
}

template <class R>
__device__ __forceinline__ void update_cell(const DevMap& M, const DevState<R>& S, const StepArgs& A,
                                            int i, double (&g)[NGLB]) {
    const double dt = A.dt;
    const R zero = R(0);
    // limiter, inflow_lspamat:456-473
    const R qr = S.qr_u[i], qf = S.qf_u[i];
    const R ri = S.rate[i];
    const int nxt = M.next[i];
    const R r = (nxt >= 0 && !(qr > zero)) ? S.rate[nxt] : ri;
    const R qrl = qr * r, qfl = qf * r;
    // upstream gather, inflow_lspamat:481-489
    double rivinf = 0.0, fldinf = 0.0;
    for (int k = M.up_ptr[i]; k < M.up_ptr[i + 1]; ++k) {
        const int u = M.up_idx[k];
        const R qu = S.qr_u[u], fu = S.qf_u[u];
        const R ru = (qu > zero) ? S.rate[u] : ri;
        rivinf = rivinf + (double)(qu * ru);
        fldinf = fldinf + (double)(fu * ru);
    }
    double pthout = 0.0;
    if (A.lpthout) {                                                  // :519-535
        for (int k = M.pout_ptr[i]; k < M.pout_ptr[i + 1]; ++k)
            pthout = pthout + (double)limited_path_sum<R>(M, S, M.pout_idx[k]);
        for (int k = M.pinf_ptr[i]; k < M.pinf_ptr[i + 1]; ++k)
            pthout = pthout - (double)limited_path_sum<R>(M, S, M.pinf_idx[k]);
    }
    const R rivinf_r = R(rivinf), fldinf_r = R(fldinf), pthout_r = R(pthout);
    if (A.last) { S.rivinf[i] = rivinf; S.fldinf[i] = fldinf; S.pthout[i] = pthout; }
    const R runoff = S.runoff[i];
    // CMF_DAMOUT_WATBAL, damout:521-536
    if (A.ldamout) {
        const int d = M.dam_of_cell[i];
        if (d >= 0 && M.damstat[d] != -9999) {
            const double inflow = ((double)rivinf_r + (double)fldinf_r) + (double)runoff;
            const double outflw = (double)qrl + (double)qfl;
            S.damsto[i] = (S.damsto[i] + inflow * dt) - outflw * dt;
        }
    }
    // CALC_VARS_PRE, physics:254-276. RIVOUT_PRE/FLDOUT_PRE share the RIVOUT/FLDOUT
    // arrays; RIVDPH/RIVDPH_PRE are a buffer pair whose roles the host swaps after
    // this kernel, so the new depth goes to the buffer of the old RIVDPH_PRE.
    double rs = S.rivsto[i], fs = S.fldsto[i];
    S.rivout[i] = qrl;
    S.fldout[i] = qfl;
    S.fldsto_pre[i] = R(fs);
    // CMF_CALC_STONXT, stonxt:54-98
    g[G_STOPRE] = rs + fs;
    g[G_RIVINF] = (double)rivinf_r * dt + (double)fldinf_r * dt;
    g[G_RIVOUT] = ((double)qrl * dt + (double)qfl * dt) + (double)pthout_r * dt;
    rs = (rs + (double)rivinf_r * dt) - (double)qrl * dt;
    if (rs < 0.0) { fs = fs + rs; rs = 0.0; }
    fs = ((fs + (double)fldinf_r * dt) - (double)qfl * dt) - (double)pthout_r * dt;
    if (fs < 0.0) { rs = vmax(rs + fs, 0.0); fs = 0.0; }
    g[G_STONXT] = rs + fs;
    const R outflw = qrl + qfl;
    const R fldfrc_old = S.fldfrc[i];
    const R drivrof = (runoff * (R(1) - fldfrc_old)) * R(dt);
    const R dfldrof = (runoff * fldfrc_old) * R(dt);
    rs = rs + (double)drivrof;
    fs = fs + (double)dfldrof;
    double levsto = A.llevee ? S.levsto[i] : 0.0;
    double st = rs + fs;
    if (A.llevee) st = st + levsto;
    g[G_STONEW] = (double)R(st);
    // CMF_PHYSICS_FLDSTG
    Stage o;
    if (A.llevee) stage_levee(M, i, rs, fs, levsto, o);
    else stage_def(M, i, rs, fs, o);
    S.rivsto[i] = o.rivsto;
    S.fldsto[i] = o.fldsto;
    const R rivdph = R(o.rivdph);
    S.rivdph_next[i] = rivdph;
    S.flddph[i] = R(o.flddph);
    S.fldfrc[i] = R(o.fldfrc);
    if (A.last) {                                                   // output-only diagnostics
        S.fldare[i] = R(o.fldare);
        S.sfcelv[i] = R(rivelv_of(M, i)) + rivdph;
        S.outflw[i] = outflw;
    }
    double stg = o.rivsto + o.fldsto;
    if (A.llevee) {
        S.levsto[i] = o.levsto;
        S.levdph[i] = R(o.levdph);
        stg = stg + o.levsto;
    }
    const R storge = R(stg);
    S.storge[i] = storge;
    g[G_STOPRE2] = A.llevee ? (rs + fs) + levsto : rs + fs;
    g[G_STONEW2] = stg;
    g[G_RIVSTO] = o.rivsto;
    g[G_FLDSTO] = o.fldsto;
    g[G_FLDARE] = o.fldare;
    g[G_LEVSTO] = A.llevee ? o.levsto : 0.0;
    // CMF_DIAG_AVEMAX_ADPSTP, diag:108-124, and GETAVE :179-186 on the last substep
    const R rdt = R(dt);
    const bool f0 = A.first;
    R a1 = (f0 ? zero : S.rivout_avg[i]) + qrl * rdt;
    R a2 = (f0 ? zero : S.fldout_avg[i]) + qfl * rdt;
    R a3 = (f0 ? zero : S.outflw_avg[i]) + outflw * rdt;
    if (A.last) {
        const R nadd = R(A.nadd);
        a1 = a1 / nadd; a2 = a2 / nadd; a3 = a3 / nadd;
        // RUNOFF is constant over the outer step: its NT additions are replayed in
        // registers, in the same order, instead of a read-modify-write per substep
        R a5 = zero;
        for (int k = 0; k < A.nrun; ++k) a5 = a5 + runoff * rdt;
        S.runoff_avg[i] = a5 / nadd;
    }
    S.rivout_avg[i] = a1; S.fldout_avg[i] = a2; S.outflw_avg[i] = a3;
    if (A.lpthout) {
        R a4 = (f0 ? zero : S.pthout_avg[i]) + pthout_r * rdt;
        if (A.last) a4 = a4 / R(A.nadd);
        S.pthout_avg[i] = a4;
    } else if (f0) {
        S.pthout_avg[i] = zero;                                     // PTHOUT = 0 without paths
    }
    S.outflw_max[i] = vmax(f0 ? zero : S.outflw_max[i], vabs(outflw));
    S.rivdph_max[i] = vmax(f0 ? zero : S.rivdph_max[i], rivdph);
    S.storge_max[i] = vmax(f0 ? zero : S.storge_max[i], storge);
    if (A.ldamout) {
        R da = (f0 ? zero : S.daminf_avg[i]) + R(S.daminf[i] * dt);
        if (A.last) da = da / R(A.nadd);
        S.daminf_avg[i] = da;
    }
}

template <class R>
__device__ __forceinline__ void update_path(const DevMap& M, const DevState<R>& S, const StepArgs& A, int p) {
    const int np = M.npth, nl = M.npthlev;
    const int ip = M.pth_upst[p], jp = M.pth_down[p];
    const bool lim = !(ip < 0 || jp < 0) && !(M.i2mask[ip] > 0 || M.i2mask[jp] > 0);
    const R zero = R(0);
    const R rdt = R(A.dt);
    R sumavg = zero;
    for (int l = 0; l < nl; ++l) {
        R q = S.pthflw_u[(size_t)l * np + p];
        if (lim) q = (q >= zero) ? q * S.rate[ip] : q * S.rate[jp];   // inflow_lspamat:503-509
        S.pthflw[(size_t)l * np + p] = q;                            // D1PTHFLW == D1PTHFLW_PRE
        R a = (A.first ? zero : S.pthflw_avg[(size_t)l * np + p]) + q * rdt;
        if (A.last) {
            a = a / R(A.nadd);
            sumavg = sumavg + a;                                     // diag:206-213
        }
        S.pthflw_avg[(size_t)l * np + p] = a;
    }
    R s = S.pthflwsum_u[p];
    if (lim) s = (s >= zero) ? s * S.rate[ip] : s * S.rate[jp];       // :511-515
    S.pthflwsum[p] = s;
    if (A.last) S.pthflwsum_avg[p] = sumavg;
}

template <class R>
__global__ __launch_bounds__(BLOCK) void k_update(DevMap M, DevState<R> S, StepArgs A, int nb_cells) {
    __shared__ double lds[NGLB * NWAVE];
    if ((int)blockIdx.x < nb_cells) {
        const int i = blockIdx.x * BLOCK + threadIdx.x;
        double g[NGLB];
#pragma unroll
        for (int k = 0; k < NGLB; ++k) g[k] = 0.0;
        if (i < M.n) update_cell<R>(M, S, A, i, g);
        block_sum<NGLB>(g, lds);
        // per-substep slot: reduced once for all substeps at the end of the outer step
        if (threadIdx.x < NGLB)
            S.partials[((size_t)A.it * nb_cells + blockIdx.x) * NGLB + threadIdx.x] = lds[threadIdx.x * NWAVE];
    } else {
        const int p = (blockIdx.x - nb_cells) * BLOCK + threadIdx.x;
        if (p < M.npth) update_path<R>(M, S, A, p);
    }
}

// deterministic second-level reduction of the per-block budget sums: block b
// reduces substep b of the outer step
__global__ __launch_bounds__(BLOCK) void k_reduce_budget(const double* partials, int nb, double* out) {
    __shared__ double lds[NGLB * NWAVE];
    const double* src = partials + (size_t)blockIdx.x * nb * NGLB;
    double v[NGLB];
#pragma unroll
    for (int k = 0; k < NGLB; ++k) v[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += BLOCK)
#pragma unroll
        for (int k = 0; k < NGLB; ++k) v[k] += src[(size_t)b * NGLB + k];
    block_sum<NGLB>(v, lds);
    if (threadIdx.x < NGLB) out[(size_t)blockIdx.x * NGLB + threadIdx.x] = lds[threadIdx.x * NWAVE];
}

// ---------------------------------------------------------------------------
// outer-step start: CMF_PHYSICS_FLDSTG and the CFL bound of CALC_ADPSTP
template <class R>
__global__ __launch_bounds__(BLOCK) void k_stage(DevMap M, DevState<R> S, StepArgs A, int with_cfl) {
    __shared__ double lds[BLOCK];
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    double c = 1.0e300;
    if (i < M.n) {
        Stage o;
        const double levsto = A.llevee ? S.levsto[i] : 0.0;
        if (A.llevee) stage_levee(M, i, S.rivsto[i], S.fldsto[i], levsto, o);
        else stage_def(M, i, S.rivsto[i], S.fldsto[i], o);
        S.rivsto[i] = o.rivsto;
        S.fldsto[i] = o.fldsto;
        const R rivdph = R(o.rivdph);
        S.rivdph[i] = rivdph;
        S.flddph[i] = R(o.flddph);
        S.fldfrc[i] = R(o.fldfrc);
        S.fldare[i] = R(o.fldare);
        S.sfcelv[i] = R(rivelv_of(M, i)) + rivdph;
        double stg = o.rivsto + o.fldsto;
        if (A.llevee) {
            S.levsto[i] = o.levsto;
            S.levdph[i] = R(o.levdph);
            stg = stg + o.levsto;
        }
        S.storge[i] = R(stg);
        if (with_cfl && M.i2mask[i] == 0) {                          // physics:178-197
            const double ddph = vmax((double)rivdph, 0.01);
            const double ddst = (M.next[i] >= 0) ? (double)M.nxtdst[i] : A.pdstmth;
            c = (A.pcadp * ddst) * (1.0 / sqrt(A.pgrv * ddph));
        }
    }
    if (!with_cfl) return;
    lds[threadIdx.x] = c;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] = vmin(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) S.cfl_part[blockIdx.x] = lds[0];
}

__global__ __launch_bounds__(BLOCK) void k_reduce_cfl(const double* part, int nb, double dt_def, double* ctl) {
    __shared__ double lds[BLOCK];
    double m = dt_def;
    for (int b = threadIdx.x; b < nb; b += BLOCK) m = vmin(m, part[b]);
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] = vmin(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) ctl[0] = lds[0];
}

// ---------------------------------------------------------------------------
// launch wrappers
template <class R>
void launch_stage(const DevMap& M, const DevState<R>& S, const StepArgs& A, int with_cfl, double dt_def,
                  hipStream_t st) {
    const int nb = (M.n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_stage<R>, dim3(nb), dim3(BLOCK), 0, st, M, S, A, with_cfl);
    if (with_cfl) hipLaunchKernelGGL(k_reduce_cfl, dim3(1), dim3(BLOCK), 0, st, S.cfl_part, nb, dt_def, S.ctl);
}

template <class R>
void launch_outflow(const DevMap& M, const DevState<R>& S, const StepArgs& A, hipStream_t st) {
    const int nbc = (M.n + BLOCK - 1) / BLOCK;
    const int nbp = A.lpthout ? (M.npth + BLOCK - 1) / BLOCK : 0;
    hipLaunchKernelGGL(k_outflow<R>, dim3(nbc + nbp), dim3(BLOCK), 0, st, M, S, A, nbc);
}

template <class R>
void launch_rate(const DevMap& M, const DevState<R>& S, const StepArgs& A, hipStream_t st) {
    const int nbc = (M.n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_rate<R>, dim3(nbc), dim3(BLOCK), 0, st, M, S, A);
}

template <class R>
void launch_update(const DevMap& M, const DevState<R>& S, const StepArgs& A, hipStream_t st) {
    const int nbc = (M.n + BLOCK - 1) / BLOCK;
    const int nbp = A.lpthout ? (M.npth + BLOCK - 1) / BLOCK : 0;
    hipLaunchKernelGGL(k_update<R>, dim3(nbc + nbp), dim3(BLOCK), 0, st, M, S, A, nbc);
}

void launch_budget(const DevMap& M, const double* partials, int nrun, double* glb_hist, hipStream_t st) {
    const int nbc = (M.n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_reduce_budget, dim3(nrun), dim3(BLOCK), 0, st, partials, nbc, glb_hist);
}

#define CMF_INSTANTIATE(R)                                                                              \
    template void launch_stage<R>(const DevMap&, const DevState<R>&, const StepArgs&, int, double,      \
                                  hipStream_t);                                                         \
    template void launch_outflow<R>(const DevMap&, const DevState<R>&, const StepArgs&, hipStream_t);   \
    template void launch_rate<R>(const DevMap&, const DevState<R>&, const StepArgs&, hipStream_t);      \
    template void launch_update<R>(const DevMap&, const DevState<R>&, const StepArgs&, hipStream_t);
CMF_INSTANTIATE(double)
CMF_INSTANTIATE(float)

}  // namespace cmf
