// cmf_kernels.hpp -- launch wrappers of the substep kernels (cmf_kernels.hip).
#pragma once
#include "cmf_device.hpp"

namespace cmf {

// block sizes the phase kernels are built for (the autotuner picks among them)
constexpr int TUNE_BS[] = {64, 128, 256, 512};

// tiles of one substep phase: cell tiles, then path tiles (paths only with LPTHOUT)
int tiles_cells(const DevMap& M, int bs);
int tiles_paths(const DevMap& M, const StepArgs& A, int bs);

// k_stage (+ CFL bound per instance into S.cfl when with_cfl)
template <class R>
void launch_stage(const DevMap& M, const DevState<R>& S, const StepArgs& A, int with_cfl, int ninst, double dt_def,
                  hipStream_t st);
// one substep phase as its own kernel: 0 = outflow, 1 = rate, 2 = update
template <class R>
void launch_phase(const DevMap& M, const DevState<R>& S, const StepArgs& A, int phase, int bs, hipStream_t st);
// the whole substep as one persistent cooperative kernel of `grid` blocks
template <class R>
hipError_t launch_substep(const DevMap& M, const DevState<R>& S, const StepArgs& A, int bs, int grid,
                          hipStream_t st);
template <class R>
const void* substep_kernel(const DevMap& M, int bs);
// per-slot cell arrays between the host layout (member planes, NSEQ order, type H)
// in a device staging buffer and the device slots (type R); perm NULL = identity
template <class R, class H>
void launch_cells_from_host(R* dst, const H* src, const int32_t* perm, int n, int lgm, int ns, int shared,
                            hipStream_t st);
template <class R, class H>
void launch_cells_to_host(H* dst, const R* src, const int32_t* perm, int n, int lgm, int ns, hipStream_t st);
// CMF_CALC_OUTPRE: RIVOUT_PRE, FLDOUT_PRE, RIVDPH_PRE, SFCELV (and D1PTHFLW_PRE with paths)
template <class R>
void launch_outpre(const DevMap& M, const DevState<R>& S, int lfldout, int lpthout, double pmanfld, double pdstmth,
                   hipStream_t st);
template <class R>
void launch_swapfix(const DevMap& M, const DevState<R>& S, hipStream_t st);
template <class R>
void launch_scale_runoff(const DevMap& M, const DevState<R>& S, const double* base, const double* factor,
                         hipStream_t st);
template <class R>
void launch_roff_interp(const DevMap& M, const DevState<R>& S, const RoffMap& Q, const double* grid,
                        hipStream_t st);
template <class R>
void launch_out_avemax(const DevMap& M, const DevState<R>& S, const OutState& O, double dt, int ldamout,
                       int npth_slots, hipStream_t st);
void launch_out_getave(const OutState& O, int ns, int npth_slots, double nadd, int ldamout, hipStream_t st);
void launch_out_reset(const OutState& O, int ns, int npth_slots, hipStream_t st);
void launch_boundary(int n, const float* elevtn, const double* meansl, const double* sealev, double* dwnelv,
                     hipStream_t st);
// total storage of every instance into out[ninst] (part: ninst * nch doubles;
// dom_off: first cell of each domain in device order, and the end)
void launch_daminf_nonzero(const int32_t* dam_of_cell, const double* daminf, int ns, int lgm, int32_t* flag,
                           hipStream_t st);
void launch_inst_storage(const double* rivsto, const double* fldsto, const double* levsto, const int32_t* dom_off,
                         int lgm, int ninst, int nch, double* part, double* out, hipStream_t st);
// reduce the per-tile budget sums of nrun substeps into glb_hist[nrun][NGLB]
void launch_budget(const double* partials, int ntiles, int nrun, double* glb_hist, hipStream_t st);

}  // namespace cmf
