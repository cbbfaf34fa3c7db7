// cmf_kernels.hpp -- launch wrappers of the substep kernels (cmf_kernels.hip).
#pragma once
#include "cmf_device.hpp"

namespace cmf {

template <class R>
void launch_stage(const DevMap& M, const DevState<R>& S, const StepArgs& A, int with_cfl, double dt_def,
                  hipStream_t st);
template <class R>
void launch_outflow(const DevMap& M, const DevState<R>& S, const StepArgs& A, hipStream_t st);
template <class R>
void launch_rate(const DevMap& M, const DevState<R>& S, const StepArgs& A, hipStream_t st);
template <class R>
void launch_update(const DevMap& M, const DevState<R>& S, const StepArgs& A, hipStream_t st);
// reduce the per-block budget sums of nrun substeps into glb_hist[nrun][NGLB]
void launch_budget(const DevMap& M, const double* partials, int nrun, double* glb_hist, hipStream_t st);

}  // namespace cmf
