// cmf_capi.hip -- C-ABI of include/cmf_hip.h: device-resident model state, map
// upload and validation, and the outer-step driver that replaces
// CMF_PHYSICS_ADVANCE / CMF_PHYSICS_FLDSTG (src/cmf_ctrl_physics_mod.F90).
#include "cmf_hip.h"
#include "cmf_kernels.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

using namespace cmf;

namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
}  // namespace

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(CMF_EHIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

struct cmf_hip_ctx {
    cmf_dims d{};
    cmf_options o{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool map_set = false;
    std::vector<void*> allocs;
    DevMap M{};
    DevState<double> S64{};
    DevState<float> S32{};
    void* vel_buf = nullptr;          // D2RIVOUT when it differs from D2RIVOUT_PRE
    bool vel_pending = false;
    double* h_ctl = nullptr;          // pinned
    int nb_cells = 0;
    int hist_cap = 0;
    // device cell order: perm[p] = NSEQ index of device position p (identity when
    // CMF_HIP_ORDER=nseq); the C-ABI always speaks NSEQ order
    std::vector<int32_t> perm, iperm;
    bool permuted = false;
    // per-kernel timing (HIP events on the handle's stream)
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double t_ms[5] = {0, 0, 0, 0, 0};
};

namespace {

template <class T>
int dalloc(cmf_hip_ctx* h, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(CMF_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    // clear on the handle's own (non-blocking) stream and wait, so the clear is
    // ordered before every later copy and kernel of this handle
    e = hipMemsetAsync(q, 0, n * sizeof(T), h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(CMF_EHIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

#define TRY(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_ != 0) return r_; \
    } while (0)

template <class R>
int alloc_state(cmf_hip_ctx* h, DevState<R>& S) {
    const size_t n = (size_t)h->d.nseqall;
    const size_t np = (size_t)(h->o.lpthout ? h->d.npthout : 0);
    const size_t npl = np * (size_t)(h->o.lpthout ? h->d.npthlev : 0);
    TRY(dalloc(h, &S.rivsto, n)); TRY(dalloc(h, &S.fldsto, n));
    TRY(dalloc(h, &S.levsto, h->o.llevee ? n : 1));
    TRY(dalloc(h, &S.damsto, h->o.ldamout ? n : 1));
    TRY(dalloc(h, &S.daminf, h->o.ldamout ? n : 1));
    TRY(dalloc(h, &S.rivout, n)); TRY(dalloc(h, &S.fldout, n));
    TRY(dalloc(h, &S.qr_u, n)); TRY(dalloc(h, &S.qf_u, n));
    TRY(dalloc(h, &S.rivdph_pre, n)); TRY(dalloc(h, &S.fldsto_pre, n));
    TRY(dalloc(h, &S.rivdph, n)); TRY(dalloc(h, &S.flddph, n)); TRY(dalloc(h, &S.fldfrc, n));
    TRY(dalloc(h, &S.fldare, n)); TRY(dalloc(h, &S.sfcelv, n)); TRY(dalloc(h, &S.storge, n));
    TRY(dalloc(h, &S.outflw, n)); TRY(dalloc(h, &S.rivvel, n));
    TRY(dalloc(h, &S.levdph, h->o.llevee ? n : 1));
    TRY(dalloc(h, &S.rivinf, n)); TRY(dalloc(h, &S.fldinf, n)); TRY(dalloc(h, &S.pthout, n));
    TRY(dalloc(h, &S.rate, n)); TRY(dalloc(h, &S.runoff, n));
    TRY(dalloc(h, &S.pthflw, npl)); TRY(dalloc(h, &S.pthflw_u, npl));
    TRY(dalloc(h, &S.pthflwsum, np)); TRY(dalloc(h, &S.pthflwsum_u, np));
    TRY(dalloc(h, &S.rivout_avg, n)); TRY(dalloc(h, &S.fldout_avg, n)); TRY(dalloc(h, &S.outflw_avg, n));
    TRY(dalloc(h, &S.rivvel_avg, n)); TRY(dalloc(h, &S.pthout_avg, n)); TRY(dalloc(h, &S.runoff_avg, n));
    TRY(dalloc(h, &S.outflw_max, n)); TRY(dalloc(h, &S.rivdph_max, n)); TRY(dalloc(h, &S.storge_max, n));
    TRY(dalloc(h, &S.daminf_avg, h->o.ldamout ? n : 1));
    TRY(dalloc(h, &S.pthflw_avg, npl)); TRY(dalloc(h, &S.pthflwsum_avg, np));
    h->hist_cap = 16;                                   // grown on demand (substeps per outer step)
    TRY(dalloc(h, &S.partials, (size_t)h->hist_cap * h->nb_cells * NGLB));
    TRY(dalloc(h, &S.cfl_part, (size_t)h->nb_cells));
    TRY(dalloc(h, &S.ctl, 4));
    TRY(dalloc(h, &S.glb_hist, (size_t)h->hist_cap * NGLB));
    S.rivout_vel = S.rivout;
    R* vb = nullptr;
    TRY(dalloc(h, &vb, n));
    h->vel_buf = vb;
    return 0;
}

bool is_f32(double x) { return (double)(float)x == x; }

// host copies with element-type conversion
// Host <-> device copies with element-type conversion. cell=true arrays are
// per-cell (NSEQ order on the host, device order on the device).
template <class R>
int put(cmf_hip_ctx* h, R* dst, const double* src, size_t n, bool cell = true) {
    if (!src || n == 0) return 0;
    const bool perm = cell && h->permuted;
    if (sizeof(R) == sizeof(double) && !perm) {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    std::vector<R> tmp(n);
    if (perm) for (size_t k = 0; k < n; ++k) tmp[k] = (R)src[h->perm[k]];
    else for (size_t k = 0; k < n; ++k) tmp[k] = (R)src[k];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * sizeof(R), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

template <class R>
int get(cmf_hip_ctx* h, double* dst, const R* src, size_t n, bool cell = true) {
    if (!dst || n == 0) return 0;
    const bool perm = cell && h->permuted;
    if (sizeof(R) == sizeof(double) && !perm) {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    std::vector<R> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (perm) for (size_t k = 0; k < n; ++k) dst[h->perm[k]] = (double)tmp[k];
    else for (size_t k = 0; k < n; ++k) dst[k] = (double)tmp[k];
    return 0;
}

// per-cell host array in device order (planes of n cells)
template <class T>
std::vector<T> to_dev(const cmf_hip_ctx* h, const T* src, size_t n, size_t planes = 1) {
    std::vector<T> out(n * planes);
    for (size_t l = 0; l < planes; ++l)
        for (size_t p = 0; p < n; ++p) out[l * n + p] = src[l * n + (h->permuted ? (size_t)h->perm[p] : p)];
    return out;
}

template <class T>
int put_raw(cmf_hip_ctx* h, T** dst, const T* src, size_t n) {
    TRY(dalloc(h, dst, n));
    if (n) {
        HIPCHK(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

// float32 copy of a map array; every reference map is read from a float32 file
int put_f32(cmf_hip_ctx* h, const float** dst, const double* src, size_t n, const char* name,
            size_t planes = 1) {
    if (!src) return fail(CMF_EINVAL, "map array %s is required", name);
    for (size_t k = 0; k < n * planes; ++k)
        if (!is_f32(src[k]))
            return fail(CMF_EMAP, "map array %s[%zu]=%.17g is not a float32 value (maps are float32 files)",
                        name, k, src[k]);
    std::vector<double> d = to_dev(h, src, n, planes);
    std::vector<float> tmp(n * planes);
    for (size_t k = 0; k < n * planes; ++k) tmp[k] = (float)d[k];
    float* p = nullptr;
    TRY(put_raw(h, &p, tmp.data(), n * planes));
    *dst = p;
    return 0;
}

// Device cell order: depth-first post-order of the river tree from each mouth,
// upstream branches in ascending NSEQ order. A cell's upstream subtree precedes
// it and a chain of cells is contiguous, so the downstream and upstream gathers of
// a wavefront mostly touch the same or adjacent cache lines.
bool dfs_order(int n, int nr, const std::vector<int32_t>& up_ptr, const std::vector<int32_t>& up_idx,
               std::vector<int32_t>& perm) {
    perm.assign((size_t)n, -1);
    std::vector<std::pair<int32_t, int32_t>> stack;
    int pos = 0;
    for (int root = nr; root < n; ++root) {
        stack.push_back({root, up_ptr[root]});
        while (!stack.empty()) {
            auto& top = stack.back();
            if (top.second < up_ptr[(size_t)top.first + 1]) {
                const int32_t c = up_idx[(size_t)top.second++];
                stack.push_back({c, up_ptr[c]});
            } else {
                if (pos >= n) return false;
                perm[(size_t)pos++] = top.first;
                stack.pop_back();
            }
        }
    }
    return pos == n;
}

// Recompute RIVSTOMAX, RIVELV (maps:716-722) and SET_FLDSTG (maps:969-999) and check
// the host tables bit for bit: the kernels derive them from the float32 sources.
int validate_tables(const cmf_hip_ctx* h, const cmf_map* mp) {
    const int n = h->d.nseqall, nl = h->d.nlfp;
    const double dfr = 1.0 / (double)nl;
    if (mp->dfrcinc != dfr) return fail(CMF_EMAP, "DFRCINC=%.17g differs from 1/NLFP", mp->dfrcinc);
    for (int i = 0; i < n; ++i) {
        const double rsm = h->o.lfplain ? (mp->rivlen[i] * mp->rivwth[i]) * mp->rivhgt[i] : (double)1.e18f;
        if (rsm != mp->rivstomax[i]) return fail(CMF_EMAP, "RIVSTOMAX(%d) differs from RIVLEN*RIVWTH*RIVHGT", i + 1);
        if (mp->elevtn[i] - mp->rivhgt[i] != mp->rivelv[i]) return fail(CMF_EMAP, "RIVELV(%d) differs from ELEVTN-RIVHGT", i + 1);
        double stopre = rsm, hpre = 0.0;
        const double dwthinc = (mp->grarea[i] * (1.0 / mp->rivlen[i])) * dfr;
        for (int l = 0; l < nl; ++l) {
            const double hl = mp->fldhgt[(size_t)l * n + i];
            const double dstonow = (mp->rivlen[i] * (mp->rivwth[i] + dwthinc * ((double)(l + 1) - 0.5))) * (hl - hpre);
            const double smax = stopre + dstonow;
            const double grd = (hl - hpre) * (1.0 / dwthinc);
            if (smax != mp->fldstomax[(size_t)l * n + i] || grd != mp->fldgrd[(size_t)l * n + i])
                return fail(CMF_EMAP, "FLDSTOMAX/FLDGRD(%d,%d) differ from the SET_FLDSTG recurrence", i + 1, l + 1);
            stopre = smax; hpre = hl;
        }
    }
    return 0;
}

void csr(int n, const std::vector<int32_t>& owner, std::vector<int32_t>& ptr, std::vector<int32_t>& idx,
         const std::vector<int32_t>& item) {
    ptr.assign((size_t)n + 1, 0);
    for (int32_t o : owner) ptr[(size_t)o + 1] += 1;
    for (int i = 0; i < n; ++i) ptr[(size_t)i + 1] += ptr[i];
    idx.assign(owner.size() ? owner.size() : 1, 0);
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (size_t k = 0; k < owner.size(); ++k) idx[(size_t)fill[owner[k]]++] = item[k];
}

template <class R>
DevState<R>& state_of(cmf_hip_ctx* h);
template <>
DevState<double>& state_of<double>(cmf_hip_ctx* h) { return h->S64; }
template <>
DevState<float>& state_of<float>(cmf_hip_ctx* h) { return h->S32; }

template <class R>
int set_state_t(cmf_hip_ctx* h, const cmf_state* st) {
    DevState<R>& S = state_of<R>(h);
    const size_t n = (size_t)h->d.nseqall;
    const size_t np = h->o.lpthout ? (size_t)h->d.npthout : 0;
    const size_t npl = np * (h->o.lpthout ? (size_t)h->d.npthlev : 0);
    auto putd = [&](double* dst, const double* src, size_t m) -> int { return put<double>(h, dst, src, m); };
    TRY(putd(S.rivsto, st->rivsto, n));
    TRY(putd(S.fldsto, st->fldsto, n));
    if (h->o.llevee) TRY(putd(S.levsto, st->levsto, n));
    if (h->o.ldamout) { TRY(putd(S.damsto, st->damsto, n)); TRY(putd(S.daminf, st->daminf, n)); }
    // D2RIVOUT_PRE lives in the RIVOUT array; a different D2RIVOUT is kept for the
    // RIVVEL of the next substep only (outflw:75 reads the pre-update RIVOUT).
    const double* rpre = st->rivout_pre ? st->rivout_pre : st->rivout;
    TRY(put<R>(h, S.rivout, rpre, n));
    if (st->rivout) {
        TRY(put<R>(h, (R*)h->vel_buf, st->rivout, n));
        h->vel_pending = true;
    }
    const double* fpre = st->fldout_pre ? st->fldout_pre : st->fldout;
    TRY(put<R>(h, S.fldout, fpre, n));
    TRY(put<R>(h, S.rivdph_pre, st->rivdph_pre, n));
    TRY(put<R>(h, S.fldsto_pre, st->fldsto_pre, n));
    if (h->o.lpthout) {
        const double* ppre = st->pthflw_pre ? st->pthflw_pre : st->pthflw;
        TRY(put<R>(h, S.pthflw, ppre, npl, false));
        TRY(put<R>(h, S.pthflwsum, st->pthflwsum, np, false));
    }
    TRY(put<R>(h, S.rivdph, st->rivdph, n));
    TRY(put<R>(h, S.flddph, st->flddph, n));
    TRY(put<R>(h, S.fldfrc, st->fldfrc, n));
    TRY(put<R>(h, S.fldare, st->fldare, n));
    TRY(put<R>(h, S.sfcelv, st->sfcelv, n));
    TRY(put<R>(h, S.storge, st->storge, n));
    TRY(put<R>(h, S.outflw, st->outflw, n));
    TRY(put<R>(h, S.rivvel, st->rivvel, n));
    if (h->o.llevee) TRY(put<R>(h, S.levdph, st->levdph, n));
    TRY(putd(S.rivinf, st->rivinf, n));
    TRY(putd(S.fldinf, st->fldinf, n));
    TRY(putd(S.pthout, st->pthout, n));
    return 0;
}

template <class R>
int get_state_t(cmf_hip_ctx* h, cmf_state* st) {
    DevState<R>& S = state_of<R>(h);
    const size_t n = (size_t)h->d.nseqall;
    const size_t np = h->o.lpthout ? (size_t)h->d.npthout : 0;
    const size_t npl = np * (h->o.lpthout ? (size_t)h->d.npthlev : 0);
    auto getd = [&](double* dst, const double* src, size_t m) -> int { return get<double>(h, dst, src, m); };
    HIPCHK(hipStreamSynchronize(h->stream));
    TRY(getd(st->rivsto, S.rivsto, n));
    TRY(getd(st->fldsto, S.fldsto, n));
    if (h->o.llevee) TRY(getd(st->levsto, S.levsto, n));
    if (h->o.ldamout) { TRY(getd(st->damsto, S.damsto, n)); TRY(getd(st->daminf, S.daminf, n)); }
    TRY(get<R>(h, st->rivout, h->vel_pending ? (const R*)h->vel_buf : S.rivout, n));
    TRY(get<R>(h, st->rivout_pre, S.rivout, n));
    TRY(get<R>(h, st->fldout, S.fldout, n));
    TRY(get<R>(h, st->fldout_pre, S.fldout, n));
    TRY(get<R>(h, st->rivdph_pre, S.rivdph_pre, n));
    TRY(get<R>(h, st->fldsto_pre, S.fldsto_pre, n));
    if (h->o.lpthout) {
        TRY(get<R>(h, st->pthflw, S.pthflw, npl, false));
        TRY(get<R>(h, st->pthflw_pre, S.pthflw, npl, false));
        TRY(get<R>(h, st->pthflwsum, S.pthflwsum, np, false));
        TRY(get<R>(h, st->pthflw_avg, S.pthflw_avg, npl, false));
        TRY(get<R>(h, st->pthflwsum_avg, S.pthflwsum_avg, np, false));
    }
    TRY(get<R>(h, st->rivdph, S.rivdph, n));
    TRY(get<R>(h, st->flddph, S.flddph, n));
    TRY(get<R>(h, st->fldfrc, S.fldfrc, n));
    TRY(get<R>(h, st->fldare, S.fldare, n));
    TRY(get<R>(h, st->sfcelv, S.sfcelv, n));
    TRY(get<R>(h, st->storge, S.storge, n));
    TRY(get<R>(h, st->outflw, S.outflw, n));
    TRY(get<R>(h, st->rivvel, S.rivvel, n));
    if (h->o.llevee) TRY(get<R>(h, st->levdph, S.levdph, n));
    TRY(getd(st->rivinf, S.rivinf, n));
    TRY(getd(st->fldinf, S.fldinf, n));
    TRY(getd(st->pthout, S.pthout, n));
    TRY(get<R>(h, st->rivout_avg, S.rivout_avg, n));
    TRY(get<R>(h, st->fldout_avg, S.fldout_avg, n));
    TRY(get<R>(h, st->outflw_avg, S.outflw_avg, n));
    TRY(get<R>(h, st->rivvel_avg, S.rivvel_avg, n));
    TRY(get<R>(h, st->pthout_avg, S.pthout_avg, n));
    TRY(get<R>(h, st->runoff_avg, S.runoff_avg, n));
    TRY(get<R>(h, st->outflw_max, S.outflw_max, n));
    TRY(get<R>(h, st->rivdph_max, S.rivdph_max, n));
    TRY(get<R>(h, st->storge_max, S.storge_max, n));
    if (h->o.ldamout) TRY(get<R>(h, st->daminf_avg, S.daminf_avg, n));
    return 0;
}

StepArgs base_args(const cmf_hip_ctx* h) {
    StepArgs A{};
    A.pgrv = h->o.pgrv; A.pmanfld = h->o.pmanfld; A.pdstmth = h->o.pdstmth;
    A.pcadp = h->o.pcadp; A.pminslp = h->o.pminslp;
    A.lfldout = h->o.lfldout; A.lpthout = h->o.lpthout; A.ldamout = h->o.ldamout; A.llevee = h->o.llevee;
    return A;
}

int ev_record(cmf_hip_ctx* h, size_t k) {
    while (h->ev.size() <= k) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        h->ev.push_back(e);
    }
    HIPCHK(hipEventRecord(h->ev[k], h->stream));
    return 0;
}

template <class R>
int advance_t(cmf_hip_ctx* h, double* dt, int32_t nt_cap, int32_t* nt_out, double* glb, int32_t glb_cap) {
    DevState<R>& S = state_of<R>(h);
    const double dt_def = *dt;
    StepArgs A = base_args(h);
    A.dt = dt_def;
    A.gdt = A.pgrv * dt_def;
    size_t ne = 0;
    if (h->timing) TRY(ev_record(h, ne++));
    // outer-step start: CMF_PHYSICS_FLDSTG (physics:54) and CALC_ADPSTP (:58)
    launch_stage<R>(h->M, S, A, h->o.ladpstp, dt_def, h->stream);
    HIPCHK(hipGetLastError());
    if (h->timing) TRY(ev_record(h, ne++));
    int32_t nt = 1;
    double dts = dt_def;
    if (h->o.ladpstp) {
        HIPCHK(hipMemcpyAsync(h->h_ctl, S.ctl, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        const double dt_min = h->h_ctl[0];
        nt = (int32_t)(dt_def * (1.0 / dt_min) - K_0P01F) + 1;       // physics:205
        dts = dt_def * (double)(1.0f / (float)nt);                   // physics:206
    }
    const int32_t nrun = (nt_cap > 0 && nt_cap < nt) ? nt_cap : nt;
    if (nrun > h->hist_cap) {                           // grow the per-substep budget buffers
        const int cap = nrun + nrun / 4;
        double *p = nullptr, *q = nullptr;
        TRY(dalloc(h, &p, (size_t)cap * NGLB));
        TRY(dalloc(h, &q, (size_t)cap * h->nb_cells * NGLB));
        (void)hipFree(S.glb_hist);
        (void)hipFree(S.partials);
        for (auto& a : h->allocs) if (a == (void*)S.glb_hist || a == (void*)S.partials) a = nullptr;
        S.glb_hist = p;
        S.partials = q;
        h->hist_cap = cap;
    }
    double nadd = 0.0;                                               // diag:108
    for (int32_t it = 0; it < nrun; ++it) nadd = nadd + dts;
    A.dt = dts;
    A.gdt = A.pgrv * dts;
    A.nadd = nadd;
    A.nrun = nrun;
    for (int32_t it = 0; it < nrun; ++it) {
        A.first = (it == 0);
        A.last = (it == nrun - 1);
        A.it = it;
        DevState<R> Sk = S;
        Sk.rivdph_next = S.rivdph_pre;
        Sk.rivout_vel = (h->vel_pending && it == 0) ? (const R*)h->vel_buf : S.rivout;
        if (h->timing) TRY(ev_record(h, ne++));
        launch_outflow<R>(h->M, Sk, A, h->stream);
        if (h->timing) { TRY(ev_record(h, ne++)); TRY(ev_record(h, ne++)); }
        launch_rate<R>(h->M, Sk, A, h->stream);
        if (h->timing) { TRY(ev_record(h, ne++)); TRY(ev_record(h, ne++)); }
        launch_update<R>(h->M, Sk, A, h->stream);
        if (h->timing) TRY(ev_record(h, ne++));
        std::swap(S.rivdph, S.rivdph_pre);                            // CALC_VARS_PRE for RIVDPH
    }
    launch_budget(h->M, S.partials, nrun, S.glb_hist, h->stream);    // CALC_WATBAL sums, all substeps
    HIPCHK(hipGetLastError());
    h->vel_pending = false;
    if (glb && glb_cap > 0) {
        const int32_t m = nrun < glb_cap ? nrun : glb_cap;
        HIPCHK(hipMemcpyAsync(glb, S.glb_hist, (size_t)m * NGLB * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->timing) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        h->t_ms[3] += ms;
        for (int32_t it = 0; it < nrun; ++it) {
            for (int k = 0; k < 3; ++k) {
                const size_t e0 = 2 + 6 * (size_t)it + 2 * k;
                HIPCHK(hipEventElapsedTime(&ms, h->ev[e0], h->ev[e0 + 1]));
                h->t_ms[k] += ms;
            }
        }
        h->t_ms[4] += nrun;
    }
    *nt_out = nt;
    *dt = dt_def;                                                    // physics:145
    return 0;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* cmf_hip_last_error(void) { return g_err.c_str(); }

int cmf_hip_create(const cmf_dims* dims, const cmf_options* opts, cmf_handle* out) {
    if (!dims || !opts || !out) return fail(CMF_EINVAL, "NULL argument");
    *out = nullptr;
    if (dims->nseqall <= 0 || dims->nseqriv < 0 || dims->nseqriv > dims->nseqall)
        return fail(CMF_EINVAL, "bad NSEQALL/NSEQRIV (%d, %d)", dims->nseqall, dims->nseqriv);
    if (dims->nlfp <= 0 || dims->nlfp > NLFP_MAX) return fail(CMF_EINVAL, "NLFP=%d outside 1..%d", dims->nlfp, NLFP_MAX);
    if (opts->lpthout && (dims->npthout <= 0 || dims->npthlev <= 0))
        return fail(CMF_EINVAL, "LPTHOUT needs NPTHOUT>0 and NPTHLEV>0");
    if (opts->ldamout && dims->ndam <= 0) return fail(CMF_EINVAL, "LDAMOUT needs NDAM>0");
    if (opts->precision != CMF_PREC_FP64 && opts->precision != CMF_PREC_MIXED)
        return fail(CMF_EINVAL, "unknown precision %d", opts->precision);
    cmf_hip_ctx* h = new cmf_hip_ctx();
    h->d = *dims;
    h->o = *opts;
    h->nb_cells = (dims->nseqall + BLOCK - 1) / BLOCK;
    int rc = 0;
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(CMF_EHIP, "no HIP device / stream");
    }
    if (hipHostMalloc((void**)&h->h_ctl, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(h->stream);
        delete h;
        return fail(CMF_EHIP, "hipHostMalloc failed");
    }
    rc = (opts->precision == CMF_PREC_FP64) ? alloc_state<double>(h, h->S64) : alloc_state<float>(h, h->S32);
    if (rc != 0) {
        std::string keep = g_err;
        cmf_hip_destroy(h);
        g_err = keep;
        return rc;
    }
    *out = h;
    return 0;
}

void cmf_hip_destroy(cmf_handle h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) if (p) (void)hipFree(p);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    if (h->h_ctl) (void)hipHostFree(h->h_ctl);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int cmf_hip_set_map(cmf_handle h, const cmf_map* mp) {
    if (!h || !mp) return fail(CMF_EINVAL, "NULL argument");
    if (h->map_set) return fail(CMF_ESTATE, "map already set for this handle");
    const int n = h->d.nseqall, nr = h->d.nseqriv, nl = h->d.nlfp;
    if (!mp->i1next || !mp->rivstomax || !mp->rivelv || !mp->dwnelv || !mp->fldhgt || !mp->fldstomax ||
        !mp->fldgrd || !mp->i2mask)
        return fail(CMF_EINVAL, "required map array missing");
    DevMap& M = h->M;
    M.n = n; M.nseqriv = nr; M.nlfp = nl;
    M.lfplain = h->o.lfplain;
    M.dfrcinc = mp->dfrcinc;
    // topology: I1NEXT -> 0-based, upstream CSR in ascending source order
    std::vector<int32_t> nxt(n);
    for (int i = 0; i < n; ++i) {
        const int32_t v = mp->i1next[i];
        if (i < nr) {
            if (v < 1 || v > n) return fail(CMF_EMAP, "I1NEXT(%d)=%d outside 1..NSEQALL for a river cell", i + 1, v);
            nxt[i] = v - 1;
        } else {
            if (v >= 0) return fail(CMF_EMAP, "I1NEXT(%d)=%d: mouth cells need a negative code", i + 1, v);
            nxt[i] = v;
        }
    }
    std::vector<int32_t> owner(nr), item(nr), up_ptr, up_idx;
    for (int i = 0; i < nr; ++i) { owner[i] = nxt[i]; item[i] = i; }
    csr(n, owner, up_ptr, up_idx, item);
    // device cell order
    const char* ord = getenv("CMF_HIP_ORDER");
    h->permuted = !(ord && std::string(ord) == "nseq");
    if (h->permuted && !dfs_order(n, nr, up_ptr, up_idx, h->perm))
        return fail(CMF_EMAP, "river network is not a forest draining to the mouth cells");
    if (!h->permuted) { h->perm.resize((size_t)n); for (int i = 0; i < n; ++i) h->perm[i] = i; }
    h->iperm.assign((size_t)n, 0);
    for (int p = 0; p < n; ++p) h->iperm[(size_t)h->perm[p]] = p;
    const std::vector<int32_t>& P = h->perm;
    const std::vector<int32_t>& IP = h->iperm;
    // topology in device positions; each upstream list keeps ascending NSEQ order
    std::vector<int32_t> dnext((size_t)n), dptr((size_t)n + 1, 0), didx(up_idx.size());
    for (int p = 0; p < n; ++p) {
        const int32_t v = nxt[(size_t)P[p]];
        dnext[(size_t)p] = v >= 0 ? IP[(size_t)v] : v;
        dptr[(size_t)p + 1] = dptr[p] + (up_ptr[(size_t)P[p] + 1] - up_ptr[(size_t)P[p]]);
    }
    for (int p = 0; p < n; ++p) {
        int32_t k = dptr[p];
        for (int32_t q = up_ptr[(size_t)P[p]]; q < up_ptr[(size_t)P[p] + 1]; ++q) didx[(size_t)k++] = IP[(size_t)up_idx[q]];
    }
    int32_t* p32 = nullptr;
    TRY(put_raw(h, &p32, dnext.data(), (size_t)n)); M.next = p32;
    TRY(put_raw(h, &p32, dptr.data(), dptr.size())); M.up_ptr = p32;
    TRY(put_raw(h, &p32, didx.data(), didx.size())); M.up_idx = p32;
    TRY(validate_tables(h, mp));
    TRY(put_f32(h, &M.grarea, mp->grarea, n, "GRAREA"));
    TRY(put_f32(h, &M.elevtn, mp->elevtn, n, "ELEVTN"));
    TRY(put_f32(h, &M.nxtdst, mp->nxtdst, n, "NXTDST"));
    TRY(put_f32(h, &M.rivlen, mp->rivlen, n, "RIVLEN"));
    TRY(put_f32(h, &M.rivwth, mp->rivwth, n, "RIVWTH"));
    TRY(put_f32(h, &M.rivhgt, mp->rivhgt, n, "RIVHGT"));
    TRY(put_f32(h, &M.rivman, mp->rivman, n, "RIVMAN"));
    TRY(put_f32(h, &M.fldhgt, mp->fldhgt, (size_t)n, "FLDHGT", (size_t)nl));
    double* pd = nullptr;
    {
        std::vector<double> v = to_dev(h, mp->dwnelv, (size_t)n);
        TRY(put_raw(h, &pd, v.data(), (size_t)n)); M.dwnelv = pd;
        std::vector<int32_t> m2 = to_dev(h, mp->i2mask, (size_t)n);
        TRY(put_raw(h, &p32, m2.data(), (size_t)n)); M.i2mask = p32;
    }
    // bifurcation paths (maps:531-639)
    if (h->o.lpthout) {
        const int np = h->d.npthout, nlv = h->d.npthlev;
        M.npth = np; M.npthlev = nlv;
        if (!mp->pth_upst || !mp->pth_down || !mp->pth_dst || !mp->pth_elv || !mp->pth_wth || !mp->pth_man)
            return fail(CMF_EINVAL, "LPTHOUT: path tables missing");
        std::vector<int32_t> pu(np), pdn(np), oo, oi, io, ii;
        for (int p = 0; p < np; ++p) {
            pu[p] = mp->pth_upst[p] > 0 ? mp->pth_upst[p] - 1 : -1;
            pdn[p] = mp->pth_down[p] > 0 ? mp->pth_down[p] - 1 : -1;
            if (pu[p] >= n || pdn[p] >= n) return fail(CMF_EMAP, "path %d endpoint outside NSEQALL", p + 1);
            if (pu[p] >= 0) pu[p] = IP[(size_t)pu[p]];
            if (pdn[p] >= 0) pdn[p] = IP[(size_t)pdn[p]];
            if (pu[p] >= 0 && pdn[p] >= 0) {
                oo.push_back(pu[p]); oi.push_back(p);
                io.push_back(pdn[p]); ii.push_back(p);
            }
        }
        std::vector<int32_t> optr, oidx, iptr, iidx;
        csr(n, oo, optr, oidx, oi);
        csr(n, io, iptr, iidx, ii);
        TRY(put_raw(h, &p32, pu.data(), (size_t)np)); M.pth_upst = p32;
        TRY(put_raw(h, &p32, pdn.data(), (size_t)np)); M.pth_down = p32;
        TRY(put_raw(h, &p32, optr.data(), optr.size())); M.pout_ptr = p32;
        TRY(put_raw(h, &p32, oidx.data(), oidx.size())); M.pout_idx = p32;
        TRY(put_raw(h, &p32, iptr.data(), iptr.size())); M.pinf_ptr = p32;
        TRY(put_raw(h, &p32, iidx.data(), iidx.size())); M.pinf_idx = p32;
        TRY(put_raw(h, &pd, mp->pth_dst, (size_t)np)); M.pth_dst = pd;
        TRY(put_raw(h, &pd, mp->pth_elv, (size_t)np * nlv)); M.pth_elv = pd;
        TRY(put_raw(h, &pd, mp->pth_wth, (size_t)np * nlv)); M.pth_wth = pd;
        TRY(put_raw(h, &pd, mp->pth_man, (size_t)nlv)); M.pth_man = pd;
    }
    // levee: compact table of the cells that carry a levee (levee:184-259)
    if (h->o.llevee) {
        const double* L[7] = {mp->levhgt, mp->levfrc, mp->bashgt, mp->levdst, mp->levbassto, mp->levtopsto, mp->levfilsto};
        for (int k = 0; k < 7; ++k) if (!L[k]) return fail(CMF_EINVAL, "LLEVEE: levee tables missing");
        std::vector<int32_t> lidx(n, -1);
        std::vector<int32_t> cells;
        for (int p = 0; p < n; ++p) {                                   // device order
            const int i = P[(size_t)p];
            const bool plain = mp->levhgt[i] == 0.0 && mp->levfrc[i] == 1.0 && mp->bashgt[i] == K_LEVSENT &&
                               mp->levdst[i] == K_LEVSENT && mp->levbassto[i] == K_LEVSENT &&
                               mp->levtopsto[i] == K_LEVSENT && mp->levfilsto[i] == K_LEVSENT;
            if (!plain) { lidx[p] = (int32_t)cells.size(); cells.push_back(i); }
        }
        const int nc = (int)cells.size();
        std::vector<double> tab((size_t)7 * (nc ? nc : 1));
        for (int k = 0; k < 7; ++k)
            for (int c = 0; c < nc; ++c) tab[(size_t)k * nc + c] = L[k][cells[c]];
        M.nlevcell = nc;
        TRY(put_raw(h, &p32, lidx.data(), (size_t)n)); M.lev_idx = p32;
        TRY(put_raw(h, &pd, tab.data(), tab.size())); M.lev_tab = pd;
    }
    // dams (damout:123-308)
    if (h->o.ldamout) {
        const int nd = h->d.ndam;
        M.ndam = nd;
        if (!mp->i1dam || !mp->damseq || !mp->damstat || !mp->dam_convol || !mp->dam_emevol || !mp->dam_adjvol ||
            !mp->dam_qa || !mp->dam_qn || !mp->dam_qf)
            return fail(CMF_EINVAL, "LDAMOUT: dam tables missing");
        std::vector<int32_t> doc(n, -1);
        for (int d = 0; d < nd; ++d) {
            if (mp->damstat[d] == -9999) continue;
            const int s = mp->damseq[d] - 1;
            if (s < 0 || s >= n) return fail(CMF_EMAP, "DamSeq(%d)=%d outside NSEQALL", d + 1, mp->damseq[d]);
            doc[(size_t)IP[(size_t)s]] = d;
        }
        std::vector<int32_t> i1d = to_dev(h, mp->i1dam, (size_t)n);
        TRY(put_raw(h, &p32, i1d.data(), (size_t)n)); M.i1dam = p32;
        TRY(put_raw(h, &p32, doc.data(), (size_t)n)); M.dam_of_cell = p32;
        TRY(put_raw(h, &p32, mp->damstat, (size_t)nd)); M.damstat = p32;
        TRY(put_raw(h, &pd, mp->dam_convol, (size_t)nd)); M.dam_convol = pd;
        TRY(put_raw(h, &pd, mp->dam_emevol, (size_t)nd)); M.dam_emevol = pd;
        TRY(put_raw(h, &pd, mp->dam_adjvol, (size_t)nd)); M.dam_adjvol = pd;
        TRY(put_raw(h, &pd, mp->dam_qa, (size_t)nd)); M.dam_qa = pd;
        TRY(put_raw(h, &pd, mp->dam_qn, (size_t)nd)); M.dam_qn = pd;
        TRY(put_raw(h, &pd, mp->dam_qf, (size_t)nd)); M.dam_qf = pd;
    }
    h->map_set = true;
    return 0;
}

int cmf_hip_set_state(cmf_handle h, const cmf_state* st) {
    if (!h || !st) return fail(CMF_EINVAL, "NULL argument");
    return h->o.precision == CMF_PREC_FP64 ? set_state_t<double>(h, st) : set_state_t<float>(h, st);
}

int cmf_hip_get_state(cmf_handle h, cmf_state* st) {
    if (!h || !st) return fail(CMF_EINVAL, "NULL argument");
    return h->o.precision == CMF_PREC_FP64 ? get_state_t<double>(h, st) : get_state_t<float>(h, st);
}

int cmf_hip_set_runoff(cmf_handle h, const double* runoff) {
    if (!h || !runoff) return fail(CMF_EINVAL, "NULL argument");
    const size_t n = (size_t)h->d.nseqall;
    return h->o.precision == CMF_PREC_FP64 ? put<double>(h, h->S64.runoff, runoff, n)
                                           : put<float>(h, h->S32.runoff, runoff, n);
}

int cmf_hip_physics_fldstg(cmf_handle h) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    StepArgs A = base_args(h);
    if (h->o.precision == CMF_PREC_FP64) launch_stage<double>(h->M, h->S64, A, 0, 0.0, h->stream);
    else launch_stage<float>(h->M, h->S32, A, 0, 0.0, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_hip_physics_advance_capped(cmf_handle h, double* dt, int32_t nt_cap, int32_t* nt_out, double* glb,
                                   int32_t glb_cap) {
    if (!h || !dt || !nt_out) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (!(*dt > 0.0)) return fail(CMF_EINVAL, "DT must be positive");
    return h->o.precision == CMF_PREC_FP64 ? advance_t<double>(h, dt, nt_cap, nt_out, glb, glb_cap)
                                           : advance_t<float>(h, dt, nt_cap, nt_out, glb, glb_cap);
}

int cmf_hip_physics_advance(cmf_handle h, double* dt, int32_t* nt_out, double* glb, int32_t glb_cap) {
    return cmf_hip_physics_advance_capped(h, dt, 0, nt_out, glb, glb_cap);
}

int cmf_hip_synchronize(cmf_handle h) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_hip_kernel_times(cmf_handle h, int32_t enable, double* out5) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (out5) for (int k = 0; k < 5; ++k) out5[k] = h->t_ms[k];
    for (int k = 0; k < 5; ++k) h->t_ms[k] = 0.0;
    h->timing = enable != 0;
    return 0;
}

}  // extern "C"
