// cmf_capi.hip -- C-ABI of include/cmf_hip.h: device-resident model state, map
// upload and validation, and the outer-step driver that replaces
// CMF_PHYSICS_ADVANCE / CMF_PHYSICS_FLDSTG (src/cmf_ctrl_physics_mod.F90).
#include "cmf_hip.h"
#include "cmf_kernels.hpp"

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

using namespace cmf;

// Scratch of the allocation-free advance (cmf_hip_physics_advance_noalloc), one copy
// per GPU: the CFL minimum, the CFL bound of every stage block, the budget rows of
// the substeps and the per-tile budget sums of one substep.
constexpr size_t NA_CFL_BLOCKS = 65536;          // stage blocks of 256 slots: NSEQALL <= 16,777,216
constexpr size_t NA_HIST = 4096;                 // substeps per outer step
constexpr size_t NA_TILES = 65536;               // update tiles of one substep
constexpr size_t NA_OFF_CFL = 0;                 // 1 value, padded to 256 bytes
constexpr size_t NA_OFF_PART_CFL = 32;
constexpr size_t NA_OFF_HIST = NA_OFF_PART_CFL + NA_CFL_BLOCKS;
constexpr size_t NA_OFF_TILES = NA_OFF_HIST + NA_HIST * NGLB;
constexpr size_t NA_DOUBLES = NA_OFF_TILES + NA_TILES * NGLB;   // 831,520 doubles = 6.65 MB
static __device__ __align__(256) double scratch[NA_DOUBLES];

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

// one registered per-slot state array (for the tuner's snapshot/restore)
struct StateArr {
    void** slot;           // address of the pointer member in DevState (follows swaps)
    size_t bytes;
};

struct cmf_hip_ctx {
    cmf_dims d{};
    cmf_options o{};
    bool hf32 = false;                // JPRB host arrays are float (jprb = 4)
    int device = 0;
    hipStream_t stream = nullptr;
    bool map_set = false;
    std::vector<void*> allocs;
    DevMap M{};
    DevState<double> S64{};
    DevState<float> S32{};
    void* vel_buf = nullptr;          // D2RIVOUT when it differs from D2RIVOUT_PRE
    bool vel_pending = false;
    void* fld_buf = nullptr;          // D2FLDOUT when it differs from D2FLDOUT_PRE (allocated when needed)
    bool fld_pending = false;
    void* pth_buf = nullptr;          // D1PTHFLW when it differs from D1PTHFLW_PRE (allocated when needed)
    bool pth_pending = false;
    double* h_ctl = nullptr;          // pinned scratch for small device->host reads
    int32_t* d_flag = nullptr;        // device flag of k_daminf_nonzero
    int32_t daminf_sparse = 1;        // DAMINF == 0 in every cell with I1DAM == 0
    // ensemble: ndom domains concatenated in NSEQ, nmem = 2^lgm members each
    int ndom = 1, nmem = 1, lgm = 0, ninst = 1;
    std::vector<int32_t> dom_off;     // NSEQ offset of each domain (+ total)
    int32_t* d_dom_off = nullptr;     // first cell of each domain in device order (+ total)
    double *d_ist_part = nullptr, *d_ist_out = nullptr;   // cmf_hip_instance_storage scratch
    int ist_nch = 0;
    InstStep* h_inst = nullptr;       // pinned [ninst]
    unsigned long long* h_cfl = nullptr;  // pinned [ninst]
    std::vector<int32_t> last_nt;     // NT of each instance in the last advance
    std::vector<double> last_dt;
    std::vector<double> last_dtmin;   // DT_MIN of CALC_ADPSTP of each instance (0 without LADPSTP)
    InstStep* d_inst = nullptr;
    double *d_rbase = nullptr, *d_rfac = nullptr;   // staging of cmf_hip_set_runoff_scaled
    RoffMap roff{};                                  // input matrix of cmf_hip_set_inpmat
    OutState out{};                                  // output-period accumulators
    double* dwnelv_dev = nullptr;                    // D2DWNELV in the device map (boundary updates)
    int upsinf_on = 0;                               // LUPSINF switched on by cmf_hip_set_upsinf
    int lkine = 0;                                   // LKINE set by cmf_hip_set_kinematic
    int lslpmix = 0;                                 // LSLPMIX set by cmf_hip_set_slopemix
    double *d_meansl = nullptr, *d_sealev = nullptr; // staging of cmf_hip_update_boundary
    double nadd_out = 0.0;
    double* d_grid = nullptr;                        // staging of cmf_hip_set_runoff_grid
    size_t grid_cap = 0;
    double* d_records = nullptr;                     // forcing fields resident in HBM
    int32_t nrecords = 0;
    size_t partial_cap = 0;           // doubles in S.partials
    int hist_cap = 0;
    // device cell order: perm[p] = NSEQ index of device position p (identity when
    // CMF_HIP_ORDER=nseq); the C-ABI always speaks NSEQ order
    std::vector<int32_t> perm, iperm;
    bool permuted = false;
    int32_t* d_perm = nullptr;        // perm on the device (NULL when not permuted)
    void* stage = nullptr;            // staging buffer of host <-> device cell-array transfers
    size_t stage_cap = 0;             // bytes
    // host index of every device slot: cells (ns), path levels (npthlev*nsp), paths (nsp)
    std::vector<int64_t> hx_cell, hx_pthl, hx_pth;
    // launch configuration
    cmf_tune tune{};
    int coop_max = 0;                 // co-resident blocks of the fused kernel at tune.bs_fused
    // graphs of whole outer-step substep sequences, keyed by (NT, buffer parity)
    struct GraphEntry { int32_t nt; int parity; hipGraphExec_t exec; };
    std::vector<GraphEntry> graphs;
    void* dph_a = nullptr;            // the RIVDPH buffer allocated first (parity reference)
    std::vector<StateArr> state_arrs;
    void* snapshot = nullptr;
    size_t snapshot_bytes = 0;
    // per-kernel timing (HIP events on the handle's stream)
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double t_ms[5] = {0, 0, 0, 0, 0};
    std::vector<double> inst_ms;      // per-instance completion of the last timed advance
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

// state array of ns/nsp slots, registered for the tuner's snapshot
template <class R, class T>
int salloc(cmf_hip_ctx* h, T** p, size_t n, bool snap = true) {
    TRY(dalloc(h, p, n));
    if (snap) h->state_arrs.push_back({reinterpret_cast<void**>(p), (n ? n : 1) * sizeof(T)});
    return 0;
}

template <class R>
int alloc_state(cmf_hip_ctx* h, DevState<R>& S) {
    const size_t ns = (size_t)h->d.nseqall * h->nmem;
    const size_t nsp = (size_t)(h->o.lpthout ? h->d.npthout : 0) * h->nmem;
    const size_t nspl = nsp * (size_t)(h->o.lpthout ? h->d.npthlev : 0);
    const size_t nlev = h->o.llevee ? ns : 1, ndm = h->o.ldamout ? ns : 1;
    TRY((salloc<R>(h, &S.rivsto, ns))); TRY((salloc<R>(h, &S.fldsto, ns)));
    TRY((salloc<R>(h, &S.levsto, nlev)));
    TRY((salloc<R>(h, &S.damsto, ndm))); TRY((salloc<R>(h, &S.daminf, ndm)));
    TRY((salloc<R>(h, &S.rivout, ns))); TRY((salloc<R>(h, &S.fldout, ns)));
    TRY((salloc<R>(h, &S.qr_u, ns, false))); TRY((salloc<R>(h, &S.qf_u, ns, false)));
    TRY((salloc<R>(h, &S.rivdph_pre, ns))); TRY((salloc<R>(h, &S.fldsto_pre, ns)));
    TRY((salloc<R>(h, &S.rivdph, ns))); TRY((salloc<R>(h, &S.flddph, ns))); TRY((salloc<R>(h, &S.fldfrc, ns)));
    TRY((salloc<R>(h, &S.fldare, ns))); TRY((salloc<R>(h, &S.sfcelv, ns))); TRY((salloc<R>(h, &S.storge, ns)));
    TRY((salloc<R>(h, &S.outflw, ns))); TRY((salloc<R>(h, &S.rivvel, ns)));
    TRY((salloc<R>(h, &S.levdph, nlev)));
    TRY((salloc<R>(h, &S.rivinf, ns))); TRY((salloc<R>(h, &S.fldinf, ns))); TRY((salloc<R>(h, &S.pthout, ns)));
    TRY((salloc<R>(h, &S.rate, ns, false))); TRY((salloc<R>(h, &S.runoff, ns)));
    TRY((salloc<R>(h, &S.pthflw, nspl))); TRY((salloc<R>(h, &S.pthflw_u, nspl, false)));
    TRY((salloc<R>(h, &S.pthflwsum, nsp))); TRY((salloc<R>(h, &S.pthflwsum_u, nsp, false)));
    TRY((salloc<R>(h, &S.rivout_avg, ns))); TRY((salloc<R>(h, &S.fldout_avg, ns)));
    TRY((salloc<R>(h, &S.outflw_avg, ns))); TRY((salloc<R>(h, &S.rivvel_avg, ns)));
    TRY((salloc<R>(h, &S.pthout_avg, ns))); TRY((salloc<R>(h, &S.runoff_avg, ns)));
    TRY((salloc<R>(h, &S.outflw_max, ns))); TRY((salloc<R>(h, &S.rivdph_max, ns)));
    TRY((salloc<R>(h, &S.storge_max, ns)));
    TRY((salloc<R>(h, &S.daminf_avg, ndm)));
    TRY((salloc<R>(h, &S.pthflw_avg, nspl))); TRY((salloc<R>(h, &S.pthflwsum_avg, nsp)));
    h->hist_cap = 16;                                   // grown on demand (substeps per outer step)
    h->partial_cap = 0;
    S.partials = nullptr;
    TRY(dalloc(h, &S.glb_hist, (size_t)h->hist_cap * NGLB));
    TRY(dalloc(h, &S.cfl, (size_t)h->ninst));
    TRY(dalloc(h, &S.cfl_part, (ns + 255) / 256));                // one value per stage block
    TRY(dalloc(h, &h->d_inst, (size_t)h->ninst));
    S.inst = h->d_inst;
    TRY(dalloc(h, &S.bar, 2));
    TRY(dalloc(h, &S.err, 1));
    S.rivout_vel = S.rivout;
    R* vb = nullptr;
    TRY(dalloc(h, &vb, ns));
    h->vel_buf = vb;
    h->dph_a = S.rivdph;
    return 0;
}

bool is_f32(double x) { return (double)(float)x == x; }

enum Kind { K_CELL, K_PTHL, K_PTH };

const std::vector<int64_t>& hx_of(const cmf_hip_ctx* h, Kind k) {
    return k == K_CELL ? h->hx_cell : (k == K_PTHL ? h->hx_pthl : h->hx_pth);
}

// Host <-> device copies of per-slot arrays with element-type conversion. Host
// arrays hold nmem member planes, each in NSEQ (or path) order; the device holds
// the member-minor slots in device cell order.
// size of one member plane of a host array of this kind
size_t plane_of(const cmf_hip_ctx* h, Kind k) {
    const size_t np = h->o.lpthout ? (size_t)h->d.npthout : 0;
    return k == K_CELL ? (size_t)h->d.nseqall : (k == K_PTHL ? np * (size_t)h->d.npthlev : np);
}

int stage_bytes(cmf_hip_ctx* h, size_t bytes) {
    if (bytes <= h->stage_cap) return 0;
    char* p = nullptr;
    TRY(dalloc(h, &p, bytes));
    if (h->stage) {
        (void)hipFree(h->stage);
        for (auto& a : h->allocs) if (a == h->stage) a = nullptr;
    }
    h->stage = p;
    h->stage_cap = bytes;
    return 0;
}

// Host -> device copy of a per-slot array of element type H into R. Cell arrays go
// to the device as the host holds them and are permuted and converted there;
// path arrays are reordered on the host. shared: the host array holds one plane
// that every member receives.
template <class R, class H>
int put_h(cmf_hip_ctx* h, R* dst, const H* src, Kind kind, bool shared = false) {
    const std::vector<int64_t>& hx = hx_of(h, kind);
    const size_t n = hx.size();
    if (!src || n == 0) return 0;
    if (shared && h->nmem == 1) shared = false;
    const bool ident = !h->permuted && h->nmem == 1;
    if (std::is_same<R, H>::value && ident) {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(R), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (kind == K_CELL) {
        const size_t hn = shared ? (size_t)h->d.nseqall : n;
        TRY(stage_bytes(h, hn * sizeof(H)));
        HIPCHK(hipMemcpyAsync(h->stage, src, hn * sizeof(H), hipMemcpyHostToDevice, h->stream));
        launch_cells_from_host<R, H>(dst, static_cast<const H*>(h->stage), h->d_perm, h->d.nseqall, h->lgm, (int)n,
                                     shared ? 1 : 0, h->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    const size_t plane = plane_of(h, kind);
    std::vector<R> tmp(n);
    for (size_t k = 0; k < n; ++k) tmp[k] = (R)src[shared ? (size_t)hx[k] % plane : (size_t)hx[k]];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * sizeof(R), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

template <class R>
int put(cmf_hip_ctx* h, R* dst, const double* src, Kind kind, bool shared = false) {
    return put_h<R, double>(h, dst, src, kind, shared);
}

// Device -> host copy of a per-slot array into host element type H (cell arrays
// permuted on the device, path arrays on the host)
template <class R, class H>
int get_h(cmf_hip_ctx* h, H* dst, const R* src, Kind kind) {
    const std::vector<int64_t>& hx = hx_of(h, kind);
    const size_t n = hx.size();
    if (!dst || n == 0) return 0;
    const bool ident = !h->permuted && h->nmem == 1;
    if (std::is_same<R, H>::value && ident) {
        HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (kind == K_CELL) {
        TRY(stage_bytes(h, n * sizeof(H)));
        launch_cells_to_host<R, H>(static_cast<H*>(h->stage), src, h->d_perm, h->d.nseqall, h->lgm, (int)n, h->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dst, h->stage, n * sizeof(H), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
    std::vector<R> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, n * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < n; ++k) dst[hx[k]] = (H)tmp[k];
    return 0;
}

template <class R>
int get(cmf_hip_ctx* h, double* dst, const R* src, Kind kind) {
    return get_h<R, double>(h, dst, src, kind);
}

// JPRB host arrays are double, or float for a handle created with jprb = 4 (the
// reference's -DSinglePrec_CMF build: D2* maps and diagnostics, D1PTHFLW, dam
// and levee tables, D2RUNOFF are JPRB; the P2* storages are JPRD = double in
// both builds). Widen gives double copies of float arrays for the duration of a
// call; Narrow gives double buffers whose values finish() rounds into the
// caller's float arrays.
struct Widen {
    const cmf_hip_ctx* h;
    std::vector<std::vector<double>> keep;
    explicit Widen(const cmf_hip_ctx* hh) : h(hh) {}
    const double* operator()(const double* p, size_t n) {
        if (!p || !h->hf32) return p;
        const float* f = reinterpret_cast<const float*>(p);
        keep.emplace_back(f, f + n);
        return keep.back().data();
    }
};
struct Narrow {
    const cmf_hip_ctx* h;
    std::vector<std::pair<float*, std::vector<double>>> keep;
    explicit Narrow(const cmf_hip_ctx* hh) : h(hh) {}
    double* operator()(double* p, size_t n) {
        if (!p || !h->hf32) return p;
        keep.emplace_back(reinterpret_cast<float*>(p), std::vector<double>(n));
        return keep.back().second.data();
    }
    void finish() {
        for (auto& k : keep)
            for (size_t i = 0; i < k.second.size(); ++i) k.first[i] = (float)k.second[i];
    }
};

// the JPRB members of cmf_state and the kind of array each is
struct StateField { double* cmf_state::*f; int kind; };   // kind: 0 cells, 1 path levels, 2 paths
const StateField JPRB_STATE[] = {
    {&cmf_state::rivout, 0}, {&cmf_state::fldout, 0}, {&cmf_state::rivout_pre, 0}, {&cmf_state::fldout_pre, 0},
    {&cmf_state::rivdph_pre, 0}, {&cmf_state::fldsto_pre, 0}, {&cmf_state::pthflw, 1}, {&cmf_state::pthflw_pre, 1},
    {&cmf_state::rivdph, 0}, {&cmf_state::flddph, 0}, {&cmf_state::fldfrc, 0}, {&cmf_state::fldare, 0},
    {&cmf_state::sfcelv, 0}, {&cmf_state::storge, 0}, {&cmf_state::outflw, 0}, {&cmf_state::rivvel, 0},
    {&cmf_state::rivinf, 0}, {&cmf_state::fldinf, 0}, {&cmf_state::pthout, 0}, {&cmf_state::levdph, 0},
    {&cmf_state::pthflwsum, 2}, {&cmf_state::rivout_avg, 0}, {&cmf_state::fldout_avg, 0},
    {&cmf_state::outflw_avg, 0}, {&cmf_state::rivvel_avg, 0}, {&cmf_state::pthout_avg, 0},
    {&cmf_state::runoff_avg, 0}, {&cmf_state::outflw_max, 0}, {&cmf_state::rivdph_max, 0},
    {&cmf_state::storge_max, 0}, {&cmf_state::daminf_avg, 0}, {&cmf_state::pthflw_avg, 1},
    {&cmf_state::pthflwsum_avg, 2}};

size_t field_size(const cmf_hip_ctx* h, int kind, size_t planes) {
    const size_t np = h->o.lpthout ? (size_t)h->d.npthout : 0;
    const size_t one = kind == 0 ? (size_t)h->d.nseqall : (kind == 1 ? np * (size_t)h->d.npthlev : np);
    return one * planes;
}

cmf_state widen_state(const cmf_hip_ctx* h, const cmf_state* st, size_t planes, Widen& w) {
    cmf_state c = *st;
    for (const StateField& f : JPRB_STATE) c.*(f.f) = const_cast<double*>(w(st->*(f.f), field_size(h, f.kind, planes)));
    return c;
}

cmf_state narrow_state(const cmf_hip_ctx* h, const cmf_state* st, Narrow& nw) {
    cmf_state c = *st;
    for (const StateField& f : JPRB_STATE) c.*(f.f) = nw(st->*(f.f), field_size(h, f.kind, (size_t)h->nmem));
    return c;
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

// Device cell order: depth-first post-order of the river tree from each mouth
// (mouths in ascending NSEQ order), upstream branches in ascending NSEQ order. A
// cell's upstream subtree precedes it and a chain of cells is contiguous, so the
// downstream and upstream gathers of a wavefront mostly touch the same or
// adjacent cache lines. Domains stay contiguous: every tree lies in one domain.
bool dfs_order(int n, const std::vector<int32_t>& nxt, const std::vector<int32_t>& up_ptr,
               const std::vector<int32_t>& up_idx, std::vector<int32_t>& perm) {
    perm.assign((size_t)n, -1);
    std::vector<std::pair<int32_t, int32_t>> stack;
    int pos = 0;
    for (int root = 0; root < n; ++root) {
        if (nxt[(size_t)root] >= 0) continue;
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
// A single-precision host (jprb = 4) computed its tables in float: they are
// checked to 1e-5 of their magnitude instead (float rounding of the same
// recurrence; a wrong table differs by far more).
int validate_tables(const cmf_hip_ctx* h, const cmf_map* mp) {
    const int n = h->d.nseqall, nl = h->d.nlfp;
    const double dfr = 1.0 / (double)nl;
    const bool f32 = h->hf32;
    auto differ = [f32](double want, double got) {
        if (!f32) return want != got;
        return std::fabs(want - got) > 1e-5 * std::fmax(std::fabs(want), std::fabs(got));
    };
    if (f32 ? mp->dfrcinc != (double)(float)dfr : mp->dfrcinc != dfr)
        return fail(CMF_EMAP, "DFRCINC=%.17g differs from 1/NLFP", mp->dfrcinc);
    for (int i = 0; i < n; ++i) {
        const double rsm = h->o.lfplain ? (mp->rivlen[i] * mp->rivwth[i]) * mp->rivhgt[i] : (double)1.e18f;
        if (differ(rsm, mp->rivstomax[i])) return fail(CMF_EMAP, "RIVSTOMAX(%d) differs from RIVLEN*RIVWTH*RIVHGT", i + 1);
        if (f32 ? std::fabs(mp->elevtn[i] - mp->rivhgt[i] - mp->rivelv[i]) > 1e-5 * (std::fabs(mp->elevtn[i]) + 1.0)
                : mp->elevtn[i] - mp->rivhgt[i] != mp->rivelv[i])
            return fail(CMF_EMAP, "RIVELV(%d) differs from ELEVTN-RIVHGT", i + 1);
        double stopre = rsm, hpre = 0.0;
        const double dwthinc = (mp->grarea[i] * (1.0 / mp->rivlen[i])) * dfr;
        for (int l = 0; l < nl; ++l) {
            const double hl = mp->fldhgt[(size_t)l * n + i];
            const double dstonow = (mp->rivlen[i] * (mp->rivwth[i] + dwthinc * ((double)(l + 1) - 0.5))) * (hl - hpre);
            const double smax = stopre + dstonow;
            const double grd = (hl - hpre) * (1.0 / dwthinc);
            if (differ(smax, mp->fldstomax[(size_t)l * n + i]) || differ(grd, mp->fldgrd[(size_t)l * n + i]))
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

void drop_graphs(cmf_hip_ctx* h);

// After an upload of DAMINF: whether every cell with I1DAM == 0 holds zero (the
// reference's initial value, CMF_PROG_INIT), which lets the update skip those
// cells' DAMINF and its average
int check_daminf(cmf_hip_ctx* h, const double* daminf) {
    if (!h->d_flag) TRY(dalloc(h, &h->d_flag, 1));
    HIPCHK(hipMemsetAsync(h->d_flag, 0, sizeof(int32_t), h->stream));
    launch_daminf_nonzero(h->M.dam_of_cell, daminf, h->M.ns, h->M.lgm, h->d_flag, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->h_ctl, h->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const int32_t sparse = *reinterpret_cast<int32_t*>(h->h_ctl) == 0;
    if (sparse != h->daminf_sparse) {
        h->daminf_sparse = sparse;
        drop_graphs(h);                                  // captured graphs hold the old arguments
    }
    return 0;
}

template <class R>
int set_state_t(cmf_hip_ctx* h, const cmf_state* st, bool shared = false) {
    DevState<R>& S = state_of<R>(h);
    auto putd = [&](double* dst, const double* src) -> int { return put<double>(h, dst, src, K_CELL, shared); };
    auto putr = [&](R* dst, const double* src, Kind k) -> int { return put<R>(h, dst, src, k, shared); };
    TRY(putd(S.rivsto, st->rivsto));
    TRY(putd(S.fldsto, st->fldsto));
    if (h->o.llevee) TRY(putd(S.levsto, st->levsto));
    if (h->o.ldamout) { TRY(putd(S.damsto, st->damsto)); TRY(putd(S.daminf, st->daminf)); }
    if (h->o.ldamout && st->daminf) TRY(check_daminf(h, S.daminf));
    // D2RIVOUT_PRE lives in the RIVOUT array; a different D2RIVOUT is kept for the
    // RIVVEL of the next substep only (outflw:75 reads the pre-update RIVOUT).
    const double* rpre = st->rivout_pre ? st->rivout_pre : st->rivout;
    TRY(putr(S.rivout, rpre, K_CELL));
    if (st->rivout) {
        TRY(putr((R*)h->vel_buf, st->rivout, K_CELL));
        h->vel_pending = true;
    }
    const double* fpre = st->fldout_pre ? st->fldout_pre : st->fldout;
    TRY(putr(S.fldout, fpre, K_CELL));
    // a D2FLDOUT given beside D2FLDOUT_PRE (a restart zeroes it, physics:254-276
    // makes the two equal again in the next substep) is kept for get_state only
    if (st->fldout && st->fldout_pre) {
        if (!h->fld_buf) {
            R* fb = nullptr;
            TRY(dalloc(h, &fb, (size_t)h->M.ns));
            h->fld_buf = fb;
        }
        TRY(putr((R*)h->fld_buf, st->fldout, K_CELL));
        h->fld_pending = true;
    } else if (fpre) {
        h->fld_pending = false;
    }
    TRY(putr(S.rivdph_pre, st->rivdph_pre, K_CELL));
    TRY(putr(S.fldsto_pre, st->fldsto_pre, K_CELL));
    if (h->o.lpthout) {
        const double* ppre = st->pthflw_pre ? st->pthflw_pre : st->pthflw;
        TRY(putr(S.pthflw, ppre, K_PTHL));
        // a D1PTHFLW given beside D1PTHFLW_PRE is kept for get_state only, as D2FLDOUT
        if (st->pthflw && st->pthflw_pre) {
            if (!h->pth_buf) {
                R* pb = nullptr;
                TRY(dalloc(h, &pb, (size_t)h->M.nsp * (size_t)h->d.npthlev));
                h->pth_buf = pb;
            }
            TRY(putr((R*)h->pth_buf, st->pthflw, K_PTHL));
            h->pth_pending = true;
        } else if (ppre) {
            h->pth_pending = false;
        }
        TRY(putr(S.pthflwsum, st->pthflwsum, K_PTH));
    }
    TRY(putr(S.rivdph, st->rivdph, K_CELL));
    TRY(putr(S.flddph, st->flddph, K_CELL));
    TRY(putr(S.fldfrc, st->fldfrc, K_CELL));
    TRY(putr(S.fldare, st->fldare, K_CELL));
    TRY(putr(S.sfcelv, st->sfcelv, K_CELL));
    TRY(putr(S.storge, st->storge, K_CELL));
    TRY(putr(S.outflw, st->outflw, K_CELL));
    TRY(putr(S.rivvel, st->rivvel, K_CELL));
    if (h->o.llevee) TRY(putr(S.levdph, st->levdph, K_CELL));
    TRY(putd(S.rivinf, st->rivinf));
    TRY(putd(S.fldinf, st->fldinf));
    TRY(putd(S.pthout, st->pthout));
    return 0;
}

template <class R>
int get_state_t(cmf_hip_ctx* h, cmf_state* st) {
    DevState<R>& S = state_of<R>(h);
    auto getd = [&](double* dst, const double* src) -> int { return get<double>(h, dst, src, K_CELL); };
    HIPCHK(hipStreamSynchronize(h->stream));
    TRY(getd(st->rivsto, S.rivsto));
    TRY(getd(st->fldsto, S.fldsto));
    if (h->o.llevee) TRY(getd(st->levsto, S.levsto));
    if (h->o.ldamout) { TRY(getd(st->damsto, S.damsto)); TRY(getd(st->daminf, S.daminf)); }
    TRY(get<R>(h, st->rivout, h->vel_pending ? (const R*)h->vel_buf : S.rivout, K_CELL));
    TRY(get<R>(h, st->rivout_pre, S.rivout, K_CELL));
    TRY(get<R>(h, st->fldout, h->fld_pending ? (const R*)h->fld_buf : S.fldout, K_CELL));
    TRY(get<R>(h, st->fldout_pre, S.fldout, K_CELL));
    TRY(get<R>(h, st->rivdph_pre, S.rivdph_pre, K_CELL));
    TRY(get<R>(h, st->fldsto_pre, S.fldsto_pre, K_CELL));
    if (h->o.lpthout) {
        TRY(get<R>(h, st->pthflw, h->pth_pending ? (const R*)h->pth_buf : S.pthflw, K_PTHL));
        TRY(get<R>(h, st->pthflw_pre, S.pthflw, K_PTHL));
        TRY(get<R>(h, st->pthflwsum, S.pthflwsum, K_PTH));
        TRY(get<R>(h, st->pthflw_avg, S.pthflw_avg, K_PTHL));
        TRY(get<R>(h, st->pthflwsum_avg, S.pthflwsum_avg, K_PTH));
    }
    TRY(get<R>(h, st->rivdph, S.rivdph, K_CELL));
    TRY(get<R>(h, st->flddph, S.flddph, K_CELL));
    TRY(get<R>(h, st->fldfrc, S.fldfrc, K_CELL));
    TRY(get<R>(h, st->fldare, S.fldare, K_CELL));
    TRY(get<R>(h, st->sfcelv, S.sfcelv, K_CELL));
    TRY(get<R>(h, st->storge, S.storge, K_CELL));
    TRY(get<R>(h, st->outflw, S.outflw, K_CELL));
    TRY(get<R>(h, st->rivvel, S.rivvel, K_CELL));
    if (h->o.llevee) TRY(get<R>(h, st->levdph, S.levdph, K_CELL));
    TRY(getd(st->rivinf, S.rivinf));
    TRY(getd(st->fldinf, S.fldinf));
    TRY(getd(st->pthout, S.pthout));
    TRY(get<R>(h, st->rivout_avg, S.rivout_avg, K_CELL));
    TRY(get<R>(h, st->fldout_avg, S.fldout_avg, K_CELL));
    TRY(get<R>(h, st->outflw_avg, S.outflw_avg, K_CELL));
    TRY(get<R>(h, st->rivvel_avg, S.rivvel_avg, K_CELL));
    TRY(get<R>(h, st->pthout_avg, S.pthout_avg, K_CELL));
    TRY(get<R>(h, st->runoff_avg, S.runoff_avg, K_CELL));
    TRY(get<R>(h, st->outflw_max, S.outflw_max, K_CELL));
    TRY(get<R>(h, st->rivdph_max, S.rivdph_max, K_CELL));
    TRY(get<R>(h, st->storge_max, S.storge_max, K_CELL));
    if (h->o.ldamout) TRY(get<R>(h, st->daminf_avg, S.daminf_avg, K_CELL));
    return 0;
}

StepArgs base_args(const cmf_hip_ctx* h) {
    StepArgs A{};
    A.pgrv = h->o.pgrv; A.pmanfld = h->o.pmanfld; A.pdstmth = h->o.pdstmth;
    A.pcadp = h->o.pcadp; A.pminslp = h->o.pminslp;
    A.lfldout = h->o.lfldout; A.lpthout = h->o.lpthout; A.ldamout = h->o.ldamout; A.llevee = h->o.llevee;
    A.budget = h->ninst == 1 && h->nmem == 1 && h->ndom == 1;   // the kernels keep budgets only then
    A.xcd = h->tune.xcd;
    A.lupsinf = h->upsinf_on;
    A.lkine = h->lkine;
    A.lslpmix = h->lslpmix && !h->lkine;                 // physics:66-72: LKINE first
    A.daminf_sparse = h->daminf_sparse;
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

bool valid_bs(int bs) { return bs == 64 || bs == 128 || bs == 256 || bs == 512; }

template <class R>
int coop_capacity(cmf_hip_ctx* h, int bs, int* out) {
    int per_cu = 0, cus = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, substep_kernel<R>(h->M, bs), bs, 0));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    *out = per_cu * cus;
    return 0;
}

// block size of the update phase (its tiles carry the budget partial sums)
int upd_bs(const cmf_hip_ctx* h) { return h->tune.fused ? h->tune.bs_fused : h->tune.bs_update; }

// captured substep graphs hold the device pointers of the state and scratch
// arrays they were recorded with: drop them whenever one of those changes
void drop_graphs(cmf_hip_ctx* h) {
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}

template <class R>
int grow_partials(cmf_hip_ctx* h, DevState<R>& S, int nrun) {
    const size_t need = (size_t)nrun * tiles_cells(h->M, upd_bs(h)) * NGLB;
    if (need > h->partial_cap) {
        // k_update of a cached graph writes into S.partials: the old buffer is freed here
        drop_graphs(h);
        const size_t cap = need + need / 4;
        double* q = nullptr;
        TRY(dalloc(h, &q, cap));
        if (S.partials) {
            (void)hipFree(S.partials);
            for (auto& a : h->allocs) if (a == (void*)S.partials) a = nullptr;
        }
        S.partials = q;
        h->partial_cap = cap;
    }
    if (nrun > h->hist_cap) {
        drop_graphs(h);
        const int cap = nrun + nrun / 4;
        double* p = nullptr;
        TRY(dalloc(h, &p, (size_t)cap * NGLB));
        (void)hipFree(S.glb_hist);
        for (auto& a : h->allocs) if (a == (void*)S.glb_hist) a = nullptr;
        S.glb_hist = p;
        h->hist_cap = cap;
    }
    return 0;
}

// the kernels of substep `it` on the stream (no host synchronisation)
template <class R>
int enqueue_substep(cmf_hip_ctx* h, const DevState<R>& Sk, StepArgs A, size_t* ne) {
    if (h->tune.fused) {
        HIPCHK(launch_substep<R>(h->M, Sk, A, h->tune.bs_fused, h->tune.grid_fused, h->stream));
        if (h->timing) TRY(ev_record(h, (*ne)++));
    } else {
        launch_phase<R>(h->M, Sk, A, 0, h->tune.bs_outflow, h->stream);
        if (h->timing) TRY(ev_record(h, (*ne)++));
        launch_phase<R>(h->M, Sk, A, 1, h->tune.bs_rate, h->stream);
        if (h->timing) TRY(ev_record(h, (*ne)++));
        launch_phase<R>(h->M, Sk, A, 2, h->tune.bs_update, h->stream);
        if (h->timing) TRY(ev_record(h, (*ne)++));
    }
    return 0;
}

// NT substeps with the RIVDPH buffer pair alternating; returns the final pair
template <class R>
int enqueue_substeps(cmf_hip_ctx* h, DevState<R>& S, int32_t ntmax, bool vel_first, size_t* ne) {
    StepArgs A = base_args(h);
    for (int32_t it = 0; it < ntmax; ++it) {
        A.it = it;
        A.prow = it;
        DevState<R> Sk = S;
        Sk.rivdph_next = S.rivdph_pre;
        Sk.rivout_vel = (vel_first && it == 0) ? (const R*)h->vel_buf : S.rivout;
        TRY(enqueue_substep<R>(h, Sk, A, ne));
        std::swap(S.rivdph, S.rivdph_pre);                            // CALC_VARS_PRE for RIVDPH
    }
    HIPCHK(hipGetLastError());
    return 0;
}

template <class R>
int run_substeps_graph(cmf_hip_ctx* h, DevState<R>& S, int32_t ntmax) {
    const int parity = S.rivdph == h->dph_a ? 0 : 1;
    hipGraphExec_t exec = nullptr;
    for (auto& g : h->graphs) if (g.nt == ntmax && g.parity == parity) exec = g.exec;
    if (!exec) {
        DevState<R> S0 = S;
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        size_t ne = 0;
        const int rc = enqueue_substeps<R>(h, S0, ntmax, false, &ne);
        const hipError_t ec = hipStreamEndCapture(h->stream, &graph);
        if (rc != 0) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (ec != hipSuccess) return fail(CMF_EHIP, "substep graph capture: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(CMF_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        if (h->graphs.size() >= 32) {
            (void)hipGraphExecDestroy(h->graphs.front().exec);
            h->graphs.erase(h->graphs.begin());
        }
        h->graphs.push_back({ntmax, parity, exec});
    }
    HIPCHK(hipGraphLaunch(exec, h->stream));
    for (int32_t it = 0; it < ntmax; ++it) std::swap(S.rivdph, S.rivdph_pre);
    return 0;
}

// NT and DT of every instance from the CFL bounds in h->h_cfl (CALC_ADPSTP,
// physics:205-206), written to the host step table h->h_inst; returns the
// largest substep count and whether an instance needs its RIVDPH pair swapped back
int plan_step(cmf_hip_ctx* h, double dt_def, int32_t nt_cap, int32_t* ntmax_out, bool* anyfix_out) {
    const int ni = h->ninst;
    int32_t ntmax = 0;
    for (int k = 0; k < ni; ++k) {
        int32_t nt = 1;
        double dts = dt_def;
        h->last_dtmin[(size_t)k] = 0.0;
        if (h->o.ladpstp) {
            double dt_min;
            std::memcpy(&dt_min, &h->h_cfl[k], sizeof dt_min);
            h->last_dtmin[(size_t)k] = dt_min;
            nt = (int32_t)(dt_def * (1.0 / dt_min) - K_0P01F) + 1;       // physics:205
            dts = dt_def * (double)(1.0f / (float)nt);                   // physics:206
        }
        h->last_nt[(size_t)k] = nt;
        h->last_dt[(size_t)k] = dts;
        const int32_t nrun = (nt_cap > 0 && nt_cap < nt) ? nt_cap : nt;
        double nadd = 0.0;                                               // diag:108
        for (int32_t it = 0; it < nrun; ++it) nadd = nadd + dts;
        h->h_inst[k] = InstStep{dts, h->o.pgrv * dts, nadd, nrun, 0};
        if (nrun > ntmax) ntmax = nrun;
    }
    bool anyfix = false;
    for (int k = 0; k < ni; ++k) {
        h->h_inst[k].swapfix = (ntmax - h->h_inst[k].nt) & 1;
        anyfix = anyfix || h->h_inst[k].swapfix;
    }
    *ntmax_out = ntmax;
    *anyfix_out = anyfix;
    return 0;
}

// One outer step of every instance. nt_cap > 0 caps the substeps (timing).
template <class R>
int advance_t(cmf_hip_ctx* h, double dt_def, int32_t nt_cap, int32_t* nt_out, double* dt_out, double* glb,
              int32_t glb_cap) {
    DevState<R>& S = state_of<R>(h);
    const int ni = h->ninst;
    StepArgs A = base_args(h);
    size_t ne = 0;
    if (h->timing) TRY(ev_record(h, ne++));
    // outer-step start: CMF_PHYSICS_FLDSTG (physics:54) and CALC_ADPSTP (:58)
    launch_stage<R>(h->M, S, A, h->o.ladpstp, ni, dt_def, h->stream);
    HIPCHK(hipGetLastError());
    if (h->timing) TRY(ev_record(h, ne++));
    if (h->o.ladpstp) {
        HIPCHK(hipMemcpyAsync(h->h_cfl, S.cfl, (size_t)ni * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    int32_t ntmax = 0;
    bool anyfix = false;
    TRY(plan_step(h, dt_def, nt_cap, &ntmax, &anyfix));
    HIPCHK(hipMemcpyAsync(h->d_inst, h->h_inst, (size_t)ni * sizeof(InstStep), hipMemcpyHostToDevice, h->stream));
    if (A.budget) TRY(grow_partials<R>(h, S, ntmax));
    const bool vel_first = h->vel_pending;
    if (h->tune.graph && !h->tune.fused && !h->timing && !vel_first) {
        TRY(run_substeps_graph<R>(h, S, ntmax));
    } else {
        TRY(enqueue_substeps<R>(h, S, ntmax, vel_first, &ne));
    }
    if (anyfix) launch_swapfix<R>(h->M, S, h->stream);
    if (A.budget)                                                    // CALC_WATBAL sums, all substeps
        launch_budget(S.partials, tiles_cells(h->M, upd_bs(h)), ntmax, S.glb_hist, h->stream);
    HIPCHK(hipGetLastError());
    h->vel_pending = false;
    h->fld_pending = false;
    h->pth_pending = false;
    if (A.budget && glb && glb_cap > 0) {
        const int32_t m = ntmax < glb_cap ? ntmax : glb_cap;
        HIPCHK(hipMemcpyAsync(glb, S.glb_hist, (size_t)m * NGLB * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    if (h->tune.fused)
        HIPCHK(hipMemcpyAsync(h->h_ctl, S.err, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->tune.fused && *reinterpret_cast<int32_t*>(h->h_ctl) != 0) {
        // clear the flag and the barrier counters (arrival count, generation) so
        // that the next fused launch starts from a consistent barrier
        HIPCHK(hipMemsetAsync(S.err, 0, sizeof(int32_t), h->stream));
        HIPCHK(hipMemsetAsync(S.bar, 0, 2 * sizeof(unsigned), h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return fail(CMF_EHIP, "grid barrier of the fused substep kernel timed out");
    }
    if (h->timing) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        h->t_ms[3] += ms;
        const int per = h->tune.fused ? 1 : 3;
        for (int32_t it = 0; it < ntmax; ++it) {
            for (int k = 0; k < per; ++k) {
                const size_t e1 = 2 + (size_t)per * it + k;
                HIPCHK(hipEventElapsedTime(&ms, h->ev[e1 - 1], h->ev[e1]));
                h->t_ms[h->tune.fused ? 2 : k] += ms;
            }
        }
        h->t_ms[4] += ntmax;
        h->inst_ms.assign((size_t)ni, 0.0);
        for (int k = 0; k < ni; ++k) {
            const size_t e1 = 1 + (size_t)per * h->h_inst[k].nt;
            HIPCHK(hipEventElapsedTime(&ms, h->ev[0], h->ev[e1]));
            h->inst_ms[(size_t)k] = ms;
        }
    }
    if (nt_out) for (int k = 0; k < ni; ++k) nt_out[k] = h->last_nt[(size_t)k];
    if (dt_out) for (int k = 0; k < ni; ++k) dt_out[k] = h->last_dt[(size_t)k];
    return 0;
}

// ---- allocation-free outer step
// Scratch of cmf_hip_physics_advance_noalloc, one copy per GPU: the CFL minimum,
// the CFL bound of every stage block, the budget rows of the substeps and the
// per-tile budget sums of one substep (reduced into its row after each substep,
// so one substep's partial sums suffice). Every value is written before it is
// read within a call, so nothing needs clearing between calls.

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[NA_MAX_DEVICES] = {};      // device address of `scratch` on each GPU
std::string g_scratch_err[NA_MAX_DEVICES];

// once per device, with that device current (called by cmf_hip_create): the
// device address of this GPU's copy of `scratch`. A failure is kept and reported
// by the allocation-free entry point only.
void scratch_setup(int dev) {
    if (dev < 0 || dev >= NA_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev] || !g_scratch_err[dev].empty()) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) g_scratch_err[dev] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    else g_scratch_dev[dev] = static_cast<double*>(p);
}

// advance_t for one instance with the scratch pointers redirected into `sc` (this
// GPU's copy of `scratch`): three phase kernels per substep, the budget of each
// substep reduced right after it, no graph, no event timing, no allocation
template <class R>
int advance_noalloc_t(cmf_hip_ctx* h, double dt_def, int32_t* nt_out, double* glb, int32_t glb_cap, double* sc) {
    DevState<R>& S0 = state_of<R>(h);
    DevState<R> S = S0;
    S.cfl = reinterpret_cast<unsigned long long*>(sc + NA_OFF_CFL);
    S.cfl_part = sc + NA_OFF_PART_CFL;
    S.glb_hist = sc + NA_OFF_HIST;
    S.partials = sc + NA_OFF_TILES;
    StepArgs A = base_args(h);
    launch_stage<R>(h->M, S, A, h->o.ladpstp, 1, dt_def, h->stream);   // physics:54-58
    HIPCHK(hipGetLastError());
    if (h->o.ladpstp) {
        HIPCHK(hipMemcpyAsync(h->h_cfl, S.cfl, sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    int32_t ntmax = 0;
    bool anyfix = false;
    TRY(plan_step(h, dt_def, 0, &ntmax, &anyfix));
    if ((size_t)ntmax > NA_HIST)
        return fail(CMF_EINVAL, "allocation-free advance: NT=%d substeps exceed its maximum of %zu (the stage "
                                "was recomputed as CMF_PHYSICS_FLDSTG does; the storages are unchanged)",
                    (int)ntmax, NA_HIST);
    HIPCHK(hipMemcpyAsync(h->d_inst, h->h_inst, sizeof(InstStep), hipMemcpyHostToDevice, h->stream));
    const int ntc = tiles_cells(h->M, h->tune.bs_update);
    const bool vel_first = h->vel_pending;
    for (int32_t it = 0; it < ntmax; ++it) {
        A.it = it;
        A.prow = 0;
        DevState<R> Sk = S;
        Sk.rivdph_next = S.rivdph_pre;
        Sk.rivout_vel = (vel_first && it == 0) ? (const R*)h->vel_buf : S.rivout;
        launch_phase<R>(h->M, Sk, A, 0, h->tune.bs_outflow, h->stream);
        launch_phase<R>(h->M, Sk, A, 1, h->tune.bs_rate, h->stream);
        launch_phase<R>(h->M, Sk, A, 2, h->tune.bs_update, h->stream);
        launch_budget(S.partials, ntc, 1, S.glb_hist + (size_t)it * NGLB, h->stream);   // CALC_WATBAL sums
        std::swap(S.rivdph, S.rivdph_pre);
    }
    HIPCHK(hipGetLastError());
    S0.rivdph = S.rivdph;
    S0.rivdph_pre = S.rivdph_pre;
    h->vel_pending = false;
    h->fld_pending = false;
    h->pth_pending = false;
    if (glb && glb_cap > 0) {
        const int32_t m = ntmax < glb_cap ? ntmax : glb_cap;
        HIPCHK(hipMemcpyAsync(glb, S.glb_hist, (size_t)m * NGLB * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    *nt_out = h->last_nt[0];
    return 0;
}

// ---- tuner: snapshot / restore of the per-slot state
int snapshot(cmf_hip_ctx* h, bool restore) {
    size_t total = 0;
    for (auto& a : h->state_arrs) total += (a.bytes + 255) & ~(size_t)255;
    if (!h->snapshot) {
        void* q = nullptr;
        HIPCHK(hipMalloc(&q, total));
        h->snapshot = q;
        h->snapshot_bytes = total;
    }
    size_t off = 0;
    for (auto& a : h->state_arrs) {
        char* snap = static_cast<char*>(h->snapshot) + off;
        if (restore) HIPCHK(hipMemcpyAsync(*a.slot, snap, a.bytes, hipMemcpyDeviceToDevice, h->stream));
        else HIPCHK(hipMemcpyAsync(snap, *a.slot, a.bytes, hipMemcpyDeviceToDevice, h->stream));
        off += (a.bytes + 255) & ~(size_t)255;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

template <class R>
int autotune_t(cmf_hip_ctx* h, double dt_def, int32_t nsub, cmf_tune* best, double* report, int32_t report_cap) {
    DevState<R>& S = state_of<R>(h);
    const cmf_tune keep = h->tune;
    const bool keep_timing = h->timing;
    const bool keep_vel = h->vel_pending, keep_fld = h->fld_pending, keep_pth = h->pth_pending;
    // an advance moves only the RIVDPH buffer pair among the state pointers;
    // scratch pointers (budget partials, history) may be reallocated by a trial
    // and must keep their new values
    R* const keep_dph = S.rivdph;
    R* const keep_dph_pre = S.rivdph_pre;
    auto restore_ptrs = [&]() { S.rivdph = keep_dph; S.rivdph_pre = keep_dph_pre; };
    double keep_t[5];
    std::memcpy(keep_t, h->t_ms, sizeof keep_t);
    TRY(snapshot(h, false));
    int nrep = 0;
    auto trial = [&](const cmf_tune& t, double* per_phase, double* total) -> int {
        h->tune = t;
        h->timing = true;
        for (int k = 0; k < 5; ++k) h->t_ms[k] = 0.0;
        int32_t nt;
        double dto;
        // one untimed outer step warms the code, then the timed one
        restore_ptrs(); h->vel_pending = keep_vel; h->fld_pending = keep_fld; h->pth_pending = keep_pth;
        TRY(snapshot(h, true));
        TRY(advance_t<R>(h, dt_def, nsub, nullptr, nullptr, nullptr, 0));
        (void)nt; (void)dto;
        for (int k = 0; k < 5; ++k) h->t_ms[k] = 0.0;
        restore_ptrs(); h->vel_pending = keep_vel; h->fld_pending = keep_fld; h->pth_pending = keep_pth;
        TRY(snapshot(h, true));
        TRY(advance_t<R>(h, dt_def, nsub, nullptr, nullptr, nullptr, 0));
        for (int k = 0; k < 3; ++k) per_phase[k] = h->t_ms[k];
        *total = h->t_ms[0] + h->t_ms[1] + h->t_ms[2];
        if (report && nrep < report_cap) {
            double* r = report + (size_t)nrep * 8;
            r[0] = t.fused; r[1] = t.fused ? t.bs_fused : t.bs_outflow; r[2] = t.xcd; r[3] = t.grid_fused;
            r[4] = per_phase[0]; r[5] = per_phase[1]; r[6] = per_phase[2]; r[7] = *total;
        }
        ++nrep;
        return 0;
    };
    cmf_tune bt = keep;
    double best_total = 1e300;
    for (int xcd = 0; xcd < 2; ++xcd) {
        // per-phase block sizes: every phase is its own kernel, so each takes its best
        double bph[3] = {1e300, 1e300, 1e300};
        int bbs[3] = {256, 256, 256};
        for (int bs : TUNE_BS) {
            cmf_tune t{};
            t.bs_outflow = t.bs_rate = t.bs_update = bs;
            t.bs_fused = 256; t.xcd = xcd; t.graph = 0; t.fused = 0;
            double ph[3], tot;
            TRY(trial(t, ph, &tot));
            for (int k = 0; k < 3; ++k) if (ph[k] < bph[k]) { bph[k] = ph[k]; bbs[k] = bs; }
        }
        const double tot = bph[0] + bph[1] + bph[2];
        if (tot < best_total) {
            best_total = tot;
            bt = cmf_tune{};
            bt.bs_outflow = bbs[0]; bt.bs_rate = bbs[1]; bt.bs_update = bbs[2];
            bt.bs_fused = 256; bt.xcd = xcd; bt.fused = 0; bt.graph = keep.graph;
        }
        for (int bs : {256, 512}) {
            cmf_tune t{};
            t.bs_outflow = t.bs_rate = t.bs_update = 256;
            t.fused = 1; t.bs_fused = bs; t.xcd = xcd;
            int cap = 0;
            TRY(coop_capacity<R>(h, bs, &cap));
            if (cap <= 0) continue;
            t.grid_fused = cap - cap % 8;
            if (t.grid_fused <= 0) t.grid_fused = cap;
            double ph[3], totf;
            TRY(trial(t, ph, &totf));
            if (totf < best_total) { best_total = totf; bt = t; bt.graph = 0; }
        }
    }
    // graph replay of the substep launches changes only the launch overhead, which
    // the kernel events do not see: time the whole capped outer step by wall clock
    // with and without it (the first call of a configuration captures the graph)
    if (!bt.fused) {
        double wall[2] = {0.0, 0.0};
        for (int g = 0; g < 2; ++g) {
            cmf_tune t = bt;
            t.graph = g;
            h->tune = t;
            h->timing = false;
            drop_graphs(h);
            for (int rep = 0; rep < 2; ++rep) {
                restore_ptrs(); h->vel_pending = keep_vel; h->fld_pending = keep_fld; h->pth_pending = keep_pth;
                TRY(snapshot(h, true));
                const auto c0 = std::chrono::steady_clock::now();
                TRY(advance_t<R>(h, dt_def, nsub, nullptr, nullptr, nullptr, 0));
                const auto c1 = std::chrono::steady_clock::now();
                wall[g] = std::chrono::duration<double, std::milli>(c1 - c0).count();
            }
            if (report && nrep < report_cap) {
                double* r = report + (size_t)nrep * 8;
                r[0] = 2 + g; r[1] = bt.bs_update; r[2] = bt.xcd; r[3] = 0;
                r[4] = r[5] = r[6] = 0.0; r[7] = wall[g];
            }
            ++nrep;
        }
        bt.graph = wall[1] < wall[0] ? 1 : 0;
        drop_graphs(h);
    }
    // restore the state and the caller's settings; keep the best configuration
    restore_ptrs();
    TRY(snapshot(h, true));
    h->vel_pending = keep_vel;
    h->fld_pending = keep_fld;
    h->pth_pending = keep_pth;
    h->timing = keep_timing;
    std::memcpy(h->t_ms, keep_t, sizeof keep_t);
    h->tune = bt;
    if (best) *best = bt;
    (void)hipFree(h->snapshot);
    h->snapshot = nullptr;
    return nrep;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* cmf_hip_last_error(void) { return g_err.c_str(); }

int cmf_hip_create_ensemble(const cmf_dims* dims, const cmf_options* opts, const cmf_ensemble* ens,
                            cmf_handle* out) {
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
    if (opts->jprb != 0 && opts->jprb != 4 && opts->jprb != 8)
        return fail(CMF_EINVAL, "jprb=%d: host real kind must be 4 or 8 bytes (0 = 8)", opts->jprb);
    int ndom = 1, nmem = 1;
    std::vector<int32_t> off{0, dims->nseqall};
    if (ens) {
        ndom = ens->ndom;
        nmem = ens->members;
        if (ndom <= 0 || !ens->dom_cells) return fail(CMF_EINVAL, "ensemble needs ndom>0 and dom_cells");
        if (nmem <= 0 || nmem > NMEM_MAX || (nmem & (nmem - 1)))
            return fail(CMF_EINVAL, "members=%d must be a power of two in 1..%d", nmem, NMEM_MAX);
        off.assign(1, 0);
        long long tot = 0;
        for (int d = 0; d < ndom; ++d) {
            if (ens->dom_cells[d] <= 0) return fail(CMF_EINVAL, "domain %d has %d cells", d + 1, ens->dom_cells[d]);
            tot += ens->dom_cells[d];
            off.push_back((int32_t)tot);
        }
        if (tot != dims->nseqall) return fail(CMF_EINVAL, "domain cells sum to %lld, NSEQALL=%d", tot, dims->nseqall);
    }
    if ((long long)dims->nseqall * nmem > 0x7fffffffLL) return fail(CMF_EINVAL, "NSEQALL*members exceeds 2^31");
    cmf_hip_ctx* h = new cmf_hip_ctx();
    h->d = *dims;
    h->o = *opts;
    h->hf32 = opts->jprb == 4;
    h->ndom = ndom; h->nmem = nmem; h->ninst = ndom * nmem;
    while ((1 << h->lgm) < nmem) ++h->lgm;
    h->dom_off = off;
    h->last_nt.assign((size_t)h->ninst, 0);
    h->last_dt.assign((size_t)h->ninst, 0.0);
    h->last_dtmin.assign((size_t)h->ninst, 0.0);
    h->tune = cmf_tune{256, 256, 256, 0, 256, 0, 0, 0};
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(CMF_EHIP, "no HIP device / stream");
    }
    if (hipHostMalloc((void**)&h->h_ctl, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->h_inst, (size_t)h->ninst * sizeof(InstStep), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->h_cfl, (size_t)h->ninst * sizeof(unsigned long long), hipHostMallocDefault) !=
            hipSuccess) {
        cmf_hip_destroy(h);
        return fail(CMF_EHIP, "hipHostMalloc failed");
    }
    scratch_setup(h->device);                       // the allocation-free advance's scratch on this GPU
    const int rc = (opts->precision == CMF_PREC_FP64) ? alloc_state<double>(h, h->S64) : alloc_state<float>(h, h->S32);
    if (rc != 0) {
        std::string keep = g_err;
        cmf_hip_destroy(h);
        g_err = keep;
        return rc;
    }
    *out = h;
    return 0;
}

int cmf_hip_create(const cmf_dims* dims, const cmf_options* opts, cmf_handle* out) {
    return cmf_hip_create_ensemble(dims, opts, nullptr, out);
}

void cmf_hip_destroy(cmf_handle h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    for (void* p : h->allocs) if (p) (void)hipFree(p);
    if (h->snapshot) (void)hipFree(h->snapshot);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    if (h->h_ctl) (void)hipHostFree(h->h_ctl);
    if (h->h_inst) (void)hipHostFree(h->h_inst);
    if (h->h_cfl) (void)hipHostFree(h->h_cfl);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int cmf_hip_set_map(cmf_handle h, const cmf_map* mp_in) {
    if (!h || !mp_in) return fail(CMF_EINVAL, "NULL argument");
    if (h->map_set) return fail(CMF_ESTATE, "map already set for this handle");
    const int n = h->d.nseqall, nl = h->d.nlfp;
    // JPRB tables of a single-precision host, widened for this call
    Widen w(h);
    cmf_map wm = *mp_in;
    {
        const size_t cn = (size_t)n, np = h->o.lpthout ? (size_t)h->d.npthout : 0;
        const size_t nlv = h->o.lpthout ? (size_t)h->d.npthlev : 0, nd = h->o.ldamout ? (size_t)h->d.ndam : 0;
        for (const double** f : {&wm.grarea, &wm.elevtn, &wm.nxtdst, &wm.rivlen, &wm.rivwth, &wm.rivhgt, &wm.rivman,
                                 &wm.rivstomax, &wm.rivelv, &wm.dwnelv, &wm.levhgt, &wm.levfrc, &wm.bashgt, &wm.levdst,
                                 &wm.levbassto, &wm.levtopsto, &wm.levfilsto})
            *f = w(*f, cn);
        for (const double** f : {&wm.fldhgt, &wm.fldstomax, &wm.fldgrd}) *f = w(*f, cn * (size_t)nl);
        wm.pth_dst = w(wm.pth_dst, np);
        wm.pth_elv = w(wm.pth_elv, np * nlv);
        wm.pth_wth = w(wm.pth_wth, np * nlv);
        wm.pth_man = w(wm.pth_man, nlv);
        for (const double** f : {&wm.dam_convol, &wm.dam_emevol, &wm.dam_adjvol, &wm.dam_qa, &wm.dam_qn, &wm.dam_qf,
                                 &wm.dam_norvol, &wm.dam_r_volupa})
            *f = w(*f, nd);
    }
    const cmf_map* mp = &wm;
    if (!mp->i1next || !mp->rivstomax || !mp->rivelv || !mp->dwnelv || !mp->fldhgt || !mp->fldstomax ||
        !mp->fldgrd || !mp->i2mask)
        return fail(CMF_EINVAL, "required map array missing");
    DevMap& M = h->M;
    M.n = n; M.nseqriv = h->d.nseqriv; M.nlfp = nl;
    M.lfplain = h->o.lfplain;
    // a single-precision host holds DFRCINC = 1/NLFP rounded to float; the device
    // tables use the FP64 value of both builds' recurrence (maps:984-995)
    M.dfrcinc = h->hf32 ? 1.0 / (double)nl : mp->dfrcinc;
    M.lgm = h->lgm; M.nmem = h->nmem;
    M.ns = n * h->nmem;
    // domain of every NSEQ cell
    std::vector<int32_t> dom((size_t)n, 0);
    for (int d = 0; d < h->ndom; ++d)
        for (int i = h->dom_off[(size_t)d]; i < h->dom_off[(size_t)d + 1]; ++i) dom[(size_t)i] = d;
    // topology: I1NEXT -> 0-based (negative codes = mouths), upstream CSR in
    // ascending source order
    std::vector<int32_t> nxt(n);
    int nriv = 0;
    for (int i = 0; i < n; ++i) {
        const int32_t v = mp->i1next[i];
        if (v > 0) {
            if (v > n) return fail(CMF_EMAP, "I1NEXT(%d)=%d outside 1..NSEQALL", i + 1, v);
            if (dom[(size_t)v - 1] != dom[(size_t)i])
                return fail(CMF_EMAP, "I1NEXT(%d)=%d leaves the cell's domain", i + 1, v);
            nxt[i] = v - 1;
            ++nriv;
        } else if (v < 0) {
            nxt[i] = v;
        } else {
            return fail(CMF_EMAP, "I1NEXT(%d)=0: river cells need a downstream cell, mouths a negative code", i + 1);
        }
    }
    if (nriv != h->d.nseqriv) return fail(CMF_EMAP, "%d cells have a downstream cell, NSEQRIV=%d", nriv, h->d.nseqriv);
    std::vector<int32_t> owner, item, up_ptr, up_idx;
    for (int i = 0; i < n; ++i) if (nxt[i] >= 0) { owner.push_back(nxt[i]); item.push_back(i); }
    csr(n, owner, up_ptr, up_idx, item);
    // device cell order
    const char* ord = getenv("CMF_HIP_ORDER");
    h->permuted = !(ord && std::string(ord) == "nseq");
    if (h->permuted && !dfs_order(n, nxt, up_ptr, up_idx, h->perm))
        return fail(CMF_EMAP, "river network is not a forest draining to the mouth cells");
    if (!h->permuted) { h->perm.resize((size_t)n); for (int i = 0; i < n; ++i) h->perm[i] = i; }
    h->iperm.assign((size_t)n, 0);
    for (int p = 0; p < n; ++p) h->iperm[(size_t)h->perm[p]] = p;
    if (h->permuted) TRY(put_raw(h, &h->d_perm, h->perm.data(), (size_t)n));
    const std::vector<int32_t>& P = h->perm;
    const std::vector<int32_t>& IP = h->iperm;
    // host index of every device slot
    {
        const int nm = h->nmem;
        h->hx_cell.resize((size_t)n * nm);
        for (int p = 0; p < n; ++p)
            for (int m = 0; m < nm; ++m) h->hx_cell[(size_t)p * nm + m] = (int64_t)m * n + P[(size_t)p];
        const int np = h->o.lpthout ? h->d.npthout : 0, nlv = h->o.lpthout ? h->d.npthlev : 0;
        h->hx_pth.resize((size_t)np * nm);
        h->hx_pthl.resize((size_t)np * nm * nlv);
        for (int q = 0; q < np; ++q)
            for (int m = 0; m < nm; ++m) {
                h->hx_pth[(size_t)q * nm + m] = (int64_t)m * np + q;
                for (int l = 0; l < nlv; ++l)
                    h->hx_pthl[(size_t)l * np * nm + (size_t)q * nm + m] = (int64_t)m * np * nlv + (int64_t)l * np + q;
            }
    }
    // topology in device positions; each upstream list keeps ascending NSEQ order
    std::vector<int32_t> dnext((size_t)n), dptr((size_t)n + 1, 0), didx(up_idx.size()), ddom((size_t)n);
    for (int p = 0; p < n; ++p) {
        const int32_t v = nxt[(size_t)P[p]];
        dnext[(size_t)p] = v >= 0 ? IP[(size_t)v] : v;
        dptr[(size_t)p + 1] = dptr[p] + (up_ptr[(size_t)P[p] + 1] - up_ptr[(size_t)P[p]]);
        ddom[(size_t)p] = dom[(size_t)P[p]];
    }
    for (int p = 0; p < n; ++p) {
        int32_t k = dptr[p];
        for (int32_t q = up_ptr[(size_t)P[p]]; q < up_ptr[(size_t)P[p] + 1]; ++q) didx[(size_t)k++] = IP[(size_t)up_idx[q]];
    }
    int32_t* p32 = nullptr;
    TRY(put_raw(h, &p32, dnext.data(), (size_t)n)); M.next = p32;
    TRY(put_raw(h, &p32, dptr.data(), dptr.size())); M.up_ptr = p32;
    TRY(put_raw(h, &p32, didx.data(), didx.size())); M.up_idx = p32;
    M.cell_dom = nullptr;
    M.ndom_arg = 0;
    if (h->ndom > 1) {
        TRY(put_raw(h, &p32, ddom.data(), (size_t)n)); M.cell_dom = p32;
        // domains stay contiguous in device order (every river tree lies in one
        // domain and the depth-first order visits mouths in NSEQ order)
        bool contiguous = true;
        for (int p = 1; p < n; ++p) contiguous = contiguous && ddom[(size_t)p] >= ddom[(size_t)p - 1];
        if (contiguous && h->ndom <= NDOM_ARG) {
            M.ndom_arg = h->ndom;
            int d = 0;
            M.dom_off_arg[0] = 0;
            for (int p = 0; p < n; ++p)
                while (d < h->ndom - 1 && ddom[(size_t)p] > d) M.dom_off_arg[++d] = p;
            while (d < h->ndom) M.dom_off_arg[++d] = n;
        }
    }
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
        h->dwnelv_dev = pd;
        std::vector<int32_t> m2 = to_dev(h, mp->i2mask, (size_t)n);
        TRY(put_raw(h, &p32, m2.data(), (size_t)n)); M.i2mask = p32;
    }
    // bifurcation paths (maps:531-639)
    M.pth_dom = nullptr;
    M.nsp = 0;
    if (h->o.lpthout) {
        const int np = h->d.npthout, nlv = h->d.npthlev;
        M.npth = np; M.npthlev = nlv;
        M.nsp = np * h->nmem;
        if (!mp->pth_upst || !mp->pth_down || !mp->pth_dst || !mp->pth_elv || !mp->pth_wth || !mp->pth_man)
            return fail(CMF_EINVAL, "LPTHOUT: path tables missing");
        std::vector<int32_t> pu(np), pdn(np), pdom(np, 0), oo, oi, io, ii;
        for (int p = 0; p < np; ++p) {
            pu[p] = mp->pth_upst[p] > 0 ? mp->pth_upst[p] - 1 : -1;
            pdn[p] = mp->pth_down[p] > 0 ? mp->pth_down[p] - 1 : -1;
            if (pu[p] >= n || pdn[p] >= n) return fail(CMF_EMAP, "path %d endpoint outside NSEQALL", p + 1);
            if (pu[p] >= 0 && pdn[p] >= 0 && dom[(size_t)pu[p]] != dom[(size_t)pdn[p]])
                return fail(CMF_EMAP, "path %d joins two domains", p + 1);
            if (pu[p] >= 0) pdom[p] = dom[(size_t)pu[p]];
            else if (pdn[p] >= 0) pdom[p] = dom[(size_t)pdn[p]];
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
        if (h->ndom > 1) { TRY(put_raw(h, &p32, pdom.data(), (size_t)np)); M.pth_dom = p32; }
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
        const bool h22 = h->o.ldamh22 != 0;
        if (!mp->i1dam || !mp->damseq || !mp->damstat || !mp->dam_convol || !mp->dam_emevol || !mp->dam_qn ||
            !mp->dam_qf || (!h22 && (!mp->dam_adjvol || !mp->dam_qa)) ||
            (h22 && (!mp->dam_norvol || !mp->dam_r_volupa)))
            return fail(CMF_EINVAL, "LDAMOUT: dam tables missing (LDAMH22=%d)", (int)h22);
        std::vector<int32_t> doc(n, -1);
        for (int d = 0; d < nd; ++d) {
            if (mp->damstat[d] == -9999) continue;
            const int s = mp->damseq[d] - 1;
            if (s < 0 || s >= n) return fail(CMF_EMAP, "DamSeq(%d)=%d outside NSEQALL", d + 1, mp->damseq[d]);
            doc[(size_t)IP[(size_t)s]] = d;
        }
        std::vector<int32_t> i1d = to_dev(h, mp->i1dam, (size_t)n);
        // the other cells whose DAMINF the dam module sets (I1DAM > 0, damout:444-465)
        for (int p = 0; p < n; ++p)
            if (i1d[(size_t)p] > 0 && doc[(size_t)p] == -1) doc[(size_t)p] = -2;
        TRY(put_raw(h, &p32, i1d.data(), (size_t)n)); M.i1dam = p32;
        TRY(put_raw(h, &p32, doc.data(), (size_t)n)); M.dam_of_cell = p32;
        TRY(put_raw(h, &p32, mp->damstat, (size_t)nd)); M.damstat = p32;
        TRY(put_raw(h, &pd, mp->dam_convol, (size_t)nd)); M.dam_convol = pd;
        TRY(put_raw(h, &pd, mp->dam_emevol, (size_t)nd)); M.dam_emevol = pd;
        const std::vector<double> zero((size_t)nd, 0.0);      // tables of the rule not in use
        TRY(put_raw(h, &pd, h22 && !mp->dam_adjvol ? zero.data() : mp->dam_adjvol, (size_t)nd)); M.dam_adjvol = pd;
        TRY(put_raw(h, &pd, h22 && !mp->dam_qa ? zero.data() : mp->dam_qa, (size_t)nd)); M.dam_qa = pd;
        TRY(put_raw(h, &pd, mp->dam_qn, (size_t)nd)); M.dam_qn = pd;
        TRY(put_raw(h, &pd, mp->dam_qf, (size_t)nd)); M.dam_qf = pd;
        M.ldamh22 = h22 ? 1 : 0;
        TRY(put_raw(h, &pd, h22 ? mp->dam_norvol : zero.data(), (size_t)nd)); M.dam_norvol = pd;
        TRY(put_raw(h, &pd, h22 ? mp->dam_r_volupa : zero.data(), (size_t)nd)); M.dam_r_volupa = pd;
    }
    {
        // first cell of every domain in device order (domains stay contiguous)
        std::vector<int32_t> doff((size_t)h->ndom + 1, 0);
        for (int p = 0; p < n; ++p) doff[(size_t)ddom[(size_t)p] + 1] += 1;
        for (int d = 0; d < h->ndom; ++d) doff[(size_t)d + 1] += doff[(size_t)d];
        for (int p = 1; p < n; ++p)
            if (ddom[(size_t)p] < ddom[(size_t)p - 1]) return fail(CMF_EMAP, "domains are not contiguous in device order");
        TRY(put_raw(h, &h->d_dom_off, doff.data(), doff.size()));
    }
    h->map_set = true;
    return 0;
}

int cmf_hip_instance_storage(cmf_handle h, double* out) {
    if (!h || !out) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (!h->d_ist_out) {
        // chunks per instance: about 2048 blocks in all, each over at least 1024 cells
        int nch = 2048 / h->ninst;
        const int maxch = (h->d.nseqall / h->ndom) / 1024;
        if (nch > maxch) nch = maxch;
        if (nch < 1) nch = 1;
        h->ist_nch = nch;
        TRY(dalloc(h, &h->d_ist_part, (size_t)h->ninst * (size_t)nch));
        TRY(dalloc(h, &h->d_ist_out, (size_t)h->ninst));
    }
    const double *rs, *fs, *ls;
    if (h->o.precision == CMF_PREC_FP64) { rs = h->S64.rivsto; fs = h->S64.fldsto; ls = h->o.llevee ? h->S64.levsto : nullptr; }
    else { rs = h->S32.rivsto; fs = h->S32.fldsto; ls = h->o.llevee ? h->S32.levsto : nullptr; }
    launch_inst_storage(rs, fs, ls, h->d_dom_off, h->lgm, h->ninst, h->ist_nch, h->d_ist_part, h->d_ist_out, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, h->d_ist_out, (size_t)h->ninst * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

namespace {
// state copies with double host arrays (the library's own callers)
int set_state_impl(cmf_hip_ctx* h, const cmf_state* st, bool shared) {
    return h->o.precision == CMF_PREC_FP64 ? set_state_t<double>(h, st, shared) : set_state_t<float>(h, st, shared);
}
int get_state_impl(cmf_hip_ctx* h, cmf_state* st) {
    return h->o.precision == CMF_PREC_FP64 ? get_state_t<double>(h, st) : get_state_t<float>(h, st);
}
}  // namespace

int cmf_hip_set_state(cmf_handle h, const cmf_state* st) {
    if (!h || !st) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    Widen w(h);
    const cmf_state c = widen_state(h, st, (size_t)h->nmem, w);
    return set_state_impl(h, &c, false);
}

int cmf_hip_set_state_shared(cmf_handle h, const cmf_state* st) {
    if (!h || !st) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    Widen w(h);
    const cmf_state c = widen_state(h, st, 1, w);
    return set_state_impl(h, &c, true);
}

int cmf_hip_get_state(cmf_handle h, cmf_state* st) {
    if (!h || !st) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    Narrow nw(h);
    cmf_state c = narrow_state(h, st, nw);
    TRY(get_state_impl(h, &c));
    nw.finish();
    return 0;
}

int cmf_hip_set_runoff(cmf_handle h, const double* runoff_in) {
    if (!h || !runoff_in) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (h->hf32) {
        const float* r = reinterpret_cast<const float*>(runoff_in);
        return h->o.precision == CMF_PREC_FP64 ? put_h<double, float>(h, h->S64.runoff, r, K_CELL)
                                               : put_h<float, float>(h, h->S32.runoff, r, K_CELL);
    }
    return h->o.precision == CMF_PREC_FP64 ? put<double>(h, h->S64.runoff, runoff_in, K_CELL)
                                           : put<float>(h, h->S32.runoff, runoff_in, K_CELL);
}

int cmf_hip_set_runoff_scaled(cmf_handle h, const double* base_in, const double* factor) {
    if (!h || !base_in || !factor) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    const size_t n = (size_t)h->d.nseqall;
    Widen w(h);
    const double* base = w(base_in, n);
    if (!h->d_rbase) {
        TRY(dalloc(h, &h->d_rbase, n));
        TRY(dalloc(h, &h->d_rfac, (size_t)h->ninst));
    }
    std::vector<double> b = to_dev(h, base, n);
    HIPCHK(hipMemcpyAsync(h->d_rbase, b.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->d_rfac, factor, (size_t)h->ninst * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (h->o.precision == CMF_PREC_FP64) launch_scale_runoff<double>(h->M, h->S64, h->d_rbase, h->d_rfac, h->stream);
    else launch_scale_runoff<float>(h->M, h->S32, h->d_rbase, h->d_rfac, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_hip_set_inpmat(cmf_handle h, int32_t inpn, int32_t nxin, int32_t nyin, const int32_t* inpx,
                       const int32_t* inpy, const double* inpa_in, double drofunit, double rmis) {
    if (!h || !inpx || !inpy || !inpa_in) return fail(CMF_EINVAL, "NULL argument");
    Widen w(h);
    const double* inpa = inpn > 0 ? w(inpa_in, (size_t)h->d.nseqall * (size_t)inpn) : inpa_in;
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (h->roff.inpx) return fail(CMF_ESTATE, "the input matrix is already set");
    if (inpn <= 0 || nxin <= 0 || nyin <= 0) return fail(CMF_EINVAL, "bad INPN/NXIN/NYIN (%d, %d, %d)", inpn, nxin, nyin);
    if (!(drofunit != 0.0)) return fail(CMF_EINVAL, "DROFUNIT must be nonzero");
    const size_t n = (size_t)h->d.nseqall, ne = n * (size_t)inpn;
    for (size_t e = 0; e < ne; ++e)
        if (inpx[e] > 0 && inpy[e] < 1)
            return fail(CMF_EINVAL, "input matrix entry %zu of cell %zu: IXIN=%d with IYIN=%d", e / n + 1, e % n + 1,
                        inpx[e], inpy[e]);
    std::vector<int32_t> x = to_dev(h, inpx, n, (size_t)inpn), y = to_dev(h, inpy, n, (size_t)inpn);
    std::vector<double> a = to_dev(h, inpa, n, (size_t)inpn);
    int32_t *dx = nullptr, *dy = nullptr;
    double* da = nullptr;
    TRY(put_raw(h, &dx, x.data(), ne));
    TRY(put_raw(h, &dy, y.data(), ne));
    TRY(put_raw(h, &da, a.data(), ne));
    h->roff = RoffMap{inpn, nxin, nyin, (int)n, dx, dy, da, drofunit, rmis};
    return 0;
}

int cmf_hip_interp_runoff(cmf_handle h, const double* d_grid) {
    if (!h || !d_grid) return fail(CMF_EINVAL, "NULL argument");
    if (!h->roff.inpx) return fail(CMF_ESTATE, "cmf_hip_set_inpmat must come first");
    if (h->o.precision == CMF_PREC_FP64) launch_roff_interp<double>(h->M, h->S64, h->roff, d_grid, h->stream);
    else launch_roff_interp<float>(h->M, h->S32, h->roff, d_grid, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_hip_set_runoff_grid(cmf_handle h, const double* grid_in) {
    if (!h || !grid_in) return fail(CMF_EINVAL, "NULL argument");
    if (!h->roff.inpx) return fail(CMF_ESTATE, "cmf_hip_set_inpmat must come first");
    const size_t ng = (size_t)h->roff.nxin * (size_t)h->roff.nyin;
    Widen w(h);
    const double* grid = w(grid_in, ng);
    if (h->grid_cap < ng) {
        TRY(dalloc(h, &h->d_grid, ng));
        h->grid_cap = ng;
    }
    HIPCHK(hipMemcpyAsync(h->d_grid, grid, ng * sizeof(double), hipMemcpyHostToDevice, h->stream));
    return cmf_hip_interp_runoff(h, h->d_grid);
}

int cmf_hip_set_runoff_records(cmf_handle h, int32_t nrec, const double* grids_in) {
    if (!h || !grids_in) return fail(CMF_EINVAL, "NULL argument");
    if (!h->roff.inpx) return fail(CMF_ESTATE, "cmf_hip_set_inpmat must come first");
    if (nrec <= 0) return fail(CMF_EINVAL, "nrec=%d must be positive", (int)nrec);
    const size_t ng = (size_t)h->roff.nxin * (size_t)h->roff.nyin, nt = ng * (size_t)nrec;
    Widen w(h);
    const double* grids = w(grids_in, nt);
    if (h->d_records) {
        (void)hipFree(h->d_records);
        for (auto& a : h->allocs) if (a == (void*)h->d_records) a = nullptr;
        h->d_records = nullptr;
        h->nrecords = 0;
    }
    TRY(dalloc(h, &h->d_records, nt));
    HIPCHK(hipMemcpyAsync(h->d_records, grids, nt * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));        // the widened copy lives until here
    h->nrecords = nrec;
    return 0;
}

int cmf_hip_use_runoff_record(cmf_handle h, int32_t k) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->d_records) return fail(CMF_ESTATE, "cmf_hip_set_runoff_records must come first");
    if (k < 0 || k >= h->nrecords)
        return fail(CMF_EINVAL, "record %d outside 0..%d", (int)k, (int)h->nrecords - 1);
    const size_t ng = (size_t)h->roff.nxin * (size_t)h->roff.nyin;
    return cmf_hip_interp_runoff(h, h->d_records + (size_t)k * ng);
}

namespace {
int npth_slots(const cmf_hip_ctx* h) { return h->o.lpthout ? h->M.nsp * h->d.npthlev : 0; }

int out_alloc(cmf_hip_ctx* h) {
    if (h->out.rivout) return 0;
    const size_t ns = (size_t)h->M.ns, np = (size_t)npth_slots(h);
    double** cell[] = {&h->out.rivout, &h->out.fldout, &h->out.outflw, &h->out.rivvel, &h->out.pthout,
                       &h->out.runoff, &h->out.daminf, &h->out.outflw_max, &h->out.rivdph_max,
                       &h->out.storge_max};
    for (double** p : cell) TRY(dalloc(h, p, ns));          // dalloc clears
    TRY(dalloc(h, &h->out.pthflw, np));
    h->nadd_out = 0.0;
    return 0;
}
}  // namespace

int cmf_hip_diag_reset_output(cmf_handle h) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    TRY(out_alloc(h));
    launch_out_reset(h->out, h->M.ns, npth_slots(h), h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->nadd_out = 0.0;
    return 0;
}

int cmf_hip_diag_avemax_output(cmf_handle h, double dt) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    TRY(out_alloc(h));
    h->nadd_out = h->nadd_out + dt;                         // diag:277
    if (h->o.precision == CMF_PREC_FP64)
        launch_out_avemax<double>(h->M, h->S64, h->out, dt, h->o.ldamout, npth_slots(h), h->stream);
    else
        launch_out_avemax<float>(h->M, h->S32, h->out, dt, h->o.ldamout, npth_slots(h), h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_hip_diag_getave_output(cmf_handle h, double* nadd_out) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->out.rivout) return fail(CMF_ESTATE, "no output period: call cmf_hip_diag_avemax_output first");
    launch_out_getave(h->out, h->M.ns, npth_slots(h), h->nadd_out, h->o.ldamout, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    if (nadd_out) *nadd_out = h->nadd_out;
    return 0;
}

namespace {
int get_output_impl(cmf_hip_ctx* h, const cmf_output* o) {
    const std::pair<double*, const double*> cell[] = {
        {o->rivout, h->out.rivout}, {o->fldout, h->out.fldout}, {o->outflw, h->out.outflw},
        {o->rivvel, h->out.rivvel}, {o->pthout, h->out.pthout}, {o->runoff, h->out.runoff},
        {o->daminf, h->out.daminf}, {o->outflw_max, h->out.outflw_max}, {o->rivdph_max, h->out.rivdph_max},
        {o->storge_max, h->out.storge_max}};
    for (const auto& c : cell) TRY(get<double>(h, c.first, c.second, K_CELL));
    if (o->pthflw && npth_slots(h) > 0) TRY(get<double>(h, o->pthflw, h->out.pthflw, K_PTHL));
    return 0;
}
}  // namespace

int cmf_hip_get_output(cmf_handle h, const cmf_output* o) {
    if (!h || !o) return fail(CMF_EINVAL, "NULL argument");
    if (!h->out.rivout) return fail(CMF_ESTATE, "no output period: call cmf_hip_diag_avemax_output first");
    Narrow nw(h);
    cmf_output c = *o;
    const size_t ns = (size_t)h->M.ns;
    for (double** f : {&c.rivout, &c.fldout, &c.outflw, &c.rivvel, &c.pthout, &c.runoff, &c.daminf, &c.outflw_max,
                       &c.rivdph_max, &c.storge_max})
        *f = nw(*f, ns);
    c.pthflw = nw(c.pthflw, field_size(h, 1, (size_t)h->nmem));
    TRY(get_output_impl(h, &c));
    nw.finish();
    return 0;
}

int cmf_hip_set_dwnelv(cmf_handle h, const double* dwnelv_in) {
    if (!h || !dwnelv_in) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    const size_t n = (size_t)h->d.nseqall;
    Widen w(h);
    const double* dwnelv = w(dwnelv_in, n);
    std::vector<double> v = to_dev(h, dwnelv, n);
    HIPCHK(hipMemcpyAsync(h->dwnelv_dev, v.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_hip_set_upsinf(cmf_handle h, const double* upsinf_in) {
    if (!h || !upsinf_in) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    const size_t ns = (size_t)h->M.ns;
    Widen w(h);
    const double* upsinf = w(upsinf_in, ns);
    if (h->o.precision == CMF_PREC_FP64) {
        if (!h->S64.upsinf) TRY(dalloc(h, &h->S64.upsinf, ns));
        TRY(put<double>(h, h->S64.upsinf, upsinf, K_CELL));
    } else {
        if (!h->S32.upsinf) TRY(dalloc(h, &h->S32.upsinf, ns));
        TRY(put<float>(h, h->S32.upsinf, upsinf, K_CELL));
    }
    if (!h->upsinf_on) {
        h->upsinf_on = 1;
        drop_graphs(h);                                  // captured graphs hold the old arguments
    }
    return 0;
}

int cmf_hip_set_kinematic(cmf_handle h, int32_t lkine) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (lkine != 0 && lkine != 1) return fail(CMF_EINVAL, "lkine=%d must be 0 or 1", (int)lkine);
    if (h->lkine != lkine) {
        h->lkine = lkine;
        drop_graphs(h);                                  // captured graphs hold the old arguments
    }
    return 0;
}

int cmf_hip_set_slopemix(cmf_handle h, int32_t lslpmix) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (lslpmix != 0 && lslpmix != 1) return fail(CMF_EINVAL, "lslpmix=%d must be 0 or 1", (int)lslpmix);
    if (h->lslpmix != lslpmix) {
        h->lslpmix = lslpmix;
        drop_graphs(h);                                  // captured graphs hold the old arguments
    }
    return 0;
}

int cmf_hip_update_boundary(cmf_handle h, const double* meansl_in, const double* sealev_in) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    const size_t n = (size_t)h->d.nseqall;
    Widen w(h);
    const double* meansl = w(meansl_in, n);
    const double* sealev = w(sealev_in, n);
    const double *dm = nullptr, *dsl = nullptr;
    if (meansl) {
        if (!h->d_meansl) TRY(dalloc(h, &h->d_meansl, n));
        std::vector<double> v = to_dev(h, meansl, n);
        HIPCHK(hipMemcpyAsync(h->d_meansl, v.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        dm = h->d_meansl;
    }
    if (sealev) {
        if (!h->d_sealev) TRY(dalloc(h, &h->d_sealev, n));
        std::vector<double> v = to_dev(h, sealev, n);
        HIPCHK(hipMemcpyAsync(h->d_sealev, v.data(), n * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        dsl = h->d_sealev;
    }
    launch_boundary(h->M.n, h->M.elevtn, dm, dsl, h->dwnelv_dev, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

namespace {
// the restart records of the handle's options, in file order (restart:418-447)
std::vector<const char*> restart_fields(const cmf_hip_ctx* h, bool lstoonly) {
    std::vector<const char*> f{"rivsto", "fldsto"};
    if (!lstoonly) { f.push_back("rivout_pre"); f.push_back("fldout_pre"); f.push_back("rivdph_pre"); f.push_back("fldsto_pre"); }
    if (h->o.ldamout) f.push_back("damsto");
    if (h->o.llevee) f.push_back("levsto");
    return f;
}

double** state_slot(cmf_state* st, const char* k) {
    if (!std::strcmp(k, "rivsto")) return &st->rivsto;
    if (!std::strcmp(k, "fldsto")) return &st->fldsto;
    if (!std::strcmp(k, "rivout_pre")) return &st->rivout_pre;
    if (!std::strcmp(k, "fldout_pre")) return &st->fldout_pre;
    if (!std::strcmp(k, "rivdph_pre")) return &st->rivdph_pre;
    if (!std::strcmp(k, "fldsto_pre")) return &st->fldsto_pre;
    if (!std::strcmp(k, "damsto")) return &st->damsto;
    return &st->levsto;
}

int check_grid(const cmf_hip_ctx* h, int32_t nx, int32_t ny, const int32_t* seqx, const int32_t* seqy) {
    if (nx <= 0 || ny <= 0 || !seqx || !seqy) return fail(CMF_EINVAL, "bad NX/NY or NULL I1SEQX/I1SEQY");
    for (int i = 0; i < h->d.nseqall; ++i)
        if (seqx[i] < 1 || seqx[i] > nx || seqy[i] < 1 || seqy[i] > ny)
            return fail(CMF_EINVAL, "cell %d at (%d, %d) outside the %d x %d grid", i + 1, seqx[i], seqy[i], nx, ny);
    return 0;
}
}  // namespace

int cmf_hip_restart_write(cmf_handle h, const char* file, int32_t nx, int32_t ny, const int32_t* seqx,
                          const int32_t* seqy, int32_t lrestdbl, int32_t lstoonly, double dmis, double rmis) {
    if (!h || !file) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (h->ninst != 1) return fail(CMF_EINVAL, "restart files hold one model: not an ensemble handle");
    TRY(check_grid(h, nx, ny, seqx, seqy));
    const size_t n = (size_t)h->d.nseqall, ng = (size_t)nx * (size_t)ny;
    const std::vector<const char*> fields = restart_fields(h, lstoonly != 0);
    std::vector<std::vector<double>> vec(fields.size(), std::vector<double>(n));
    cmf_state st{};
    for (size_t k = 0; k < fields.size(); ++k) *state_slot(&st, fields[k]) = vec[k].data();
    std::vector<double> pth;
    if (h->o.lpthout) { pth.resize((size_t)h->d.npthout * h->d.npthlev); st.pthflw_pre = pth.data(); }
    TRY(get_state_impl(h, &st));
    FILE* f = std::fopen(file, "wb");
    if (!f) return fail(CMF_EIO, "cannot open %s for writing", file);
    bool ok = true;
    std::vector<double> m8;
    std::vector<float> m4;
    for (size_t k = 0; k < fields.size() && ok; ++k) {                // WRTE_BIN_MAP (restart:484-517)
        if (lrestdbl) {
            m8.assign(ng, dmis);
            for (size_t i = 0; i < n; ++i) m8[(size_t)(seqy[i] - 1) * nx + (seqx[i] - 1)] = vec[k][i];
            ok = std::fwrite(m8.data(), sizeof(double), ng, f) == ng;
        } else {
            m4.assign(ng, (float)rmis);
            for (size_t i = 0; i < n; ++i) m4[(size_t)(seqy[i] - 1) * nx + (seqx[i] - 1)] = (float)vec[k][i];
            ok = std::fwrite(m4.data(), sizeof(float), ng, f) == ng;
        }
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(CMF_EIO, "writing %s failed", file);
    if (h->o.lpthout) {                                               // restart:450-475
        const std::string pf = std::string(file) + ".pth";
        FILE* g = std::fopen(pf.c_str(), "wb");
        if (!g) return fail(CMF_EIO, "cannot open %s for writing", pf.c_str());
        if (lrestdbl) ok = std::fwrite(pth.data(), sizeof(double), pth.size(), g) == pth.size();
        else {
            std::vector<float> p4(pth.begin(), pth.end());
            ok = std::fwrite(p4.data(), sizeof(float), p4.size(), g) == p4.size();
        }
        ok = (std::fclose(g) == 0) && ok;
        if (!ok) return fail(CMF_EIO, "writing %s failed", pf.c_str());
    }
    return 0;
}

int cmf_hip_output_write(cmf_handle h, const char* var, const char* file, int32_t nx, int32_t ny,
                         const int32_t* seqx, const int32_t* seqy, int32_t loutvec, double rmis) {
    if (!h || !var || !file) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (h->ninst != 1) return fail(CMF_EINVAL, "output files hold one model: not an ensemble handle");
    TRY(check_grid(h, nx, ny, seqx, seqy));
    const std::string v(var);
    const size_t n = (size_t)h->d.nseqall;
    std::vector<double> vec;
    // output_mod:489-579: which array a name writes
    const bool avg = v == "rivout" || v == "fldout" || v == "outflw" || v == "totout" || v == "rivvel" ||
                     v == "pthout" || v == "runoff" || v == "rofsfc" || v == "daminf" || v == "maxflw" ||
                     v == "maxdph" || v == "maxsto" || v == "pthflw";
    if ((v == "pthout" || v == "pthflw") && !h->o.lpthout) return 0;
    if ((v == "damsto" || v == "daminf") && !h->o.ldamout) return 0;
    if ((v == "levsto" || v == "levdph") && !h->o.llevee) return 0;
    if (avg) {
        if (!h->out.rivout) return fail(CMF_ESTATE, "%s is an output-period value: no output period yet", var);
        cmf_output o{};
        vec.resize(v == "pthflw" ? (size_t)h->d.npthout * h->d.npthlev : n);
        double* p = vec.data();
        if (v == "rivout") o.rivout = p;
        else if (v == "fldout") o.fldout = p;
        else if (v == "outflw" || v == "totout") o.outflw = p;
        else if (v == "rivvel") o.rivvel = p;
        else if (v == "pthout") o.pthout = p;
        else if (v == "runoff" || v == "rofsfc") o.runoff = p;
        else if (v == "daminf") o.daminf = p;
        else if (v == "maxflw") o.outflw_max = p;
        else if (v == "maxdph") o.rivdph_max = p;
        else if (v == "maxsto") o.storge_max = p;
        else o.pthflw = p;
        TRY(get_output_impl(h, &o));
    } else {
        cmf_state st{};
        vec.resize(n);
        double* p = vec.data();
        if (v == "rivsto") st.rivsto = p;
        else if (v == "fldsto") st.fldsto = p;
        else if (v == "rivdph") st.rivdph = p;
        else if (v == "flddph") st.flddph = p;
        else if (v == "fldfrc") st.fldfrc = p;
        else if (v == "fldare") st.fldare = p;
        else if (v == "sfcelv") st.sfcelv = p;
        else if (v == "storge" || v == "totsto") st.storge = p;
        else if (v == "damsto") st.damsto = p;
        else if (v == "levsto") st.levsto = p;
        else if (v == "levdph") st.levdph = p;
        else return fail(CMF_EINVAL, "unknown output variable %s", var);
        TRY(get_state_impl(h, &st));
    }
    std::vector<float> rec;
    if (v == "pthflw" || loutvec) {                                   // WRTE_OUTPTH / WRTE_OUTVEC
        rec.assign(vec.begin(), vec.end());
    } else {                                                          // vecD2mapR + WRTE_OUTBIN
        rec.assign((size_t)nx * (size_t)ny, (float)rmis);
        for (size_t i = 0; i < n; ++i) rec[(size_t)(seqy[i] - 1) * nx + (seqx[i] - 1)] = (float)vec[i];
    }
    FILE* f = std::fopen(file, "ab");
    if (!f) return fail(CMF_EIO, "cannot open %s for appending", file);
    bool ok = std::fwrite(rec.data(), sizeof(float), rec.size(), f) == rec.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(CMF_EIO, "writing %s failed", file);
    return 0;
}

int cmf_hip_restart_read(cmf_handle h, const char* file, int32_t nx, int32_t ny, const int32_t* seqx,
                         const int32_t* seqy, int32_t lrestdbl, int32_t lstoonly) {
    if (!h || !file) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (h->ninst != 1) return fail(CMF_EINVAL, "restart files hold one model: not an ensemble handle");
    TRY(check_grid(h, nx, ny, seqx, seqy));
    const size_t n = (size_t)h->d.nseqall, ng = (size_t)nx * (size_t)ny;
    const std::vector<const char*> fields = restart_fields(h, lstoonly != 0);
    std::vector<std::vector<double>> vec(fields.size(), std::vector<double>(n));
    FILE* f = std::fopen(file, "rb");
    if (!f) return fail(CMF_EIO, "cannot open %s", file);
    bool ok = true;
    std::vector<double> m8(lrestdbl ? ng : 0);
    std::vector<float> m4(lrestdbl ? 0 : ng);
    for (size_t k = 0; k < fields.size() && ok; ++k) {
        ok = lrestdbl ? std::fread(m8.data(), sizeof(double), ng, f) == ng
                      : std::fread(m4.data(), sizeof(float), ng, f) == ng;
        for (size_t i = 0; ok && i < n; ++i) {
            const size_t g = (size_t)(seqy[i] - 1) * nx + (seqx[i] - 1);
            vec[k][i] = lrestdbl ? m8[g] : (double)m4[g];
        }
    }
    std::fclose(f);
    if (!ok) return fail(CMF_EIO, "%s is shorter than %zu records of %d x %d", file, fields.size(), nx, ny);
    // CMF_RESTART_INIT (restart:100-132): every array the file does not carry is
    // zero (D2RIVOUT and D2FLDOUT always, so the first substep's RIVVEL uses 0;
    // the four t-1 records and D1PTHFLW_PRE with LSTOONLY), then with LSTOONLY
    // D2FLDSTO_PRE = P2FLDSTO
    cmf_state st{};
    for (size_t k = 0; k < fields.size(); ++k) *state_slot(&st, fields[k]) = vec[k].data();
    std::vector<double> zero(n, 0.0), fpre;
    st.rivout = zero.data();
    st.fldout = zero.data();
    if (lstoonly) {
        fpre = vec[1];                                                // P2FLDSTO
        st.rivout_pre = zero.data();
        st.fldout_pre = zero.data();
        st.rivdph_pre = zero.data();
        st.fldsto_pre = fpre.data();
    }
    std::vector<double> pth, pth0;
    if (h->o.lpthout) {                                               // D1PTHFLW = 0 (restart:111-114)
        pth0.assign((size_t)h->d.npthout * h->d.npthlev, 0.0);
        st.pthflw = pth0.data();
    }
    if (h->o.lpthout && lstoonly) {                                   // restart:212-214
        pth.assign((size_t)h->d.npthout * h->d.npthlev, 0.0);
        st.pthflw_pre = pth.data();
    } else if (h->o.lpthout) {
        pth.resize((size_t)h->d.npthout * h->d.npthlev);
        const std::string pf = std::string(file) + ".pth";
        FILE* g = std::fopen(pf.c_str(), "rb");
        if (!g) return fail(CMF_EIO, "cannot open %s", pf.c_str());
        if (lrestdbl) ok = std::fread(pth.data(), sizeof(double), pth.size(), g) == pth.size();
        else {
            std::vector<float> p4(pth.size());
            ok = std::fread(p4.data(), sizeof(float), p4.size(), g) == p4.size();
            for (size_t k = 0; k < p4.size(); ++k) pth[k] = (double)p4[k];
        }
        std::fclose(g);
        if (!ok) return fail(CMF_EIO, "%s is too short", pf.c_str());
        st.pthflw_pre = pth.data();
    }
    return set_state_impl(h, &st, false);
}

int cmf_hip_calc_outpre(cmf_handle h) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    const size_t ns = (size_t)h->M.ns;
    auto run = [&](auto& S) -> int {
        using R = typename std::remove_reference<decltype(*S.rivout)>::type;
        // OUTPRE writes D2RIVOUT_PRE only: keep D2RIVOUT (the RIVVEL input of the
        // next substep) in the separate buffer before the shared array changes
        if (!h->vel_pending) {
            HIPCHK(hipMemcpyAsync(h->vel_buf, S.rivout, ns * sizeof(R), hipMemcpyDeviceToDevice, h->stream));
            h->vel_pending = true;
        }
        if (!h->fld_pending) {                          // and D2FLDOUT, for get_state
            if (!h->fld_buf) {
                R* fb = nullptr;
                TRY(dalloc(h, &fb, ns));
                h->fld_buf = fb;
            }
            HIPCHK(hipMemcpyAsync(h->fld_buf, S.fldout, ns * sizeof(R), hipMemcpyDeviceToDevice, h->stream));
            h->fld_pending = true;
        }
        const size_t npl = (size_t)h->M.nsp * (size_t)h->d.npthlev;
        if (h->o.lpthout && npl > 0 && !h->pth_pending) {        // and D1PTHFLW
            if (!h->pth_buf) {
                R* pb = nullptr;
                TRY(dalloc(h, &pb, npl));
                h->pth_buf = pb;
            }
            HIPCHK(hipMemcpyAsync(h->pth_buf, S.pthflw, npl * sizeof(R), hipMemcpyDeviceToDevice, h->stream));
            h->pth_pending = true;
        }
        launch_outpre<R>(h->M, S, h->o.lfldout, h->o.lpthout, h->o.pmanfld, h->o.pdstmth, h->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    };
    return h->o.precision == CMF_PREC_FP64 ? run(h->S64) : run(h->S32);
}

int cmf_hip_physics_fldstg(cmf_handle h) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    StepArgs A = base_args(h);
    if (h->o.precision == CMF_PREC_FP64) launch_stage<double>(h->M, h->S64, A, 0, h->ninst, 0.0, h->stream);
    else launch_stage<float>(h->M, h->S32, A, 0, h->ninst, 0.0, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_hip_physics_advance_capped(cmf_handle h, double* dt, int32_t nt_cap, int32_t* nt_out, double* glb,
                                   int32_t glb_cap) {
    if (!h || !dt || !nt_out) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (h->ninst != 1) return fail(CMF_EINVAL, "ensemble handle: use cmf_hip_ensemble_advance");
    if (!(*dt > 0.0)) return fail(CMF_EINVAL, "DT must be positive");
    const double dt_def = *dt;
    const int rc = h->o.precision == CMF_PREC_FP64 ? advance_t<double>(h, dt_def, nt_cap, nt_out, nullptr, glb, glb_cap)
                                                   : advance_t<float>(h, dt_def, nt_cap, nt_out, nullptr, glb, glb_cap);
    *dt = dt_def;                                                    // physics:145
    return rc;
}

int cmf_hip_physics_advance(cmf_handle h, double* dt, int32_t* nt_out, double* glb, int32_t glb_cap) {
    return cmf_hip_physics_advance_capped(h, dt, 0, nt_out, glb, glb_cap);
}

int cmf_hip_physics_advance_noalloc(cmf_handle h, double* dt, int32_t* nt_out, double* glb, int32_t glb_cap) {
    if (!h || !dt || !nt_out) return fail(CMF_EINVAL, "NULL argument");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (h->ninst != 1) return fail(CMF_EINVAL, "allocation-free advance: single-model handles only");
    if (!(*dt > 0.0)) return fail(CMF_EINVAL, "DT must be positive");
    if (h->device < 0 || h->device >= NA_MAX_DEVICES) return fail(CMF_EINVAL, "device %d out of range", h->device);
    double* sc = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_err[h->device].empty()) return fail(CMF_EHIP, "%s", g_scratch_err[h->device].c_str());
        sc = g_scratch_dev[h->device];
    }
    if (!sc) return fail(CMF_ESTATE, "allocation-free advance: scratch of device %d not set up", h->device);
    const size_t blocks = ((size_t)h->M.ns + 255) / 256;
    if (blocks > NA_CFL_BLOCKS)
        return fail(CMF_EINVAL, "allocation-free advance: NSEQALL=%d exceeds its maximum of %zu cells", h->M.ns,
                    NA_CFL_BLOCKS * 256);
    const size_t tiles = (size_t)tiles_cells(h->M, h->tune.bs_update);
    if (tiles > NA_TILES)
        return fail(CMF_EINVAL, "allocation-free advance: %zu update tiles (NSEQALL=%d, block %d) exceed its "
                                "maximum of %zu", tiles, h->M.ns, h->tune.bs_update, NA_TILES);
    const double dt_def = *dt;
    const int rc = h->o.precision == CMF_PREC_FP64 ? advance_noalloc_t<double>(h, dt_def, nt_out, glb, glb_cap, sc)
                                                   : advance_noalloc_t<float>(h, dt_def, nt_out, glb, glb_cap, sc);
    *dt = dt_def;                                                    // physics:145
    return rc;
}

int cmf_hip_ensemble_advance(cmf_handle h, double dt, int32_t nt_cap, int32_t* nt_out, double* dt_out) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (!(dt > 0.0)) return fail(CMF_EINVAL, "DT must be positive");
    return h->o.precision == CMF_PREC_FP64 ? advance_t<double>(h, dt, nt_cap, nt_out, dt_out, nullptr, 0)
                                           : advance_t<float>(h, dt, nt_cap, nt_out, dt_out, nullptr, 0);
}

int cmf_hip_get_dt_min(cmf_handle h, double* dt_min) {
    if (!h || !dt_min) return fail(CMF_EINVAL, "NULL argument");
    for (int k = 0; k < h->ninst; ++k) dt_min[k] = h->last_dtmin[(size_t)k];
    return 0;
}

int cmf_hip_instance_times(cmf_handle h, double* ms) {
    if (!h || !ms) return fail(CMF_EINVAL, "NULL argument");
    if (h->inst_ms.size() != (size_t)h->ninst) return fail(CMF_ESTATE, "no timed advance yet");
    for (int k = 0; k < h->ninst; ++k) ms[k] = h->inst_ms[(size_t)k];
    return 0;
}

int cmf_hip_set_tune(cmf_handle h, const cmf_tune* t) {
    if (!h || !t) return fail(CMF_EINVAL, "NULL argument");
    if (!valid_bs(t->bs_outflow) || !valid_bs(t->bs_rate) || !valid_bs(t->bs_update) || !valid_bs(t->bs_fused))
        return fail(CMF_EINVAL, "block sizes must be 64, 128, 256 or 512");
    cmf_tune v = *t;
    if (v.fused) {
        int cap = 0;
        const int rc = h->o.precision == CMF_PREC_FP64 ? coop_capacity<double>(h, v.bs_fused, &cap)
                                                       : coop_capacity<float>(h, v.bs_fused, &cap);
        if (rc != 0) return rc;
        if (v.grid_fused <= 0) v.grid_fused = cap - cap % 8 > 0 ? cap - cap % 8 : cap;
        if (v.grid_fused > cap)
            return fail(CMF_EINVAL, "fused grid %d exceeds the %d co-resident blocks", v.grid_fused, cap);
    }
    if (v.graph != h->tune.graph || v.fused != h->tune.fused || v.bs_outflow != h->tune.bs_outflow ||
        v.bs_rate != h->tune.bs_rate || v.bs_update != h->tune.bs_update || v.bs_fused != h->tune.bs_fused ||
        v.grid_fused != h->tune.grid_fused || v.xcd != h->tune.xcd) {
        drop_graphs(h);
    }
    h->tune = v;
    return 0;
}

int cmf_hip_get_tune(cmf_handle h, cmf_tune* t) {
    if (!h || !t) return fail(CMF_EINVAL, "NULL argument");
    *t = h->tune;
    return 0;
}

int cmf_hip_autotune(cmf_handle h, double dt, int32_t nsub, cmf_tune* best, double* report, int32_t report_cap) {
    if (!h) return fail(CMF_EINVAL, "NULL handle");
    if (!h->map_set) return fail(CMF_ESTATE, "cmf_hip_set_map must come first");
    if (!(dt > 0.0) || nsub <= 0) return fail(CMF_EINVAL, "autotune needs DT>0 and nsub>0");
    drop_graphs(h);
    return h->o.precision == CMF_PREC_FP64 ? autotune_t<double>(h, dt, nsub, best, report, report_cap)
                                           : autotune_t<float>(h, dt, nsub, best, report, report_cap);
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
