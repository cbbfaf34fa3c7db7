// cmf_mlp.hip -- training and inference kernels of the spin-up surrogate, a
// per-cell multilayer perceptron (n_in -> H -> H -> H -> n_out, ReLU) that maps
// catchment attributes and runoff climatology to equilibrium river and
// floodplain storage. No reference counterpart: the reference spins up by
// repeating the forcing year (gosh/test1-glb_15min.sh:311-345).
//
// Arithmetic: BF16 operands on the matrix cores (v_mfma_f32_16x16x32_bf16),
// FP32 accumulation, FP32 master weights and Adam moments. Quantisation points
// (each is a round-to-nearest-even FP32 -> BF16 conversion):
//   forward : X, W_l              -> Z_l = X W_l + b_l (FP32), H_l = bf16(relu(Z_l))
//   output  : Y = H_3 W_4 + b_4 (FP32)
//   backward: G_Y = bf16(dL/dY); dW_l = H_{l-1}^T G_l (FP32); db_l = sum G_l (FP32);
//             G_{l-1} = bf16((G_l W_l^T) * [H_{l-1} > 0])
// The CPU oracle restates the same quantisation points.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cmf_mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

thread_local std::string m_err;
int mfail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    m_err = buf;
    return code;
}
#define MCHK(x)                                                                                  \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return mfail(CMF_MLP_EHIP, "%s: %s", #x, hipGetErrorString(e_));   \
    } while (0)
#define MTRY(x)                 \
    do {                        \
        int r_ = (x);           \
        if (r_ != 0) return r_; \
    } while (0)

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------
// C[M,N] (+)= op(A)[M,K] op(B)[K,N], A and B BF16, C FP32.
// op(A)(m,k) = ta ? A[k*lda+m] : A[m*lda+k];  op(B)(k,n) = tb ? B[n*ldb+k] : B[k*ldb+n].
// Block tile 64x64, K step 32, 256 threads = 4 wavefronts in 2x2, each 32x32 =
// 2x2 MFMA 16x16 tiles. Tiles are staged in LDS with k contiguous for both
// operands, so every lane reads its 8 consecutive k of one row/column with one
// 16-byte LDS load. MFMA 16x16x32 lane layout: lane l supplies A(row l%16,
// k 8*(l/16)..+7) and B(k 8*(l/16)..+7, col l%16) and receives
// D(row 4*(l/16)+r, col l%16), r = 0..3.
constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;   // +8 bf16 pad against bank conflicts

__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                   int ta, const uint16_t* __restrict__ B, int ldb, int tb,
                                                   float* __restrict__ C, int ldc, int accumulate) {
    __shared__ __attribute__((aligned(16))) uint16_t As[TM * LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[TN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += TK) {
        // stage op(A)[m0..+64, k0..+32] and op(B)[k0..+32, n0..+64]: 2048 values each
        for (int e = tid; e < TM * TK; e += 256) {
            int r, k;
            if (ta) { r = e % TM; k = e / TM; }     // consecutive threads walk m (contiguous in memory)
            else { r = e / TK; k = e % TK; }        // consecutive threads walk k
            const int gm = m0 + r, gk = k0 + k;
            uint16_t v = 0;
            if (gm < M && gk < K) v = ta ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[r * LDK + k] = v;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            int c, k;
            if (tb) { c = e / TK; k = e % TK; }     // B[n*ldb+k]: k contiguous
            else { c = e % TN; k = e / TN; }        // B[k*ldb+n]: n contiguous
            const int gn = n0 + c, gk = k0 + k;
            uint16_t v = 0;
            if (gn < N && gk < K) v = tb ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bs[c * LDK + k] = v;
        }
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[(wm + 16 * i + fr) * LDK + fk]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[(wn + 16 * j + fr) * LDK + fk]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + 16 * i + rq + r, gn = n0 + wn + 16 * j + col;
                if (gm < M && gn < N) {
                    float* c = &C[(size_t)gm * ldc + gn];
                    *c = accumulate ? *c + acc[i][j][r] : acc[i][j][r];
                }
            }
}

// FP32 -> BF16 copy
__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// rows gathered by index (training batch) and converted to BF16
__global__ void k_gather_bf16(const float* __restrict__ x, const int64_t* __restrict__ idx, int rows, int cols,
                              uint16_t* __restrict__ y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    y[i] = f2bf(x[(size_t)idx[r] * cols + c]);
}
__global__ void k_gather_f32(const float* __restrict__ x, const int64_t* __restrict__ idx, int rows, int cols,
                             float* __restrict__ y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    y[i] = x[(size_t)idx[r] * cols + c];
}

// H = bf16(relu(Z + b)) (hidden layers); Y = Z + b (output layer, FP32 in place)
__global__ void k_bias_relu_bf16(const float* __restrict__ z, const float* __restrict__ b, int rows, int cols,
                                 uint16_t* __restrict__ h) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const float v = z[i] + b[i % cols];
    h[i] = f2bf(v > 0.f ? v : 0.f);
}
__global__ void k_bias(float* __restrict__ z, const float* __restrict__ b, int rows, int cols) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (long long)rows * cols) z[i] = z[i] + b[i % cols];
}

// weighted mean squared error over rows*cols values, weight w_over where the
// prediction exceeds the target and 1 elsewhere: G = bf16(2 w (Y - T) / (rows*cols)),
// per-block partial sums of w (Y - T)^2 into part[blockIdx.x]
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t, int n,
                                             float w_over, uint16_t* __restrict__ g, float* __restrict__ part) {
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float e2 = 0.f;
    if (i < n) {
        const float d = y[i] - t[i];
        const float w = d > 0.f ? w_over : 1.f;
        e2 = w * (d * d);
        g[i] = f2bf((2.f * w) * d / (float)n);
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// G = bf16(D * [H > 0])
__global__ void k_relu_grad(const float* __restrict__ d, const uint16_t* __restrict__ h, long long n,
                            uint16_t* __restrict__ g) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] = (h[i] & 0x7fffu) != 0 && !(h[i] & 0x8000u) ? f2bf(d[i]) : (uint16_t)0;
}

// db[c] = sum over rows of G[r][c] (BF16 values, FP32 sums in a fixed order):
// chunk sums of COLSUM_CHUNK rows, then the chunk sums in chunk order
constexpr int COLSUM_CHUNK = 128;
__global__ void k_colsum_part(const uint16_t* __restrict__ g, int rows, int cols, float* __restrict__ part) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    const int r0 = blockIdx.y * COLSUM_CHUNK;
    const int r1 = r0 + COLSUM_CHUNK < rows ? r0 + COLSUM_CHUNK : rows;
    float s = 0.f;
    for (int r = r0; r < r1; ++r) s += bf2f(g[(size_t)r * cols + c]);
    part[(size_t)blockIdx.y * cols + c] = s;
}
__global__ void k_colsum_final(const float* __restrict__ part, int nchunk, int cols, float* __restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int k = 0; k < nchunk; ++k) s += part[(size_t)k * cols + c];
    db[c] = s;
}

// Adam with bias correction, FP32 master parameters
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
}

inline unsigned nblk(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ===========================================================================
struct cmf_mlp {
    int n_in = 0, hid = 0, n_out = 0, max_batch = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    long long np = 0;                      // parameter count
    long long off_w[4], off_b[4];          // offsets in the packed parameter vector
    int din[4], dout[4];
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // FP32 master, grads, Adam moments
    uint16_t* wbf = nullptr;               // BF16 copy of the parameters (weights used by the GEMMs)
    // batch buffers
    uint16_t* xb = nullptr;                // [B][n_in]
    uint16_t* hb[3] = {nullptr, nullptr, nullptr};   // [B][H]
    float* z = nullptr;                    // [B][max(H, n_out)] FP32 pre-activation / products
    float* y = nullptr;                    // [B][n_out]
    float* t = nullptr;                    // [B][n_out] targets
    uint16_t* gy = nullptr;                // [B][n_out]
    uint16_t* gh[2] = {nullptr, nullptr};  // [B][H] ping-pong
    float* part = nullptr;                 // loss partial sums
    float* colpart = nullptr;              // bias-gradient chunk sums
    int64_t* idx = nullptr;                // batch row indices
    // resident training set
    float *tx = nullptr, *tt = nullptr;
    long long tn = 0;
    float* h_scratch = nullptr;            // pinned
    float w_over = 1.f;                    // loss weight of over-predictions
};

namespace {

template <class T>
int malloc_t(cmf_mlp* h, T** p, size_t n) {
    void* q = nullptr;
    MCHK(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
    MCHK(hipMemsetAsync(q, 0, (n ? n : 1) * sizeof(T), h->stream));
    MCHK(hipStreamSynchronize(h->stream));
    h->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return 0;
}

void gemm(cmf_mlp* h, int M, int N, int K, const uint16_t* A, int lda, int ta, const uint16_t* B, int ldb, int tb,
          float* C, int ldc, int acc) {
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, h->stream, M, N, K, A, lda, ta, B, ldb, tb, C, ldc, acc);
}

// forward pass of `rows` BF16 input rows in h->xb; output FP32 in h->y
void forward(cmf_mlp* h, int rows) {
    const uint16_t* in = h->xb;
    for (int l = 0; l < 3; ++l) {
        gemm(h, rows, h->dout[l], h->din[l], in, h->din[l], 0, h->wbf + h->off_w[l], h->dout[l], 0, h->z,
             h->dout[l], 0);
        hipLaunchKernelGGL(k_bias_relu_bf16, dim3(nblk((long long)rows * h->dout[l])), dim3(256), 0, h->stream, h->z,
                           h->p + h->off_b[l], rows, h->dout[l], h->hb[l]);
        in = h->hb[l];
    }
    gemm(h, rows, h->n_out, h->hid, in, h->hid, 0, h->wbf + h->off_w[3], h->n_out, 0, h->y, h->n_out, 0);
    hipLaunchKernelGGL(k_bias, dim3(nblk((long long)rows * h->n_out)), dim3(256), 0, h->stream, h->y,
                       h->p + h->off_b[3], rows, h->n_out);
}

void refresh_bf16(cmf_mlp* h) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(h->np)), dim3(256), 0, h->stream, h->p, h->wbf, h->np);
}

// forward + backward of a batch whose BF16 inputs are in xb and FP32 targets in t;
// gradients into h->g, mean squared error into *loss (host)
int grad_batch(cmf_mlp* h, int rows, float* loss) {
    forward(h, rows);
    const int ny = rows * h->n_out;
    const unsigned nb = nblk(ny);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, h->stream, h->y, h->t, ny, h->w_over, h->gy, h->part);
    // output layer: dW4 = H3^T GY, db4 = colsum GY, D = GY W4^T
    const uint16_t* G = h->gy;
    int gcols = h->n_out;
    for (int l = 3; l >= 0; --l) {
        const uint16_t* Hin = l == 0 ? h->xb : h->hb[l - 1];
        gemm(h, h->din[l], h->dout[l], rows, Hin, h->din[l], 1, G, gcols, 0, h->g + h->off_w[l], h->dout[l], 0);
        const int nchunk = (rows + COLSUM_CHUNK - 1) / COLSUM_CHUNK;
        hipLaunchKernelGGL(k_colsum_part, dim3(nblk(gcols, 64), nchunk), dim3(64), 0, h->stream, G, rows, gcols,
                           h->colpart);
        hipLaunchKernelGGL(k_colsum_final, dim3(nblk(gcols, 64)), dim3(64), 0, h->stream, h->colpart, nchunk, gcols,
                           h->g + h->off_b[l]);
        if (l == 0) break;
        gemm(h, rows, h->din[l], h->dout[l], G, gcols, 0, h->wbf + h->off_w[l], h->dout[l], 1, h->z, h->din[l], 0);
        uint16_t* Gn = h->gh[l & 1];
        hipLaunchKernelGGL(k_relu_grad, dim3(nblk((long long)rows * h->din[l])), dim3(256), 0, h->stream, h->z,
                           h->hb[l - 1], (long long)rows * h->din[l], Gn);
        G = Gn;
        gcols = h->din[l];
    }
    if (loss) {
        std::vector<float> part(nb);
        MCHK(hipMemcpyAsync(part.data(), h->part, nb * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        MCHK(hipStreamSynchronize(h->stream));
        double s = 0.0;
        for (unsigned b = 0; b < nb; ++b) s += part[b];
        *loss = (float)(s / ny);
    }
    MCHK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

const char* cmf_mlp_last_error(void) { return m_err.c_str(); }

int cmf_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, int32_t max_batch, cmf_mlp_handle* out) {
    if (!out) return mfail(CMF_MLP_EINVAL, "NULL argument");
    *out = nullptr;
    if (n_in <= 0 || hidden <= 0 || n_out <= 0 || max_batch <= 0)
        return mfail(CMF_MLP_EINVAL, "sizes must be positive");
    cmf_mlp* h = new cmf_mlp();
    h->n_in = n_in; h->hid = hidden; h->n_out = n_out; h->max_batch = max_batch;
    const int din[4] = {n_in, hidden, hidden, hidden}, dout[4] = {hidden, hidden, hidden, n_out};
    long long o = 0;
    for (int l = 0; l < 4; ++l) {
        h->din[l] = din[l]; h->dout[l] = dout[l];
        h->off_w[l] = o; o += (long long)din[l] * dout[l];
        h->off_b[l] = o; o += dout[l];
    }
    h->np = o;
    int rc = 0;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return mfail(CMF_MLP_EHIP, "hipStreamCreate failed");
    }
    const size_t B = (size_t)max_batch, H = (size_t)hidden, wide = (size_t)(hidden > n_out ? hidden : n_out);
    if ((rc = malloc_t(h, &h->p, h->np)) || (rc = malloc_t(h, &h->g, h->np)) || (rc = malloc_t(h, &h->m1, h->np)) ||
        (rc = malloc_t(h, &h->m2, h->np)) || (rc = malloc_t(h, &h->wbf, h->np)) ||
        (rc = malloc_t(h, &h->xb, B * n_in)) || (rc = malloc_t(h, &h->hb[0], B * H)) ||
        (rc = malloc_t(h, &h->hb[1], B * H)) || (rc = malloc_t(h, &h->hb[2], B * H)) ||
        (rc = malloc_t(h, &h->z, B * wide)) || (rc = malloc_t(h, &h->y, B * n_out)) ||
        (rc = malloc_t(h, &h->t, B * n_out)) || (rc = malloc_t(h, &h->gy, B * n_out)) ||
        (rc = malloc_t(h, &h->gh[0], B * H)) || (rc = malloc_t(h, &h->gh[1], B * H)) ||
        (rc = malloc_t(h, &h->part, (B * n_out + 255) / 256)) || (rc = malloc_t(h, &h->idx, B)) ||
        (rc = malloc_t(h, &h->colpart, ((B + COLSUM_CHUNK - 1) / COLSUM_CHUNK) * wide))) {
        std::string keep = m_err;
        cmf_mlp_destroy(h);
        m_err = keep;
        return rc;
    }
    if (hipHostMalloc((void**)&h->h_scratch, 1024 * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        cmf_mlp_destroy(h);
        return mfail(CMF_MLP_EHIP, "hipHostMalloc failed");
    }
    *out = h;
    return 0;
}

void cmf_mlp_destroy(cmf_mlp_handle h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* q : h->allocs) (void)hipFree(q);
    if (h->h_scratch) (void)hipHostFree(h->h_scratch);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int64_t cmf_mlp_param_count(cmf_mlp_handle h) { return h ? h->np : -1; }

int cmf_mlp_set_loss(cmf_mlp_handle h, float w_over) {
    if (!h || !(w_over > 0.f)) return mfail(CMF_MLP_EINVAL, "loss weight must be positive");
    h->w_over = w_over;
    return 0;
}

int cmf_mlp_set_params(cmf_mlp_handle h, const float* params) {
    if (!h || !params) return mfail(CMF_MLP_EINVAL, "NULL argument");
    MCHK(hipMemcpyAsync(h->p, params, h->np * sizeof(float), hipMemcpyHostToDevice, h->stream));
    MCHK(hipMemsetAsync(h->m1, 0, h->np * sizeof(float), h->stream));
    MCHK(hipMemsetAsync(h->m2, 0, h->np * sizeof(float), h->stream));
    refresh_bf16(h);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_mlp_get_params(cmf_mlp_handle h, float* params, float* grads) {
    if (!h) return mfail(CMF_MLP_EINVAL, "NULL handle");
    if (params) MCHK(hipMemcpyAsync(params, h->p, h->np * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    if (grads) MCHK(hipMemcpyAsync(grads, h->g, h->np * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int cmf_mlp_forward(cmf_mlp_handle h, const float* x, int64_t n, float* y) {
    if (!h || !x || !y || n < 0) return mfail(CMF_MLP_EINVAL, "bad argument");
    float* xd = nullptr;
    MCHK(hipMalloc(&xd, (size_t)h->max_batch * h->n_in * sizeof(float)));
    int rc = 0;
    for (int64_t r0 = 0; r0 < n && rc == 0; r0 += h->max_batch) {
        const int rows = (int)((n - r0) < h->max_batch ? (n - r0) : h->max_batch);
        hipError_t e = hipMemcpyAsync(xd, x + (size_t)r0 * h->n_in, (size_t)rows * h->n_in * sizeof(float),
                                      hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) { rc = mfail(CMF_MLP_EHIP, "copy: %s", hipGetErrorString(e)); break; }
        hipLaunchKernelGGL(k_to_bf16, dim3(nblk((long long)rows * h->n_in)), dim3(256), 0, h->stream, xd, h->xb,
                           (long long)rows * h->n_in);
        forward(h, rows);
        e = hipMemcpyAsync(y + (size_t)r0 * h->n_out, h->y, (size_t)rows * h->n_out * sizeof(float),
                           hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = mfail(CMF_MLP_EHIP, "forward: %s", hipGetErrorString(e));
    }
    (void)hipFree(xd);
    return rc;
}

int cmf_mlp_grad(cmf_mlp_handle h, const float* x, const float* t, int32_t rows, float* loss) {
    if (!h || !x || !t || rows <= 0 || rows > h->max_batch) return mfail(CMF_MLP_EINVAL, "bad argument");
    float* xd = nullptr;
    MCHK(hipMalloc(&xd, (size_t)rows * h->n_in * sizeof(float)));
    hipError_t e = hipMemcpyAsync(xd, x, (size_t)rows * h->n_in * sizeof(float), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->t, t, (size_t)rows * h->n_out * sizeof(float), hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) { (void)hipFree(xd); return mfail(CMF_MLP_EHIP, "copy: %s", hipGetErrorString(e)); }
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk((long long)rows * h->n_in)), dim3(256), 0, h->stream, xd, h->xb,
                       (long long)rows * h->n_in);
    const int rc = grad_batch(h, rows, loss);
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(xd);
    return rc;
}

int cmf_mlp_set_data(cmf_mlp_handle h, const float* x, const float* t, int64_t n) {
    if (!h || !x || !t || n <= 0) return mfail(CMF_MLP_EINVAL, "bad argument");
    for (void*& q : h->allocs)
        if (q == (void*)h->tx || q == (void*)h->tt) { (void)hipFree(q); q = nullptr; }
    std::vector<void*> keep;
    for (void* q : h->allocs) if (q) keep.push_back(q);
    h->allocs = keep;
    h->tx = h->tt = nullptr;
    MTRY(malloc_t(h, &h->tx, (size_t)n * h->n_in));
    MTRY(malloc_t(h, &h->tt, (size_t)n * h->n_out));
    MCHK(hipMemcpyAsync(h->tx, x, (size_t)n * h->n_in * sizeof(float), hipMemcpyHostToDevice, h->stream));
    MCHK(hipMemcpyAsync(h->tt, t, (size_t)n * h->n_out * sizeof(float), hipMemcpyHostToDevice, h->stream));
    MCHK(hipStreamSynchronize(h->stream));
    h->tn = n;
    return 0;
}

int cmf_mlp_train(cmf_mlp_handle h, const int64_t* order, int64_t n_order, int32_t batch, float lr, float beta1,
                  float beta2, float eps, int32_t step0, float* loss_out) {
    if (!h || !order || n_order <= 0 || batch <= 0 || batch > h->max_batch)
        return mfail(CMF_MLP_EINVAL, "bad argument");
    if (!h->tx) return mfail(CMF_MLP_ESTATE, "cmf_mlp_set_data must come first");
    for (int64_t k = 0; k < n_order; ++k)
        if (order[k] < 0 || order[k] >= h->tn) return mfail(CMF_MLP_EINVAL, "order[%lld] outside the data", (long long)k);
    double lsum = 0.0;
    int32_t step = step0;
    int64_t nb = 0;
    for (int64_t r0 = 0; r0 + batch <= n_order; r0 += batch) {
        MCHK(hipMemcpyAsync(h->idx, order + r0, (size_t)batch * sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_gather_bf16, dim3(nblk((long long)batch * h->n_in)), dim3(256), 0, h->stream, h->tx,
                           h->idx, batch, h->n_in, h->xb);
        hipLaunchKernelGGL(k_gather_f32, dim3(nblk((long long)batch * h->n_out)), dim3(256), 0, h->stream, h->tt,
                           h->idx, batch, h->n_out, h->t);
        float l = 0.f;
        MTRY(grad_batch(h, batch, &l));
        ++step;
        const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
        hipLaunchKernelGGL(k_adam, dim3(nblk(h->np)), dim3(256), 0, h->stream, h->p, h->g, h->m1, h->m2, h->np, lr,
                           beta1, beta2, eps, c1, c2);
        refresh_bf16(h);
        lsum += l;
        ++nb;
    }
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(h->stream));
    if (loss_out) *loss_out = nb ? (float)(lsum / nb) : 0.f;
    return (int)nb;
}

int cmf_mlp_gemm_test(int32_t M, int32_t N, int32_t K, const uint16_t* A, int32_t lda, int32_t ta, const uint16_t* B,
                      int32_t ldb, int32_t tb, float* C) {
    if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0) return mfail(CMF_MLP_EINVAL, "bad argument");
    const size_t na = (size_t)(ta ? (size_t)K * lda : (size_t)M * lda);
    const size_t nbv = (size_t)(tb ? (size_t)N * ldb : (size_t)K * ldb);
    if ((ta ? lda < M : lda < K) || (tb ? ldb < K : ldb < N)) return mfail(CMF_MLP_EINVAL, "leading dimension too small");
    uint16_t *da = nullptr, *db = nullptr;
    float* dc = nullptr;
    MCHK(hipMalloc(&da, na * 2));
    MCHK(hipMalloc(&db, nbv * 2));
    MCHK(hipMalloc(&dc, (size_t)M * N * 4));
    MCHK(hipMemcpy(da, A, na * 2, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(db, B, nbv * 2, hipMemcpyHostToDevice));
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, 0, M, N, K, da, lda, ta, db, ldb, tb, dc, N, 0);
    MCHK(hipGetLastError());
    MCHK(hipMemcpy(C, dc, (size_t)M * N * 4, hipMemcpyDeviceToHost));
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dc);
    return 0;
}

}  // extern "C"
