// FP32 multilayer perceptron for the learned schedule cost model: training
// and inference on the GPU.
//
// Network: y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3 with x in R^F,
// two hidden layers of H units, one output (predicted log run time).
// Loss: group-centred squared error.  Rows come in groups (the candidate
// schedules of one pattern on one graph); only the differences between
// predictions inside a group matter for choosing a schedule, so the loss is
// sum_g sum_{i in g} ((y_i - t_i) - mean_{j in g}(y_j - t_j))^2 / rows.
// Optimiser: Adam, full batch; one training step (3 forward GEMMs, the loss,
// 5 backward GEMMs, 3 column sums, the update) is captured once into a
// hipGraph and replayed.
//
// Every matrix product goes through one MFMA kernel (v_mfma_f32_16x16x4_f32,
// fp32 in, fp32 accumulate): 64x64 block tiles of 4 waves, each wave 2x2
// tiles of 16x16, K staged through LDS 16 at a time.  Operands are addressed
// through element strides, so transposed products need no copies.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "graphindor_amd.h"

namespace {

constexpr int kBM = 64, kBN = 64, kBK = 16, kThreads = 256;

typedef float float4_t __attribute__((ext_vector_type(4)));

struct GemmArgs {
    const float* A;  // A(m, k) = A[m * sam + k * sak]
    const float* B;  // B(k, n) = B[k * sbk + n * sbn]
    float* C;        // C(m, n) = C[m * ldc + n]
    const float* bias;  // [N] or null
    const float* mask;  // C(m, n) *= (mask[m * ldc + n] > 0) if non-null
    int64_t sam, sak, sbk, sbn, ldc;
    int32_t M, N, K;
    int32_t relu;
};

// C = op(A) op(B) (+ bias) (relu) (* mask).  Grid: ceil(N/64) x ceil(M/64).
__global__ __launch_bounds__(kThreads) void gemm_f32_kernel(GemmArgs g) {
    __shared__ float As[kBK][kBM + 4];
    __shared__ float Bs[kBK][kBN + 4];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int bm = blockIdx.y * kBM, bn = blockIdx.x * kBN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    float4_t acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += kBK) {
        // Stage a 64x16 tile of A and a 16x64 tile of B; out-of-range reads 0.
        for (int idx = tid; idx < kBM * kBK; idx += kThreads) {
            const int m = idx / kBK, k = idx % kBK;
            const int gm = bm + m, gk = k0 + k;
            As[k][m] = (gm < g.M && gk < g.K) ? g.A[(int64_t)gm * g.sam + (int64_t)gk * g.sak] : 0.f;
        }
        for (int idx = tid; idx < kBK * kBN; idx += kThreads) {
            const int k = idx / kBN, n = idx % kBN;
            const int gk = k0 + k, gn = bn + n;
            Bs[k][n] = (gk < g.K && gn < g.N) ? g.B[(int64_t)gk * g.sbk + (int64_t)gn * g.sbn] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            // 16x16x4 operand layout: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].
            const int kr = kk + (lane >> 4), c = lane & 15;
            const float a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
            const float b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // Result layout: lane l, register j holds D[4 * (l / 16) + j][l % 16].
    for (int ti = 0; ti < 2; ++ti)
        for (int tj = 0; tj < 2; ++tj)
            for (int j = 0; j < 4; ++j) {
                const int m = bm + wm + ti * 16 + 4 * (lane >> 4) + j;
                const int n = bn + wn + tj * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[ti][tj][j];
                if (g.bias) v += g.bias[n];
                if (g.relu) v = v > 0.f ? v : 0.f;
                if (g.mask && !(g.mask[(int64_t)m * g.ldc + n] > 0.f)) v = 0.f;
                g.C[(int64_t)m * g.ldc + n] = v;
            }
}

// out[n] = sum_m X[m * ld + n], one thread per column (N <= a few hundred).
__global__ void colsum_kernel(const float* X, int32_t M, int32_t N, int64_t ld, float* out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int32_t m = 0; m < M; ++m) s += X[(int64_t)m * ld + n];
    out[n] = s;
}

// Group-centred residuals: dy[i] = 2 (r_i - mean_g r) / rows with r = y - t;
// loss_g[g] = sum_{i in g} (r_i - mean_g r)^2.  One block per group.
__global__ void centred_loss_kernel(const float* y, const float* t, const int32_t* goff, int32_t rows, float* dy,
                                    float* loss_g) {
    __shared__ float red[kThreads];
    const int gi = blockIdx.x;
    const int32_t b = goff[gi], e = goff[gi + 1];
    float s = 0.f;
    for (int32_t i = b + threadIdx.x; i < e; i += blockDim.x) s += y[i] - t[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const float mean = e > b ? red[0] / (float)(e - b) : 0.f;
    __syncthreads();
    float l = 0.f;
    for (int32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const float c = (y[i] - t[i]) - mean;
        dy[i] = 2.f * c / (float)rows;
        l += c * c;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_g[gi] = red[0];
}

// Adam with decoupled weight decay; the step number lives on the device so a
// replayed graph advances it (step_kernel runs after every update).
__global__ void adam_kernel(float* p, const float* grad, float* m1, float* m2, int64_t n, const int32_t* step, float lr,
                            float wd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int32_t t = *step + 1;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float b = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    const float ah = a / (1.f - powf(b1, (float)t));
    const float bh = b / (1.f - powf(b2, (float)t));
    p[i] -= lr * (ah / (sqrtf(bh) + eps) + wd * p[i]);
}

__global__ void step_kernel(int32_t* step) { *step += 1; }

thread_local std::string t_err;

int mlp_fail(int code, const std::string& msg) {
    t_err = msg;
    return code;
}

#define MLP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return mlp_fail(GM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

struct gm_mlp {
    int32_t F = 0, H = 0;
    int device = 0;
    int64_t nparams = 0;
    // Offsets of W1[F*H], b1[H], W2[H*H], b2[H], W3[H], b3[1] in the parameter buffer.
    int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
    float* d_p = nullptr;   // parameters
    float* d_g = nullptr;   // gradients, same layout
    float* d_m1 = nullptr;  // Adam moments
    float* d_m2 = nullptr;
    int32_t* d_step = nullptr;
    // Activations for the rows of the current buffers.
    int32_t rows_cap = 0, groups_cap = 0;
    float *d_x = nullptr, *d_h1 = nullptr, *d_h2 = nullptr, *d_y = nullptr, *d_t = nullptr, *d_dy = nullptr;
    float *d_dh1 = nullptr, *d_dh2 = nullptr, *d_loss = nullptr;
    int32_t* d_goff = nullptr;
};

namespace {

void free_rows(gm_mlp* m) {
    float** fs[] = {&m->d_x, &m->d_h1, &m->d_h2, &m->d_y, &m->d_t, &m->d_dy, &m->d_dh1, &m->d_dh2, &m->d_loss};
    for (float** f : fs) {
        if (*f) (void)hipFree(*f);
        *f = nullptr;
    }
    if (m->d_goff) (void)hipFree(m->d_goff);
    m->d_goff = nullptr;
    m->rows_cap = m->groups_cap = 0;
}

int ensure_rows(gm_mlp* m, int32_t rows, int32_t groups) {
    if (rows <= m->rows_cap && groups <= m->groups_cap) return GM_OK;
    free_rows(m);
    const int32_t R = std::max(rows, 64), G = std::max(groups, 1);
    const size_t H = (size_t)m->H;
    MLP_TRY(hipMalloc(&m->d_x, sizeof(float) * R * m->F));
    MLP_TRY(hipMalloc(&m->d_h1, sizeof(float) * R * H));
    MLP_TRY(hipMalloc(&m->d_h2, sizeof(float) * R * H));
    MLP_TRY(hipMalloc(&m->d_dh1, sizeof(float) * R * H));
    MLP_TRY(hipMalloc(&m->d_dh2, sizeof(float) * R * H));
    MLP_TRY(hipMalloc(&m->d_y, sizeof(float) * R));
    MLP_TRY(hipMalloc(&m->d_t, sizeof(float) * R));
    MLP_TRY(hipMalloc(&m->d_dy, sizeof(float) * R));
    MLP_TRY(hipMalloc(&m->d_loss, sizeof(float) * G));
    MLP_TRY(hipMalloc(&m->d_goff, sizeof(int32_t) * (G + 1)));
    m->rows_cap = R;
    m->groups_cap = G;
    return GM_OK;
}

void gemm(hipStream_t s, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C,
          int64_t ldc, int32_t M, int32_t N, int32_t K, const float* bias, int relu, const float* mask) {
    GemmArgs g;
    g.A = A;
    g.B = B;
    g.C = C;
    g.bias = bias;
    g.mask = mask;
    g.sam = sam;
    g.sak = sak;
    g.sbk = sbk;
    g.sbn = sbn;
    g.ldc = ldc;
    g.M = M;
    g.N = N;
    g.K = K;
    g.relu = relu;
    dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM);
    hipLaunchKernelGGL(gemm_f32_kernel, grid, dim3(kThreads), 0, s, g);
}

// Forward over `rows` rows of d_x into d_h1, d_h2, d_y.
void forward(gm_mlp* m, hipStream_t s, int32_t rows) {
    const int32_t F = m->F, H = m->H;
    const float* p = m->d_p;
    gemm(s, m->d_x, F, 1, p + m->oW1, H, 1, m->d_h1, H, rows, H, F, p + m->ob1, 1, nullptr);
    gemm(s, m->d_h1, H, 1, p + m->oW2, H, 1, m->d_h2, H, rows, H, H, p + m->ob2, 1, nullptr);
    gemm(s, m->d_h2, H, 1, p + m->oW3, 1, 1, m->d_y, 1, rows, 1, H, p + m->ob3, 0, nullptr);
}

// One full-batch training step (forward, loss, backward, Adam).
void train_step(gm_mlp* m, hipStream_t s, int32_t rows, int32_t groups, float lr, float wd) {
    const int32_t F = m->F, H = m->H;
    const float* p = m->d_p;
    float* g = m->d_g;
    forward(m, s, rows);
    hipLaunchKernelGGL(centred_loss_kernel, dim3(groups), dim3(kThreads), 0, s, m->d_y, m->d_t, m->d_goff, rows, m->d_dy,
                       m->d_loss);
    // Output layer: dW3 = H2^T dy, db3 = sum dy, dH2 = dy W3^T masked by H2 > 0.
    gemm(s, m->d_h2, 1, H, m->d_dy, 1, 1, g + m->oW3, 1, H, 1, rows, nullptr, 0, nullptr);
    hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(64), 0, s, m->d_dy, rows, 1, (int64_t)1, g + m->ob3);
    gemm(s, m->d_dy, 1, 1, p + m->oW3, 1, 1, m->d_dh2, H, rows, H, 1, nullptr, 0, m->d_h2);
    // Second layer: dW2 = H1^T dH2, db2 = colsum dH2, dH1 = dH2 W2^T masked by H1 > 0.
    gemm(s, m->d_h1, 1, H, m->d_dh2, H, 1, g + m->oW2, H, H, H, rows, nullptr, 0, nullptr);
    hipLaunchKernelGGL(colsum_kernel, dim3((H + 255) / 256), dim3(256), 0, s, m->d_dh2, rows, H, (int64_t)H, g + m->ob2);
    gemm(s, m->d_dh2, H, 1, p + m->oW2, 1, H, m->d_dh1, H, rows, H, H, nullptr, 0, m->d_h1);
    // First layer: dW1 = X^T dH1, db1 = colsum dH1.
    gemm(s, m->d_x, 1, F, m->d_dh1, H, 1, g + m->oW1, H, F, H, rows, nullptr, 0, nullptr);
    hipLaunchKernelGGL(colsum_kernel, dim3((H + 255) / 256), dim3(256), 0, s, m->d_dh1, rows, H, (int64_t)H, g + m->ob1);
    const int64_t n = m->nparams;
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m->d_p, g, m->d_m1, m->d_m2, n,
                       m->d_step, lr, wd);
    hipLaunchKernelGGL(step_kernel, dim3(1), dim3(1), 0, s, m->d_step);
}

}  // namespace

extern "C" {

const char* gm_mlp_last_error(void) { return t_err.c_str(); }

int gm_mlp_create(int32_t in_dim, int32_t hidden, uint64_t seed, gm_mlp** out) {
    if (!out || in_dim <= 0 || in_dim > 4096 || hidden <= 0 || hidden > 4096)
        return mlp_fail(GM_ERR_INVALID, "gm_mlp_create: bad dimensions");
    *out = nullptr;
    gm_mlp* m = new gm_mlp();
    m->F = in_dim;
    m->H = hidden;
    (void)hipGetDevice(&m->device);
    const int64_t F = in_dim, H = hidden;
    m->oW1 = 0;
    m->ob1 = m->oW1 + F * H;
    m->oW2 = m->ob1 + H;
    m->ob2 = m->oW2 + H * H;
    m->oW3 = m->ob2 + H;
    m->ob3 = m->oW3 + H;
    m->nparams = m->ob3 + 1;
    // He-uniform weights from SplitMix64, zero biases.
    std::vector<float> p((size_t)m->nparams, 0.f);
    uint64_t st = seed;
    auto fill = [&](int64_t off, int64_t count, int64_t fan_in) {
        const double lim = std::sqrt(6.0 / (double)fan_in);
        for (int64_t i = 0; i < count; ++i) {
            const double u = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[(size_t)(off + i)] = (float)((2.0 * u - 1.0) * lim);
        }
    };
    fill(m->oW1, F * H, F);
    fill(m->oW2, H * H, H);
    fill(m->oW3, H, H);
    auto bail = [&](int code) {
        gm_mlp_destroy(m);
        return code;
    };
    const size_t bytes = sizeof(float) * (size_t)m->nparams;
    if (hipMalloc(&m->d_p, bytes) != hipSuccess || hipMalloc(&m->d_g, bytes) != hipSuccess ||
        hipMalloc(&m->d_m1, bytes) != hipSuccess || hipMalloc(&m->d_m2, bytes) != hipSuccess ||
        hipMalloc(&m->d_step, sizeof(int32_t)) != hipSuccess)
        return bail(mlp_fail(GM_ERR_OOM, "gm_mlp_create: device allocation failed"));
    if (hipMemcpy(m->d_p, p.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->d_g, 0, bytes) != hipSuccess || hipMemset(m->d_m1, 0, bytes) != hipSuccess ||
        hipMemset(m->d_m2, 0, bytes) != hipSuccess || hipMemset(m->d_step, 0, sizeof(int32_t)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return bail(mlp_fail(GM_ERR_HIP, "gm_mlp_create: initialisation failed"));
    *out = m;
    return GM_OK;
}

void gm_mlp_destroy(gm_mlp* m) {
    if (!m) return;
    free_rows(m);
    if (m->d_p) (void)hipFree(m->d_p);
    if (m->d_g) (void)hipFree(m->d_g);
    if (m->d_m1) (void)hipFree(m->d_m1);
    if (m->d_m2) (void)hipFree(m->d_m2);
    if (m->d_step) (void)hipFree(m->d_step);
    delete m;
}

int64_t gm_mlp_param_count(const gm_mlp* m) { return m ? m->nparams : 0; }

int gm_mlp_get_params(const gm_mlp* m, float* out) {
    if (!m || !out) return mlp_fail(GM_ERR_INVALID, "gm_mlp_get_params: null argument");
    MLP_TRY(hipMemcpy(out, m->d_p, sizeof(float) * (size_t)m->nparams, hipMemcpyDeviceToHost));
    return GM_OK;
}

int gm_mlp_get_grads(const gm_mlp* m, float* out) {
    if (!m || !out) return mlp_fail(GM_ERR_INVALID, "gm_mlp_get_grads: null argument");
    MLP_TRY(hipMemcpy(out, m->d_g, sizeof(float) * (size_t)m->nparams, hipMemcpyDeviceToHost));
    return GM_OK;
}

int gm_mlp_set_params(gm_mlp* m, const float* in) {
    if (!m || !in) return mlp_fail(GM_ERR_INVALID, "gm_mlp_set_params: null argument");
    const size_t bytes = sizeof(float) * (size_t)m->nparams;
    MLP_TRY(hipMemcpy(m->d_p, in, bytes, hipMemcpyHostToDevice));
    // New parameters start a new optimisation.
    MLP_TRY(hipMemset(m->d_m1, 0, bytes));
    MLP_TRY(hipMemset(m->d_m2, 0, bytes));
    MLP_TRY(hipMemset(m->d_step, 0, sizeof(int32_t)));
    MLP_TRY(hipDeviceSynchronize());
    return GM_OK;
}

int gm_mlp_predict(gm_mlp* m, const float* x, int32_t rows, float* y) {
    if (!m || !x || !y || rows < 0) return mlp_fail(GM_ERR_INVALID, "gm_mlp_predict: bad argument");
    if (rows == 0) return GM_OK;
    int rc = ensure_rows(m, rows, 1);
    if (rc != GM_OK) return rc;
    MLP_TRY(hipMemcpy(m->d_x, x, sizeof(float) * (size_t)rows * m->F, hipMemcpyHostToDevice));
    forward(m, nullptr, rows);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipMemcpy(y, m->d_y, sizeof(float) * (size_t)rows, hipMemcpyDeviceToHost));
    return GM_OK;
}

int gm_mlp_train(gm_mlp* m, const float* x, const float* t, const int32_t* group_off, int32_t groups, int32_t steps,
                 float lr, float weight_decay, float* final_loss) {
    if (!m || !x || !t || !group_off || groups <= 0 || steps < 0)
        return mlp_fail(GM_ERR_INVALID, "gm_mlp_train: bad argument");
    if (group_off[0] != 0) return mlp_fail(GM_ERR_INVALID, "gm_mlp_train: group_off[0] must be 0");
    for (int32_t i = 0; i < groups; ++i)
        if (group_off[i + 1] < group_off[i]) return mlp_fail(GM_ERR_INVALID, "gm_mlp_train: group offsets decrease");
    const int32_t rows = group_off[groups];
    if (rows <= 0) return mlp_fail(GM_ERR_INVALID, "gm_mlp_train: no rows");
    int rc = ensure_rows(m, rows, groups);
    if (rc != GM_OK) return rc;
    MLP_TRY(hipMemcpy(m->d_x, x, sizeof(float) * (size_t)rows * m->F, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->d_t, t, sizeof(float) * (size_t)rows, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->d_goff, group_off, sizeof(int32_t) * (size_t)(groups + 1), hipMemcpyHostToDevice));
    hipStream_t s = nullptr;
    MLP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        train_step(m, s, rows, groups, lr, weight_decay);
        hipError_t e2 = hipStreamEndCapture(s, &graph);
        e = e2 != hipSuccess ? e2 : hipGetLastError();
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int32_t i = 0; i < steps && e == hipSuccess; ++i) e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return mlp_fail(GM_ERR_HIP, std::string("gm_mlp_train: ") + hipGetErrorString(e));
    if (final_loss) {
        // Loss of the parameters the last step started from.
        std::vector<float> lg((size_t)groups);
        MLP_TRY(hipMemcpy(lg.data(), m->d_loss, sizeof(float) * (size_t)groups, hipMemcpyDeviceToHost));
        double sum = 0;
        for (float v : lg) sum += v;
        *final_loss = steps > 0 ? (float)(sum / rows) : 0.f;
    }
    return GM_OK;
}

}  // extern "C"
