// mlp.hip -- the learned warm start's network on MFMA: a BF16 multilayer
// perceptron (in -> hidden x layers -> out, ReLU) trained and evaluated on
// the GPU by hand-written kernels.
//
//   k_gemm_bf16   C(M x N, fp32) = op(A)(M x K) op(B)(K x N), BF16 operands,
//                 FP32 accumulation on v_mfma_f32_32x32x8bf16_1k; one wave per
//                 32 x 32 output tile, operand strips staged in LDS 64 k at a
//                 time; strided operand access gives both transposes the
//                 backward pass needs
//   k_bias_act    + bias, ReLU, BF16 copy of the activation (forward)
//   k_mse_grad    dY = 2 (Y - T) / (batch * out), loss partials
//   k_relu_back   dZ = dA * (Z > 0), BF16 copy
//   k_colsum      bias gradient (column sums of dZ, 16 row groups per column)
//   k_adam        Adam step on the FP32 master weights + BF16 refresh
//
// Shapes are multiples of 32 (rows, columns) and 8 (reduction); the host
// pads the batch.  BF16 values travel as raw 16-bit patterns; conversion
// from FP32 rounds to nearest even.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mcpdlp.h"
#include "../../include/mcpdlp_mlp.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// One wave per 32 x 32 tile of C.  Operand element (i, k) of A at
// A[i*sai + k*sak], (k, j) of B at B[k*sbk + j*sbj].  The wave stages its
// 32 x KC strip of A and KC x 32 strip of B in LDS (both k-contiguous, B
// transposed; rows padded against bank conflicts), every load of a chunk in
// flight at once -- one memory latency per KC = 64 instead of one per MFMA
// step -- with lanes running along whichever index of an operand has unit
// stride, so the global loads coalesce for every transpose.  MFMA 32x32x8
// layout: lane l supplies A(l%32, 4*(l/32)+u) and B(4*(l/32)+u, l%32),
// u = 0..3, and holds C(8*(r/4) + 4*(l/32) + r%4, l%32) in accumulator
// register r; the k steps run in increasing order.  Waves past the last row
// tile stage a valid tile and store nothing, so every wave reaches the
// workgroup barriers.
constexpr int kGemmKC = 64;               // k per staged chunk
constexpr int kGemmRow = kGemmKC + 4;     // padded LDS row, BF16 elements

__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A,
                                                   int64_t sai, int64_t sak, const uint16_t* __restrict__ B,
                                                   int64_t sbk, int64_t sbj, float* __restrict__ C, int ldc,
                                                   int accumulate) {
    __shared__ uint16_t As[4][32 * kGemmRow];
    __shared__ uint16_t Bs[4][32 * kGemmRow];  // B strip transposed: [column][k]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int tn = N >> 5;
    const int ti = wave / tn, tj = wave % tn;
    const bool live = ti < (M >> 5);
    const int i0 = (live ? ti : 0) * 32, j0 = tj * 32;
    uint16_t* as = As[wv];
    uint16_t* bs = Bs[wv];
    const bool a_rows_unit = sai == 1 && sak != 1;  // A transposed: unit stride along rows
    const bool b_cols_unit = sbj == 1;              // B row-major: unit stride along columns
    const int kq = 4 * (lane >> 5), rl = lane & 31;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += kGemmKC) {
        const int kc = min(kGemmKC, K - k0);  // a multiple of 8
        // element e of a 32 x 64 strip: (e % 32, e / 32) along the unit stride
        // of rows/columns, else (e / 64, e % 64); k >= kc is not loaded.  All 32
        // loads of each operand are issued before the first LDS store.
        uint16_t ra[2048 / 64], rb[2048 / 64];
#pragma unroll
        for (int q = 0; q < 2048 / 64; ++q) {
            const int e = lane + 64 * q;
            const int ar = a_rows_unit ? (e & 31) : (e >> 6), ak = a_rows_unit ? (e >> 5) : (e & 63);
            const int bc = b_cols_unit ? (e & 31) : (e >> 6), bk = b_cols_unit ? (e >> 5) : (e & 63);
            ra[q] = ak < kc ? A[(int64_t)(i0 + ar) * sai + (int64_t)(k0 + ak) * sak] : (uint16_t)0;
            rb[q] = bk < kc ? B[(int64_t)(k0 + bk) * sbk + (int64_t)(j0 + bc) * sbj] : (uint16_t)0;
        }
#pragma unroll
        for (int q = 0; q < 2048 / 64; ++q) {
            const int e = lane + 64 * q;
            const int ar = a_rows_unit ? (e & 31) : (e >> 6), ak = a_rows_unit ? (e >> 5) : (e & 63);
            const int bc = b_cols_unit ? (e & 31) : (e >> 6), bk = b_cols_unit ? (e >> 5) : (e & 63);
            as[ar * kGemmRow + ak] = ra[q];
            bs[bc * kGemmRow + bk] = rb[q];
        }
        __syncthreads();
        for (int ks = 0; ks < kc; ks += 8) {
            bf16x4 a, b;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = (short)as[rl * kGemmRow + ks + kq + u];
                b[u] = (short)bs[rl * kGemmRow + ks + kq + u];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    if (!live) return;
    const int j = j0 + rl;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = i0 + 8 * (r >> 2) + kq + (r & 3);
        float* c = C + (int64_t)row * ldc + j;
        *c = accumulate ? *c + acc[r] : acc[r];
    }
}

// Z = H + bias (in place, fp32); A = act(Z) as BF16 (relu != 0: ReLU).
__global__ void k_bias_act(int rows, int cols, float* Z, const float* bias, uint16_t* Abf, int relu) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)rows * cols) return;
    const float z = Z[k] + bias[k % cols];
    Z[k] = z;
    if (Abf) Abf[k] = f2bf(relu ? fmaxf(z, 0.0f) : z);
}

// dY = 2 (Y - T) / (valid * cols) on rows < valid (padding rows give 0);
// part[block] = sum of squared errors of this block's elements.
__global__ void k_mse_grad(int rows, int valid, int cols, const float* Y, const float* T, float* dY,
                           uint16_t* dYbf, float* part) {
    __shared__ float red[256];
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.0f;
    if (k < (int64_t)rows * cols) {
        const bool ok = (k / cols) < valid;
        const float e = ok ? Y[k] - T[k] : 0.0f;
        const float g = 2.0f * e / (float)((int64_t)valid * cols);
        dY[k] = g;
        dYbf[k] = f2bf(g);
        e2 = e * e;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dZ = dA * (Z > 0) (fp32 in place) and its BF16 copy.
__global__ void k_relu_back(int64_t n, float* dA, const float* Z, uint16_t* dZbf) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float g = Z[k] > 0.0f ? dA[k] : 0.0f;
    dA[k] = g;
    dZbf[k] = f2bf(g);
}

// db[j] = sum_i dZ[i, j]: a workgroup takes 64 columns with 16 row groups;
// a thread sums rows g, g + 16, ... of its column in order, and the 16
// partials are added in group order (a fixed order: deterministic).
constexpr int kColsumGroups = 16;
__global__ __launch_bounds__(64 * kColsumGroups) void k_colsum(int rows, int cols, const float* __restrict__ dZ,
                                                               float* __restrict__ db) {
    __shared__ float part[kColsumGroups][64];
    const int tx = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + tx;
    float s = 0.0f;
    if (j < cols)
        for (int i = g; i < rows; i += kColsumGroups) s += dZ[(int64_t)i * cols + j];
    part[g][tx] = s;
    __syncthreads();
    if (g == 0 && j < cols) {
        float t = 0.0f;
        for (int q = 0; q < kColsumGroups; ++q) t += part[q][tx];
        db[j] = t;
    }
}

// Adam (Kingma & Ba 2015) on fp32 master params; refreshes the BF16 weights.
__global__ void k_adam(int64_t n, float* p, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                       float eps, float c1, float c2, uint16_t* pbf) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.0f - b1) * gk;
    const float v = b2 * m2[k] + (1.0f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = v;
    const float upd = lr * (a / c1) / (sqrtf(v / c2) + eps);
    p[k] -= upd;
    if (pbf) pbf[k] = f2bf(p[k]);
}

__global__ void k_to_bf16(int64_t n, const float* src, uint16_t* dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = f2bf(src[k]);
}

// BF16 copy of the first `valid` rows of x (rows x cols, row-major) with
// zero padding rows up to `rows`.
__global__ void k_rows_to_bf16(int valid, int rows, int cols, const float* x, uint16_t* dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)rows * cols) return;
    dst[k] = (k / cols) < valid ? f2bf(x[k]) : (uint16_t)0;
}

__global__ void k_copy_f32(int64_t n, const float* src, float* dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

// Device scratch of mcp_mlp_forward_noalloc: the padded BF16 input, the
// pre-activation of the current layer (FP32) and the BF16 activations, ping-
// ponged between layers.  One copy per device, held for the life of the
// process; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device.
constexpr size_t kNoAllocScratchBytes = 16u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

thread_local std::string g_mlp_err;

struct MlpError : std::runtime_error {
    int code;
    MlpError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MCHK(expr)                                                                                   \
    do {                                                                                             \
        const hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                        \
            throw MlpError(e_ == hipErrorOutOfMemory ? MC_ERR_OUT_OF_MEMORY : MC_ERR_DEVICE,        \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                       \
    } while (0)

template <class F>
int mguard(F&& f) {
    try {
        f();
        return MC_OK;
    } catch (const MlpError& e) {
        g_mlp_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_mlp_err = e.what();
        return MC_ERR_INTERNAL;
    }
}

int blocks(int64_t n) { return (int)((n + 255) / 256); }

// Device address of `scratch`, looked up once per device with that device
// current (each device holds its own copy at its own address).
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

void* scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        MCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw MlpError(MC_ERR_DEVICE, "no device address for the forward-pass scratch");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

struct mcp_mlp {
    int device = 0;
    hipStream_t s = nullptr;
    std::vector<int> dims;  // in, h, ..., h, out
    int cap = 0;            // batch capacity (multiple of 32)
    std::vector<int64_t> woff, boff;  // offsets into the flat parameter vector
    int64_t nparam = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // fp32 params / grads / Adam moments
    uint16_t* wbf = nullptr;                                          // BF16 weights (same offsets)
    std::vector<float*> Z;        // pre-activations per layer [cap x dims[l+1]]
    std::vector<uint16_t*> Abf;   // BF16 activations; Abf[0] = input
    float *dA = nullptr, *dB = nullptr, *T = nullptr, *part = nullptr;
    uint16_t *dbf = nullptr;
    int64_t step = 0;
    std::vector<void*> owned;

    template <class T_>
    T_* alloc(int64_t count) {
        void* q = nullptr;
        MCHK(hipMalloc(&q, std::max<int64_t>(count, 1) * sizeof(T_)));
        owned.push_back(q);
        return static_cast<T_*>(q);
    }
    ~mcp_mlp() {
        if (s) (void)hipStreamSynchronize(s);
        for (void* q : owned) (void)hipFree(q);
        if (s) (void)hipStreamDestroy(s);
    }
    int layers() const { return (int)dims.size() - 1; }
    int maxw() const {
        int w = 0;
        for (int d : dims) w = std::max(w, d);
        return w;
    }

    void gemm(int M, int N, int K, const uint16_t* A, int64_t sai, int64_t sak, const uint16_t* B, int64_t sbk,
              int64_t sbj, float* C, int ldc, bool acc) {
        const int waves = (M / 32) * (N / 32);
        hipLaunchKernelGGL(k_gemm_bf16, dim3((waves + 3) / 4), dim3(256), 0, s, M, N, K, A, sai, sak, B, sbk, sbj, C,
                           ldc, acc ? 1 : 0);
    }

    // forward on `rows` padded rows (input already in Abf[0])
    void forward(int rows) {
        for (int l = 0; l < layers(); ++l) {
            const int K = dims[l], N = dims[l + 1];
            // Z = A_l (rows x K) * W_l (K x N); W stored row-major K x N
            gemm(rows, N, K, Abf[l], K, 1, wbf + woff[l], N, 1, Z[l], N, false);
            const bool last = l == layers() - 1;
            hipLaunchKernelGGL(k_bias_act, dim3(blocks((int64_t)rows * N)), dim3(256), 0, s, rows, N, Z[l],
                               p + boff[l], last ? nullptr : Abf[l + 1], last ? 0 : 1);
        }
    }

    // The forward pass of forward() with every intermediate in caller-given
    // memory: a0/a1 BF16 activation ping-pong (rows x maxw), z FP32 (rows x
    // maxw); x (valid x in, FP32) and y (valid x out, FP32) are device arrays.
    void forward_into(const float* x, int valid, int rows, uint16_t* a0, uint16_t* a1, float* z, float* y) {
        const int in = dims[0], out = dims.back();
        hipLaunchKernelGGL(k_rows_to_bf16, dim3(blocks((int64_t)rows * in)), dim3(256), 0, s, valid, rows, in, x, a0);
        for (int l = 0; l < layers(); ++l) {
            const int K = dims[l], N = dims[l + 1];
            gemm(rows, N, K, a0, K, 1, wbf + woff[l], N, 1, z, N, false);
            const bool last = l == layers() - 1;
            hipLaunchKernelGGL(k_bias_act, dim3(blocks((int64_t)rows * N)), dim3(256), 0, s, rows, N, z, p + boff[l],
                               last ? nullptr : a1, last ? 0 : 1);
            std::swap(a0, a1);
        }
        hipLaunchKernelGGL(k_copy_f32, dim3(blocks((int64_t)valid * out)), dim3(256), 0, s, (int64_t)valid * out, z,
                           y);
    }

    // backward from dY (fp32 in dA, bf16 in dbf) -> grads in g
    void backward(int rows) {
        for (int l = layers() - 1; l >= 0; --l) {
            const int K = dims[l], N = dims[l + 1];
            // dW_l = A_l^T (K x rows) * dZ (rows x N)
            gemm(K, N, rows, Abf[l], 1, K, dbf, N, 1, g + woff[l], N, false);
            hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64), dim3(64 * kColsumGroups), 0, s, rows, N, dA,
                               g + boff[l]);
            if (l == 0) break;
            // dA_{l} = dZ (rows x N) * W_l^T (N x K)
            gemm(rows, K, N, dbf, N, 1, wbf + woff[l], 1, N, dB, K, false);
            // dZ_{l-1} = dA * relu'(Z_{l-1})
            hipLaunchKernelGGL(k_relu_back, dim3(blocks((int64_t)rows * K)), dim3(256), 0, s, (int64_t)rows * K, dB,
                               Z[l - 1], dbf);
            std::swap(dA, dB);
        }
    }

    void load_input(const float* x, int n, int rows) {
        const int in = dims[0];
        std::vector<uint16_t> h((size_t)rows * in, 0);
        for (int64_t k = 0; k < (int64_t)n * in; ++k) h[k] = f2bf(x[k]);
        MCHK(hipMemcpyAsync(Abf[0], h.data(), h.size() * 2, hipMemcpyHostToDevice, s));
    }

    float loss_and_grad(const float* x, const float* t, int n) {
        const int rows = (n + 31) / 32 * 32;
        if (rows > cap) throw MlpError(MC_ERR_INVALID_ARGUMENT, "batch larger than the network's capacity");
        const int out = dims.back();
        load_input(x, n, rows);
        std::vector<float> th((size_t)rows * out, 0.0f);
        std::memcpy(th.data(), t, sizeof(float) * (size_t)n * out);
        MCHK(hipMemcpyAsync(T, th.data(), th.size() * 4, hipMemcpyHostToDevice, s));
        forward(rows);
        const int nb = blocks((int64_t)rows * out);
        hipLaunchKernelGGL(k_mse_grad, dim3(nb), dim3(256), 0, s, rows, n, out, Z[layers() - 1], T, dA, dbf, part);
        backward(rows);
        std::vector<float> ph(nb);
        MCHK(hipMemcpyAsync(ph.data(), part, nb * 4, hipMemcpyDeviceToHost, s));
        MCHK(hipStreamSynchronize(s));
        double l = 0.0;
        for (float v : ph) l += v;
        return (float)(l / ((double)n * out));
    }
};

extern "C" {

const char* mcp_mlp_last_error(void) { return g_mlp_err.c_str(); }

int mcp_mlp_create(int32_t device, int32_t in_dim, int32_t hidden, int32_t hidden_layers, int32_t out_dim,
                   int32_t batch_capacity, uint64_t seed, mcp_mlp** out) {
    *out = nullptr;
    return mguard([&] {
        if (in_dim % 32 || hidden % 32 || out_dim % 32 || in_dim <= 0 || hidden <= 0 || out_dim <= 0 ||
            hidden_layers < 1 || batch_capacity < 32)
            throw MlpError(MC_ERR_INVALID_ARGUMENT, "MLP widths must be positive multiples of 32");
        MCHK(hipSetDevice(device));
        auto m = std::make_unique<mcp_mlp>();
        m->device = device;
        MCHK(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
        m->dims.push_back(in_dim);
        for (int l = 0; l < hidden_layers; ++l) m->dims.push_back(hidden);
        m->dims.push_back(out_dim);
        m->cap = (batch_capacity + 31) / 32 * 32;
        for (int l = 0; l < m->layers(); ++l) {
            m->woff.push_back(m->nparam);
            m->nparam += (int64_t)m->dims[l] * m->dims[l + 1];
            m->boff.push_back(m->nparam);
            m->nparam += m->dims[l + 1];
        }
        m->p = m->alloc<float>(m->nparam);
        m->g = m->alloc<float>(m->nparam);
        m->m1 = m->alloc<float>(m->nparam);
        m->m2 = m->alloc<float>(m->nparam);
        m->wbf = m->alloc<uint16_t>(m->nparam);
        m->Abf.push_back(m->alloc<uint16_t>((int64_t)m->cap * in_dim));
        for (int l = 0; l < m->layers(); ++l) {
            m->Z.push_back(m->alloc<float>((int64_t)m->cap * m->dims[l + 1]));
            if (l + 1 < m->layers()) m->Abf.push_back(m->alloc<uint16_t>((int64_t)m->cap * m->dims[l + 1]));
        }
        const int64_t wmax = (int64_t)m->cap * m->maxw();
        m->dA = m->alloc<float>(wmax);
        m->dB = m->alloc<float>(wmax);
        m->dbf = m->alloc<uint16_t>(wmax);
        m->T = m->alloc<float>((int64_t)m->cap * out_dim);
        m->part = m->alloc<float>(blocks((int64_t)m->cap * out_dim));
        // He-uniform weights, zero biases (seeded)
        std::vector<float> h(m->nparam, 0.0f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < m->layers(); ++l) {
            const double lim = std::sqrt(6.0 / m->dims[l]);
            for (int64_t k = 0; k < (int64_t)m->dims[l] * m->dims[l + 1]; ++k)
                h[m->woff[l] + k] = (float)(lim * (2.0 * ((rng() >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
        }
        MCHK(hipMemcpy(m->p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        MCHK(hipMemset(m->m1, 0, m->nparam * 4));
        MCHK(hipMemset(m->m2, 0, m->nparam * 4));
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m->nparam)), dim3(256), 0, m->s, m->nparam, m->p, m->wbf);
        MCHK(hipStreamSynchronize(m->s));
        *out = m.release();
    });
}

void mcp_mlp_destroy(mcp_mlp* m) { delete m; }

int mcp_mlp_param_count(const mcp_mlp* m, int64_t* n) {
    if (!m || !n) return MC_ERR_INVALID_ARGUMENT;
    *n = m->nparam;
    return MC_OK;
}

int mcp_mlp_get_params(mcp_mlp* m, float* params) {
    return mguard([&] {
        MCHK(hipMemcpy(params, m->p, m->nparam * 4, hipMemcpyDeviceToHost));
    });
}

int mcp_mlp_set_params(mcp_mlp* m, const float* params) {
    return mguard([&] {
        MCHK(hipMemcpy(m->p, params, m->nparam * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m->nparam)), dim3(256), 0, m->s, m->nparam, m->p, m->wbf);
        MCHK(hipStreamSynchronize(m->s));
    });
}

int mcp_mlp_forward(mcp_mlp* m, const float* x, int32_t n, float* y) {
    return mguard([&] {
        MCHK(hipSetDevice(m->device));
        const int out = m->dims.back();
        for (int off = 0; off < n; off += m->cap) {
            const int k = std::min(m->cap, n - off);
            const int rows = (k + 31) / 32 * 32;
            m->load_input(x + (int64_t)off * m->dims[0], k, rows);
            m->forward(rows);
            MCHK(hipMemcpyAsync(y + (int64_t)off * out, m->Z[m->layers() - 1], (size_t)k * out * 4,
                                hipMemcpyDeviceToHost, m->s));
            MCHK(hipStreamSynchronize(m->s));
        }
    });
}

int mcp_mlp_forward_noalloc(mcp_mlp* m, const float* x_dev, int32_t n, float* y_dev) {
    return mguard([&] {
        if (!m || n < 0 || (n > 0 && (!x_dev || !y_dev))) throw MlpError(MC_ERR_INVALID_ARGUMENT, "null argument");
        MCHK(hipSetDevice(m->device));
        char* base = static_cast<char*>(scratch_on(m->device));
        if (n == 0) return;
        const int rows = (n + 31) / 32 * 32;
        const size_t act = align256((size_t)rows * m->maxw() * sizeof(uint16_t));
        const size_t pre = align256((size_t)rows * m->maxw() * sizeof(float));
        if (2 * act + pre > kNoAllocScratchBytes)
            throw MlpError(MC_ERR_INVALID_ARGUMENT,
                           "mcp_mlp_forward_noalloc: " + std::to_string(n) + " samples need " +
                               std::to_string(2 * act + pre) + " bytes of scratch; the maximum is " +
                               std::to_string(kNoAllocScratchBytes) + " (" +
                               std::to_string(kNoAllocScratchBytes / (8 * (size_t)m->maxw()) / 32 * 32) +
                               " samples at this width)");
        uint16_t* a0 = reinterpret_cast<uint16_t*>(base);
        uint16_t* a1 = reinterpret_cast<uint16_t*>(base + act);
        float* z = reinterpret_cast<float*>(base + 2 * act);
        m->forward_into(x_dev, n, rows, a0, a1, z, y_dev);
        MCHK(hipGetLastError());
        MCHK(hipStreamSynchronize(m->s));
    });
}

int mcp_mlp_loss_grad(mcp_mlp* m, const float* x, const float* target, int32_t n, float* loss, float* grad) {
    return mguard([&] {
        MCHK(hipSetDevice(m->device));
        *loss = m->loss_and_grad(x, target, n);
        if (grad) MCHK(hipMemcpy(grad, m->g, m->nparam * 4, hipMemcpyDeviceToHost));
    });
}

int mcp_mlp_train(mcp_mlp* m, const float* x, const float* target, int32_t n, int32_t batch, int32_t epochs,
                  float lr, uint64_t seed, float* final_loss, double* train_ms) {
    return mguard([&] {
        MCHK(hipSetDevice(m->device));
        if (batch < 1 || batch > m->cap) throw MlpError(MC_ERR_INVALID_ARGUMENT, "batch outside 1..capacity");
        const int in = m->dims[0], out = m->dims.back();
        std::vector<int> order(n);
        for (int k = 0; k < n; ++k) order[k] = k;
        std::mt19937_64 rng(seed);
        std::vector<float> xb((size_t)batch * in), tb((size_t)batch * out);
        hipEvent_t e0, e1;
        MCHK(hipEventCreate(&e0));
        MCHK(hipEventCreate(&e1));
        MCHK(hipEventRecord(e0, m->s));
        double last = 0.0;
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            double acc = 0.0;
            int cnt = 0;
            for (int off = 0; off + batch <= n; off += batch) {
                for (int k = 0; k < batch; ++k) {
                    std::memcpy(&xb[(size_t)k * in], x + (int64_t)order[off + k] * in, sizeof(float) * in);
                    std::memcpy(&tb[(size_t)k * out], target + (int64_t)order[off + k] * out, sizeof(float) * out);
                }
                acc += m->loss_and_grad(xb.data(), tb.data(), batch);
                ++cnt;
                ++m->step;
                const float b1 = 0.9f, b2 = 0.999f;
                const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
                hipLaunchKernelGGL(k_adam, dim3(blocks(m->nparam)), dim3(256), 0, m->s, m->nparam, m->p, m->g, m->m1,
                                   m->m2, lr, b1, b2, 1e-8f, c1, c2, m->wbf);
            }
            last = cnt ? acc / cnt : 0.0;
        }
        MCHK(hipEventRecord(e1, m->s));
        MCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        MCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (final_loss) *final_loss = (float)last;
        if (train_ms) *train_ms = ms;
    });
}

int mcp_gemm_bf16_test(int32_t device, int32_t M, int32_t N, int32_t K, const float* A, const float* B, int32_t transpose_a,
                       int32_t transpose_b, float* C) {
    return mguard([&] {
        if (M % 32 || N % 32 || K % 8 || M <= 0 || N <= 0 || K <= 0)
            throw MlpError(MC_ERR_INVALID_ARGUMENT, "M, N multiples of 32 and K of 8");
        MCHK(hipSetDevice(device));
        std::vector<uint16_t> a((size_t)M * K), b((size_t)K * N);
        for (size_t k = 0; k < a.size(); ++k) a[k] = f2bf(A[k]);
        for (size_t k = 0; k < b.size(); ++k) b[k] = f2bf(B[k]);
        uint16_t *da = nullptr, *db = nullptr;
        float* dc = nullptr;
        MCHK(hipMalloc(&da, a.size() * 2));
        MCHK(hipMalloc(&db, b.size() * 2));
        MCHK(hipMalloc(&dc, (size_t)M * N * 4));
        MCHK(hipMemcpy(da, a.data(), a.size() * 2, hipMemcpyHostToDevice));
        MCHK(hipMemcpy(db, b.data(), b.size() * 2, hipMemcpyHostToDevice));
        // A given as M x K row-major (or K x M when transposed), B as K x N (or N x K)
        const int64_t sai = transpose_a ? 1 : K, sak = transpose_a ? M : 1;
        const int64_t sbk = transpose_b ? 1 : N, sbj = transpose_b ? K : 1;
        const int waves = (M / 32) * (N / 32);
        hipLaunchKernelGGL(k_gemm_bf16, dim3((waves + 3) / 4), dim3(256), 0, 0, M, N, K, da, sai, sak, db, sbk, sbj,
                           dc, N, 0);
        MCHK(hipGetLastError());
        MCHK(hipMemcpy(C, dc, (size_t)M * N * 4, hipMemcpyDeviceToHost));
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dc);
    });
}

}  // extern "C"
