// The excess-attenuation surrogate network on gfx950: BF16 MFMA products (v_mfma_f32_32x32x16_bf16)
// with FP32 accumulation, for training (forward with saved pre-activations, backward, Adam) and
// for inference. No counterpart in the reference; the network's shape and conventions are stated
// in include/qm_surrogate.h.
//
// One GEMM kernel serves every product of training: C[M][N] = A(m, k) B(k, n) with A and B BF16
// addressed by strides (so W, W^T, dZ^T need no transposed copies), 64x64 tiles per 4-wave
// workgroup (each wave one 32x32 MFMA block), K staged through LDS 32 at a time, and an epilogue
// per use: bias + GELU for a hidden layer (pre-activation FP32 kept, activation stored BF16), bias
// for the output, GELU' for a backward layer (gradient stored BF16), or plain FP32 partials for the
// weight gradients, whose K (the batch) is split across workgroups and summed in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/qm_surrogate.h"
#include "qm_internal.hpp"

typedef __bf16 bf16x8_t __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x16_t __attribute__((__vector_size__(16 * sizeof(float))));

namespace {

constexpr int LAYERS = QM_MLP_LAYERS;
constexpr int FAN_IN[LAYERS] = {QM_MLP_INPUTS, QM_MLP_HIDDEN, QM_MLP_HIDDEN, QM_MLP_HIDDEN, QM_MLP_HIDDEN};
constexpr int FAN_OUT[LAYERS] = {QM_MLP_HIDDEN, QM_MLP_HIDDEN, QM_MLP_HIDDEN, QM_MLP_HIDDEN, QM_MLP_OUTPUTS};
constexpr float GELU_C = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float GELU_A = 0.044715f;

// Offsets of each layer's weights and biases in the flat parameter vector.
struct Layout {
    size_t w[LAYERS], b[LAYERS];
    constexpr Layout() : w(), b() {
        size_t k = 0;
        for (int l = 0; l < LAYERS; ++l) {
            w[l] = k;
            k += (size_t)FAN_IN[l] * FAN_OUT[l];
            b[l] = k;
            k += FAN_OUT[l];
        }
    }
};
constexpr Layout LAYOUT;
static_assert(LAYOUT.b[LAYERS - 1] + QM_MLP_OUTPUTS == QM_MLP_PARAMS, "parameter count");

__device__ __forceinline__ uint16_t bf16_bits(float f) {
    // Round to nearest, ties to even (finite values).
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + tanhf(GELU_C * (z + GELU_A * z * z * z)));
}
__device__ __forceinline__ float gelu_grad(float z) {
    const float t = tanhf(GELU_C * (z + GELU_A * z * z * z));
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * GELU_C * (1.0f + 3.0f * GELU_A * z * z);
}

enum Epilogue : int {
    FORWARD_HIDDEN = 0,  // Z = acc + bias (FP32 kept), act = bf16(gelu(Z))
    FORWARD_OUTPUT = 1,  // Z = acc + bias
    BACKWARD_HIDDEN = 2, // act = bf16(acc * gelu'(Zin))
    PARTIAL = 3,         // Z[split] = acc
    INFER_HIDDEN = 4,    // act = bf16(gelu(acc + bias))
};

struct GemmArgs {
    int M, N, K, k_per_split;
    const uint16_t* A;
    long long sAm, sAk;
    const uint16_t* B;
    long long sBk, sBn;
    int mode;
    const float* bias;
    float* Z;
    long long ldz;
    uint16_t* act;
    long long ldact;
    const float* zin;
    long long ldzin;
};

constexpr int BM = 64, BN = 64, BK = 32, PAD = 8;

__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[BM][BK + PAD];  // [m][k]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN][BK + PAD];  // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int k_begin = blockIdx.z * g.k_per_split;
    const int k_end = min(g.K, k_begin + g.k_per_split);
    f32x16_t acc = {};
    for (int k0 = k_begin; k0 < k_end; k0 += BK) {
        if (g.sAk == 1) {
            const int m = tid >> 2, kq = (tid & 3) * 8;
            const bool row = m0 + m < g.M;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + kq + j;
                As[m][kq + j] = (row && k < k_end) ? g.A[(long long)(m0 + m) * g.sAm + k] : (uint16_t)0;
            }
        } else {
            const int k = tid >> 3, mq = (tid & 7) * 8;
            const bool in_k = k0 + k < k_end;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int m = m0 + mq + j;
                As[mq + j][k] = (in_k && m < g.M) ? g.A[(long long)m * g.sAm + (long long)(k0 + k) * g.sAk] : (uint16_t)0;
            }
        }
        if (g.sBk == 1) {
            const int n = tid >> 2, kq = (tid & 3) * 8;
            const bool col = n0 + n < g.N;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = k0 + kq + j;
                Bs[n][kq + j] = (col && k < k_end) ? g.B[k + (long long)(n0 + n) * g.sBn] : (uint16_t)0;
            }
        } else {
            const int k = tid >> 3, nq = (tid & 7) * 8;
            const bool in_k = k0 + k < k_end;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int n = n0 + nq + j;
                Bs[nq + j][k] = (in_k && n < g.N) ? g.B[(long long)(k0 + k) * g.sBk + (long long)n * g.sBn] : (uint16_t)0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&As[wm + (lane & 31)][s * 16 + 8 * (lane >> 5)]);
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&Bs[wn + (lane & 31)][s * 16 + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D[i][j] of the 32x32 block: lane holds column j = lane % 32 and rows 8*(r/4) + 4*(lane/32) + r%4.
    const int n = n0 + wn + (lane & 31);
    if (n >= g.N) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (m >= g.M) continue;
        const float v = acc[r];
        switch (g.mode) {
            case FORWARD_HIDDEN: {
                const float z = v + g.bias[n];
                g.Z[(long long)m * g.ldz + n] = z;
                g.act[(long long)m * g.ldact + n] = bf16_bits(gelu(z));
                break;
            }
            case FORWARD_OUTPUT:
                g.Z[(long long)m * g.ldz + n] = v + g.bias[n];
                break;
            case BACKWARD_HIDDEN:
                g.act[(long long)m * g.ldact + n] = bf16_bits(v * gelu_grad(g.zin[(long long)m * g.ldzin + n]));
                break;
            case PARTIAL:
                g.Z[(long long)blockIdx.z * g.M * g.ldz + (long long)m * g.ldz + n] = v;
                break;
            default:
                g.act[(long long)m * g.ldact + n] = bf16_bits(gelu(v + g.bias[n]));
                break;
        }
    }
}

// Inference, all five layers in one workgroup per 128 rows: the activations stay in one LDS
// buffer (128 x 264 BF16; a layer's products are held in the accumulators until every wave has
// read its inputs, then the epilogue overwrites them). Each of the four waves computes 64 output
// features of a hidden layer for all rows as D = W_block * act_block^T (2 feature blocks x 4 row
// blocks of 32x32 MFMA accumulators): a lane then holds one row and runs of four consecutive
// features, which the epilogue converts in pairs (v_cvt_pk_bf16_f32) and stores 8 bytes at a time.
// Weights are read from L2 a chunk of four K steps ahead; GELU is z * sigmoid(2u), the tanh
// form's identity, with the hardware exp2 and reciprocal.
#ifndef QM_INFER_ROWS
#define QM_INFER_ROWS 128
#endif
#ifndef QM_INFER_WAVES_PER_EU
#define QM_INFER_WAVES_PER_EU 2
#endif
constexpr int INFER_ROWS = QM_INFER_ROWS;
constexpr int INFER_PITCH = QM_MLP_HIDDEN + 8;
constexpr int RB = INFER_ROWS / 32;

typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));

// exp(-2u) = 2^(z (k + k A z^2)) with k = -2 sqrt(2/pi) log2(e): one multiply, one FMA and one
// multiply before the hardware exp2.
constexpr float GELU_K = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
constexpr float GELU_KA = GELU_K * 0.044715f;

__device__ __forceinline__ float gelu_infer(float z) {
    const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(z * z, GELU_KA, GELU_K));
    return z * __builtin_amdgcn_rcpf(1.0f + e);
}

// One hidden layer's products for a wave: features f0 .. f0+63 (two 32-feature blocks) over all
// INFER_ROWS rows, K in chunks of four 16-wide steps, the next chunk's weights loaded while this
// chunk's products run (the scheduling barriers keep the loads a whole chunk ahead).
template <int K>
__device__ __forceinline__ void hidden_products(const uint16_t (*act)[INFER_PITCH], const uint16_t* W0,
                                                const uint16_t* W1, const float* bias, int lr, int lk,
                                                f32x16_t (&acc)[2][RB]) {
    constexpr int STEPS = K / 16;
    constexpr int CHUNK = STEPS < 4 ? STEPS : 4;
    constexpr int CHUNKS = STEPS / CHUNK;
    static_assert(STEPS % CHUNK == 0, "K steps per chunk");
    bf16x8_t w[2][CHUNK][2];
#pragma unroll
    for (int i = 0; i < CHUNK; ++i) {
        w[0][i][0] = *reinterpret_cast<const bf16x8_t*>(W0 + 16 * i);
        w[0][i][1] = *reinterpret_cast<const bf16x8_t*>(W1 + 16 * i);
    }
    // The bias enters as the first step's accumulator input: element r of block fb is feature
    // 32 fb + 8 (r / 4) + r % 4 of the lane's run (bias already offset to the run's start).
    f32x16_t b[2];
#pragma unroll
    for (int fb = 0; fb < 2; ++fb)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float4 v = *reinterpret_cast<const float4*>(bias + 32 * fb + 8 * g);
            b[fb][4 * g + 0] = v.x;
            b[fb][4 * g + 1] = v.y;
            b[fb][4 * g + 2] = v.z;
            b[fb][4 * g + 3] = v.w;
        }
#pragma unroll
    for (int c = 0; c < CHUNKS; ++c) {
        if (c + 1 < CHUNKS) {
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                w[(c + 1) & 1][i][0] = *reinterpret_cast<const bf16x8_t*>(W0 + 16 * ((c + 1) * CHUNK + i));
                w[(c + 1) & 1][i][1] = *reinterpret_cast<const bf16x8_t*>(W1 + 16 * ((c + 1) * CHUNK + i));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < CHUNK; ++i) {
            const int st = c * CHUNK + i;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const bf16x8_t x = *reinterpret_cast<const bf16x8_t*>(&act[rb * 32 + lr][16 * st + lk]);
                acc[0][rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[c & 1][i][0], x, st ? acc[0][rb] : b[0], 0, 0, 0);
                acc[1][rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[c & 1][i][1], x, st ? acc[1][rb] : b[1], 0, 0, 0);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// `rows` rows, or 2 * *rows_pairs when that count is given (the rows a device kernel produced).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QM_INFER_WAVES_PER_EU))) void mlp_infer_kernel(
    const uint16_t* __restrict__ x, size_t rows, const uint32_t* __restrict__ rows_pairs,
    const uint16_t* __restrict__ wb, const float* __restrict__ w, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t act[INFER_ROWS][INFER_PITCH];
    if (rows_pairs) rows = 2 * (size_t)*rows_pairs;
    const size_t row0 = (size_t)blockIdx.x * INFER_ROWS;
    if (row0 >= rows) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int K0 = 48;  // the 40 inputs padded to three MFMA K steps
    for (int e = tid; e < INFER_ROWS * K0; e += 256) {
        const int r = e / K0, k = e % K0;
        act[r][k] = (k < QM_MLP_INPUTS && row0 + r < rows) ? x[(row0 + r) * QM_MLP_INPUTS + k] : (uint16_t)0;
    }
    __syncthreads();
    const int lr = lane & 31, lk = 8 * (lane >> 5), lq = 4 * (lane >> 5);
#pragma unroll 1
    for (int l = 0; l < LAYERS - 1; ++l) {
        const int fan_in = FAN_IN[l];
        const int f0 = wave * 64;
        const uint16_t* W0 = wb + LAYOUT.w[l] + (size_t)(f0 + lr) * fan_in + lk;
        const uint16_t* W1 = W0 + (size_t)32 * fan_in;
        f32x16_t acc[2][RB];
        const float* bias = w + LAYOUT.b[l] + f0 + lq;
        if (l == 0) hidden_products<K0>(act, W0, W1, bias, lr, lk, acc);
        else hidden_products<QM_MLP_HIDDEN>(act, W0, W1, bias, lr, lk, acc);
        __syncthreads();  // every wave has read this layer's inputs
#pragma unroll
        for (int fb = 0; fb < 2; ++fb)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                // Features f0 + 32 fb + 8 g + lq .. +3 of row (rb, lr), bias included.
                const int feature = f0 + fb * 32 + 8 * g + lq;
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    bf16x4_t v;
                    v.x = (__bf16)gelu_infer(acc[fb][rb][4 * g + 0]);
                    v.y = (__bf16)gelu_infer(acc[fb][rb][4 * g + 1]);
                    v.z = (__bf16)gelu_infer(acc[fb][rb][4 * g + 2]);
                    v.w = (__bf16)gelu_infer(acc[fb][rb][4 * g + 3]);
                    *reinterpret_cast<bf16x4_t*>(&act[rb * 32 + lr][feature]) = v;
                }
            }
        __syncthreads();
    }
    // The 8 outputs: 32-row blocks dealt over the waves, D = W5 (8 of 32 feature rows) * act^T.
    for (int rb = wave; rb < RB; rb += 4) {
        const uint16_t* W = wb + LAYOUT.w[LAYERS - 1];
        const float* bias = w + LAYOUT.b[LAYERS - 1];
        f32x16_t acc = {};
        const bf16x8_t zero = {};
#pragma unroll 4
        for (int k0 = 0; k0 < QM_MLP_HIDDEN; k0 += 16) {
            const bf16x8_t xa = *reinterpret_cast<const bf16x8_t*>(&act[rb * 32 + lr][k0 + lk]);
            const bf16x8_t wa =
                lr < QM_MLP_OUTPUTS ? *reinterpret_cast<const bf16x8_t*>(W + (size_t)lr * QM_MLP_HIDDEN + k0 + lk) : zero;
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, xa, acc, 0, 0, 0);
        }
        // acc[0..3]: features lq .. lq+3 of row (rb, lr).
        const size_t row = row0 + rb * 32 + lr;
        if (row < rows) {
            const float4 bc = *reinterpret_cast<const float4*>(bias + lq);
            *reinterpret_cast<float4*>(out + row * QM_MLP_OUTPUTS + lq) =
                make_float4(acc[0] + bc.x, acc[1] + bc.y, acc[2] + bc.z, acc[3] + bc.w);
        }
    }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, size_t n, uint16_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = bf16_bits(x[i]);
}

// dL/dy of the mean squared error, stored BF16, and per-block partial sums of the squared error.
__global__ __launch_bounds__(256) void output_gradient_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                                              size_t n, float scale, uint16_t* __restrict__ dy,
                                                              float* __restrict__ partial) {
    __shared__ float sums[256];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.0f;
    if (i < n) {
        const float d = y[i] - t[i];
        dy[i] = bf16_bits(2.0f * d * scale);
        sq = d * d;
    }
    sums[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sums[threadIdx.x] += sums[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sums[0];
}

// Sums of the split-K partials: out[i] = sum over s of partial[s][i], in split order.
__global__ void sum_splits_kernel(const float* __restrict__ partial, int splits, size_t n, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += partial[(size_t)k * n + i];
    out[i] = s;
}

// Column sums of a BF16 [rows][cols] matrix (the bias gradients): 32 columns a block, 8 row
// groups per column reduced in a fixed order.
__global__ __launch_bounds__(256) void column_sum_kernel(const uint16_t* __restrict__ d, int rows, int cols,
                                                         float* __restrict__ out) {
    __shared__ float part[8][32];
    const int c = blockIdx.x * 32 + (threadIdx.x & 31), grp = threadIdx.x >> 5;
    float s = 0.0f;
    if (c < cols)
        for (int r = grp; r < rows; r += 8) s += bf16_value(d[(size_t)r * cols + c]);
    part[grp][threadIdx.x & 31] = s;
    __syncthreads();
    if (grp == 0 && c < cols) {
        float t = 0.0f;
        for (int k = 0; k < 8; ++k) t += part[k][threadIdx.x & 31];
        out[c] = t;
    }
}

__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, uint16_t* __restrict__ wb, size_t n, float lr, float beta1,
                            float beta2, float eps, float correction1, float correction2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / correction1) / (sqrtf(vi / correction2) + eps);
    w[i] = wi;
    wb[i] = bf16_bits(wi);
}

int fail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

struct qm_mlp {
    qm_gpu* gpu = nullptr;
    size_t max_batch = 0;
    int step = 0;
    float* w = nullptr;       // FP32 parameters
    float* m = nullptr;       // Adam first moments
    float* v = nullptr;       // Adam second moments
    float* grad = nullptr;    // gradient of the last backward pass
    uint16_t* wb = nullptr;   // BF16 copy of the parameters (the MFMA operands)
    uint16_t* xb = nullptr;   // inputs, BF16 [rows][40]
    float* z[LAYERS - 1] = {};    // hidden pre-activations [rows][256]
    uint16_t* a[LAYERS - 1] = {}; // hidden activations, BF16 [rows][256]
    float* y = nullptr;           // outputs [rows][8]
    uint16_t* dy = nullptr;       // BF16 output gradient [rows][8]
    uint16_t* dz[2] = {};         // BF16 pre-activation gradients [rows][256]
    float* partial = nullptr;     // split-K partials of the weight gradients
    float* loss_partial = nullptr;
    int splits = 1;
};

namespace {

void free_all(qm_mlp* p) {
    void* bufs[] = {p->w, p->m, p->v, p->grad, p->wb, p->xb, p->y, p->dy, p->dz[0], p->dz[1], p->partial,
                    p->loss_partial};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l < LAYERS - 1; ++l) {
        if (p->z[l]) (void)hipFree(p->z[l]);
        if (p->a[l]) (void)hipFree(p->a[l]);
    }
}

hipError_t launch_gemm(const GemmArgs& g, int splits, hipStream_t stream) {
    dim3 grid(blocks_for(g.N, BN), blocks_for(g.M, BM), splits);
    hipLaunchKernelGGL(mlp_gemm_kernel, grid, dim3(256), 0, stream, g);
    return hipGetLastError();
}

GemmArgs gemm(int M, int N, int K, const uint16_t* A, long long sAm, long long sAk, const uint16_t* B, long long sBk,
              long long sBn, int mode) {
    GemmArgs g{};
    g.M = M;
    g.N = N;
    g.K = K;
    g.k_per_split = K;
    g.A = A;
    g.sAm = sAm;
    g.sAk = sAk;
    g.B = B;
    g.sBk = sBk;
    g.sBn = sBn;
    g.mode = mode;
    return g;
}

// The forward pass of `rows` training rows (inputs already in p->xb): p->z, p->a and p->y.
hipError_t forward(qm_mlp* p, int rows, hipStream_t s) {
    hipError_t e = hipSuccess;
    const uint16_t* in = p->xb;
    for (int l = 0; l < LAYERS && e == hipSuccess; ++l) {
        const int K = FAN_IN[l], N = FAN_OUT[l];
        GemmArgs g = gemm(rows, N, K, in, K, 1, p->wb + LAYOUT.w[l], 1, K,
                          l + 1 < LAYERS ? FORWARD_HIDDEN : FORWARD_OUTPUT);
        g.bias = p->w + LAYOUT.b[l];
        if (l + 1 < LAYERS) {
            g.Z = p->z[l];
            g.ldz = N;
            g.act = p->a[l];
            g.ldact = N;
            in = p->a[l];
        } else {
            g.Z = p->y;
            g.ldz = N;
        }
        e = launch_gemm(g, 1, s);
    }
    return e;
}

// The backward pass from p->dy into `grad` (QM_MLP_PARAMS floats).
hipError_t backward(qm_mlp* p, int rows, float* grad, hipStream_t s) {
    hipError_t e = hipSuccess;
    const uint16_t* d = p->dy;  // gradient at this layer's output, BF16 [rows][N]
    const int split_rows = 256;  // batch rows per split of the weight-gradient products
    for (int l = LAYERS - 1; l >= 0 && e == hipSuccess; --l) {
        const int K_in = FAN_IN[l], N = FAN_OUT[l];
        const uint16_t* a_prev = l > 0 ? p->a[l - 1] : p->xb;
        // dW[o][i] = sum_b d[b][o] a_prev[b][i]
        const int splits = (rows + split_rows - 1) / split_rows;
        GemmArgs gw = gemm(N, K_in, rows, d, 1, N, a_prev, K_in, 1, PARTIAL);
        gw.k_per_split = split_rows;
        gw.Z = p->partial;
        gw.ldz = K_in;
        e = launch_gemm(gw, splits, s);
        if (e != hipSuccess) break;
        const size_t nw = (size_t)N * K_in;
        hipLaunchKernelGGL(sum_splits_kernel, dim3(blocks_for(nw, 256)), dim3(256), 0, s, p->partial, splits, nw,
                           grad + LAYOUT.w[l]);
        hipLaunchKernelGGL(column_sum_kernel, dim3(blocks_for(N, 32)), dim3(256), 0, s, d, rows, N,
                           grad + LAYOUT.b[l]);
        e = hipGetLastError();
        if (e != hipSuccess || l == 0) break;
        // dA_prev[b][i] = sum_o d[b][o] W[o][i]; dZ_prev = dA_prev * gelu'(Z_prev)
        uint16_t* out = p->dz[l & 1];
        GemmArgs gd = gemm(rows, K_in, N, d, N, 1, p->wb + LAYOUT.w[l], K_in, 1, BACKWARD_HIDDEN);
        gd.act = out;
        gd.ldact = K_in;
        gd.zin = p->z[l - 1];
        gd.ldzin = K_in;
        e = launch_gemm(gd, 1, s);
        d = out;
    }
    return e;
}

// Inputs to BF16, forward, output gradient and loss partials, backward into `grad`.
hipError_t gradient(qm_mlp* p, const float* x, const float* t, int rows, float* grad, hipStream_t s) {
    const size_t nx = (size_t)rows * QM_MLP_INPUTS, ny = (size_t)rows * QM_MLP_OUTPUTS;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(nx, 256)), dim3(256), 0, s, x, nx, p->xb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = forward(p, rows, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(output_gradient_kernel, dim3(blocks_for(ny, 256)), dim3(256), 0, s, p->y, t, ny,
                       1.0f / (float)ny, p->dy, p->loss_partial);
    e = hipGetLastError();
    if (e == hipSuccess) e = backward(p, rows, grad, s);
    return e;
}

int read_loss(qm_mlp* p, int rows, hipStream_t s, float* loss, char* err, size_t errlen) {
    const size_t ny = (size_t)rows * QM_MLP_OUTPUTS;
    const unsigned nb = blocks_for(ny, 256);
    float* host = new (std::nothrow) float[nb];
    if (!host) return fail(err, errlen, "out of host memory");
    hipError_t e = hipMemcpyAsync(host, p->loss_partial, sizeof(float) * nb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    double sum = 0.0;
    for (unsigned i = 0; i < nb; ++i) sum += host[i];
    delete[] host;
    if (e != hipSuccess) return fail(err, errlen, "loss: %s", hipGetErrorString(e));
    *loss = (float)(sum / (double)ny);
    return 0;
}

}  // namespace

extern "C" int qm_mlp_new(qm_gpu* gpu, const float* params, size_t max_batch, qm_mlp** out, char* err,
                          size_t errlen) {
    if (!gpu || !params || !out) return fail(err, errlen, "null argument");
    if (max_batch > (1u << 24)) return fail(err, errlen, "max_batch %zu above %u", max_batch, 1u << 24);
    *out = nullptr;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e != hipSuccess) return fail(err, errlen, "hipSetDevice: %s", hipGetErrorString(e));
    qm_mlp* p = new (std::nothrow) qm_mlp;
    if (!p) return fail(err, errlen, "out of host memory");
    p->gpu = gpu;
    p->max_batch = max_batch;
    const size_t P = QM_MLP_PARAMS, B = max_batch;
    auto alloc = [&](void** ptr, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMalloc(ptr, bytes);
    };
    alloc((void**)&p->w, P * 4);
    alloc((void**)&p->m, P * 4);
    alloc((void**)&p->v, P * 4);
    alloc((void**)&p->grad, P * 4);
    alloc((void**)&p->wb, P * 2);
    if (B) {
        alloc((void**)&p->xb, B * QM_MLP_INPUTS * 2);
        for (int l = 0; l < LAYERS - 1; ++l) {
            alloc((void**)&p->z[l], B * QM_MLP_HIDDEN * 4);
            alloc((void**)&p->a[l], B * QM_MLP_HIDDEN * 2);
        }
        alloc((void**)&p->y, B * QM_MLP_OUTPUTS * 4);
        alloc((void**)&p->dy, B * QM_MLP_OUTPUTS * 2);
        alloc((void**)&p->dz[0], B * QM_MLP_HIDDEN * 2);
        alloc((void**)&p->dz[1], B * QM_MLP_HIDDEN * 2);
        alloc((void**)&p->partial, ((B + 255) / 256) * QM_MLP_HIDDEN * QM_MLP_HIDDEN * 4);
        alloc((void**)&p->loss_partial, blocks_for(B * QM_MLP_OUTPUTS, 256) * 4);
    }
    if (e == hipSuccess) e = hipMemcpy(p->w, params, P * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p->m, 0, P * 4);
    if (e == hipSuccess) e = hipMemset(p->v, 0, P * 4);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(P, 256)), dim3(256), 0, 0, p->w, P, p->wb);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_all(p);
        delete p;
        return fail(err, errlen, "qm_mlp_new: %s", hipGetErrorString(e));
    }
    *out = p;
    return 0;
}

extern "C" int qm_mlp_params(qm_mlp* p, float* params, char* err, size_t errlen) {
    if (!p || !params) return fail(err, errlen, "null argument");
    hipError_t e = hipSetDevice(p->gpu->ordinal);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(params, p->w, QM_MLP_PARAMS * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(err, errlen, "qm_mlp_params: %s", hipGetErrorString(e));
    return 0;
}

extern "C" void qm_mlp_free(qm_mlp* p) {
    if (!p) return;
    (void)hipSetDevice(p->gpu->ordinal);
    (void)hipDeviceSynchronize();
    free_all(p);
    delete p;
}

extern "C" int qm_mlp_gradient_device(qm_mlp* p, const float* d_inputs, const float* d_targets, size_t rows,
                                      float* d_gradient, float* loss, void* stream, char* err, size_t errlen) {
    if (!p || !d_inputs || !d_targets || !d_gradient) return fail(err, errlen, "null argument");
    if (rows == 0 || rows > p->max_batch) return fail(err, errlen, "rows %zu outside 1..%zu", rows, p->max_batch);
    hipError_t e = hipSetDevice(p->gpu->ordinal);
    if (e == hipSuccess) e = gradient(p, d_inputs, d_targets, (int)rows, d_gradient, (hipStream_t)stream);
    if (e != hipSuccess) return fail(err, errlen, "qm_mlp_gradient_device: %s", hipGetErrorString(e));
    return loss ? read_loss(p, (int)rows, (hipStream_t)stream, loss, err, errlen) : 0;
}

extern "C" int qm_mlp_train_step_device(qm_mlp* p, const float* d_inputs, const float* d_targets, size_t rows,
                                        float lr, float* loss, void* stream, char* err, size_t errlen) {
    if (!p || !d_inputs || !d_targets) return fail(err, errlen, "null argument");
    if (rows == 0 || rows > p->max_batch) return fail(err, errlen, "rows %zu outside 1..%zu", rows, p->max_batch);
    const hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSetDevice(p->gpu->ordinal);
    if (e == hipSuccess) e = gradient(p, d_inputs, d_targets, (int)rows, p->grad, s);
    if (e != hipSuccess) return fail(err, errlen, "qm_mlp_train_step_device: %s", hipGetErrorString(e));
    if (loss && read_loss(p, (int)rows, s, loss, err, errlen)) return 1;
    p->step += 1;
    const float c1 = 1.0f - powf(0.9f, (float)p->step), c2 = 1.0f - powf(0.999f, (float)p->step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(QM_MLP_PARAMS, 256)), dim3(256), 0, s, p->w, p->m, p->v, p->grad,
                       p->wb, (size_t)QM_MLP_PARAMS, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(err, errlen, "adam_kernel: %s", hipGetErrorString(e));
    return 0;
}

int qm::mlp_infer_rows(qm_mlp* p, const uint16_t* d_rows, size_t rows, const uint32_t* d_rows_pairs, float* d_outputs,
                       hipStream_t s, char* err, size_t errlen) {
    if (rows == 0) return 0;
    const size_t blocks = (rows + INFER_ROWS - 1) / INFER_ROWS;
    if (blocks > 0x7fffffffu) return fail(err, errlen, "too many rows");
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_rows, rows, d_rows_pairs, p->wb,
                       p->w, d_outputs);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(err, errlen, "mlp_infer_kernel: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int qm_mlp_infer_device(qm_mlp* p, const float* d_inputs, size_t rows, float* d_outputs, void* stream,
                                   char* err, size_t errlen) {
    if (!p || (rows && (!d_inputs || !d_outputs))) return fail(err, errlen, "null argument");
    if (rows == 0) return 0;
    if (rows > p->max_batch) return fail(err, errlen, "rows %zu above %zu", rows, p->max_batch);
    const hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSetDevice(p->gpu->ordinal);
    const size_t nx = rows * QM_MLP_INPUTS;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(nx, 256)), dim3(256), 0, s, d_inputs, nx, p->xb);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(err, errlen, "qm_mlp_infer_device: %s", hipGetErrorString(e));
    return qm::mlp_infer_rows(p, p->xb, rows, nullptr, d_outputs, s, err, errlen);
}
