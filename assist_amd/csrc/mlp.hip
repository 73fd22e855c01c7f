// Learned IAS15 start: a multilayer perceptron that maps a system's state to the step size
// IAS15 settles on and a warm start for its b coefficients, with its forward pass, backward
// pass and Adam update written here as HIP kernels for gfx950.
//
// Network: 16 inputs (10 used) -> 128 -> 128 -> 128 (tanh) -> 32 outputs (22 used).  Weights
// are kept as an FP32 master copy for the optimizer and a BF16 copy that the forward pass
// feeds to the matrix cores; every product runs on v_mfma_f32_16x16x16_bf16 (BF16 operands,
// FP32 accumulation).  Activations and gradients are stored in FP32 and rounded to BF16
// (nearest even) where they enter a product.
//
// GEMM mapping: one wave computes one 16 x 16 tile of C = op(A) op(B) (four waves, four
// tiles, per workgroup); each MFMA consumes a 16 x 16 slice of the reduction dimension.  For
// the weight gradients (reduction over the batch) the reduction is split across workgroups
// (blockIdx.z) and the partial tiles are added with atomics.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/assist.h"

namespace assist_gpu {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

template <typename T>
__device__ __forceinline__ uint16_t load_bf(const T* p, size_t i);
template <>
__device__ __forceinline__ uint16_t load_bf<float>(const float* p, size_t i) { return f2bf(p[i]); }
template <>
__device__ __forceinline__ uint16_t load_bf<uint16_t>(const uint16_t* p, size_t i) { return p[i]; }

struct GemmArgs {
    int M, N, K;
    const void* A;  // element (i, k): TA ? A[k * lda + i] : A[i * lda + k]
    int lda;
    const void* B;  // element (k, j): TB ? B[j * ldb + k] : B[k * ldb + j]
    int ldb;
    float* C;       // C[i * ldc + j]
    int ldc;
    const float* bias;  // [N] added before the activation, or null
    int act;            // kActNone or kActTanh
    int atomic;         // add into C with atomics (split reduction) instead of storing
    int kchunk;         // reduction length per blockIdx.z (a multiple of 16)
};

enum : int { kActNone = 0, kActTanh = 1 };

// MFMA 16x16x16 BF16 operand layout (one wave): lane l holds A[l % 16][4 (l / 16) + j] and
// B[4 (l / 16) + j][l % 16] for j = 0..3, and receives C[4 (l / 16) + j][l % 16].
template <typename TAe, typename TBe, bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int tiles_n = (g.N + 15) >> 4;
    const int tile = blockIdx.x * 4 + wave;
    const int tiles_m = (g.M + 15) >> 4;
    if (tile >= tiles_m * tiles_n) return;  // whole waves exit
    const int i0 = (tile / tiles_n) << 4;
    const int j0 = (tile % tiles_n) << 4;
    const int k_beg = blockIdx.z * g.kchunk;
    const int k_end = min(g.K, k_beg + g.kchunk);
    const TAe* A = (const TAe*)g.A;
    const TBe* B = (const TBe*)g.B;
    const int r = lane & 15;
    const int q = (lane >> 4) << 2;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = k_beg; k0 < k_end; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + q + j;
            const int ia = i0 + r, jb = j0 + r;
            uint16_t av = 0, bv = 0;
            if (ia < g.M && k < k_end) av = load_bf<TAe>(A, TA ? (size_t)k * g.lda + ia : (size_t)ia * g.lda + k);
            if (jb < g.N && k < k_end) bv = load_bf<TBe>(B, TB ? (size_t)jb * g.ldb + k : (size_t)k * g.ldb + jb);
            a[j] = (short)av;
            b[j] = (short)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int jc = j0 + r;
    if (jc >= g.N) return;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int ic = i0 + q + j;
        if (ic >= g.M) continue;
        float v = acc[j];
        if (g.bias) v += g.bias[jc];
        if (g.act == kActTanh) v = tanhf(v);
        float* dst = g.C + (size_t)ic * g.ldc + jc;
        if (g.atomic) atomicAdd(dst, v);
        else *dst = v;
    }
}

// dZ = dH * (1 - H^2) (tanh derivative from its output), in place on dH.
__global__ void tanh_backward_kernel(float* dH, const float* H, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dH[i] *= 1.f - H[i] * H[i];
}

// Column sums of X[M][N] added into out[N] (bias gradients).
__global__ void colsum_kernel(const float* X, int M, int N, float* out) {
    const int j = blockIdx.y;
    float s = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) s += X[(size_t)i * N + j];
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        atomicAdd(out + j, t);
    }
}

// Weighted squared error: loss += sum w_j (Y - T)^2 / M, dY = 2 w_j (Y - T) / M.
__global__ void mse_kernel(const float* Y, const float* T, const float* w, int M, int N, float* dY, float* loss) {
    const size_t n = (size_t)M * N;
    float s = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(i % N);
        const float d = Y[i] - T[i];
        s += w[j] * d * d;
        dY[i] = 2.f * w[j] * d / (float)M;
    }
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += part[k];
        atomicAdd(loss, t / (float)M);
    }
}

// Adam (Kingma & Ba 2015) with bias correction; refreshes the BF16 copy of the weights.
__global__ void adam_kernel(float* p, uint16_t* p16, const float* g, float* m, float* v, size_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    p16[i] = f2bf(pi);
}

__global__ void to_bf16_kernel(const float* p, uint16_t* p16, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p16[i] = f2bf(p[i]);
}

__global__ void fill_kernel(float* p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- prediction: the four layers fused in one kernel ----------------------------------------
// The forward pass for prediction keeps no activations: each wave takes 16 rows at a time
// through all four layers, with the BF16 weights staged once per workgroup in LDS (transposed,
// [out][in], so that a lane's four consecutive inputs of one output are one 8-byte read) and the
// 16 x 128 activation tile in the wave's own LDS area, through which the MFMA result layout
// (four rows of one column per lane) becomes the next operand layout (four columns of one row).
// Every product is the same v_mfma_f32_16x16x16_bf16 over the same 16-wide slices of the
// reduction in the same order as gemm_bf16_kernel, with the same bias, tanh and BF16 rounding,
// so the outputs equal assist_mlp_forward's bit for bit.
// (LDS rows are padded by 4 elements, 8 bytes: the 16 lanes that read one operand slice then
// fall in distinct banks instead of all in one)
constexpr int kPIn = ASSIST_MLP_IN, kPH = ASSIST_MLP_HIDDEN, kPOut = ASSIST_MLP_OUT;
constexpr int kPad = 4;
constexpr int kPW1 = 0, kPW2 = kPW1 + kPH * (kPIn + kPad), kPW3 = kPW2 + kPH * (kPH + kPad),
              kPW4 = kPW3 + kPH * (kPH + kPad);
constexpr int kPWords = kPW4 + kPOut * (kPH + kPad);  // BF16 weights in LDS
constexpr int kPBias = 3 * kPH + kPOut;                 // FP32 biases in LDS
constexpr int kPWaves = 16;  // four waves per SIMD share one copy of the weights
constexpr int kPTileRow = kPH + kPad;
constexpr size_t kPLdsBytes = (size_t)kPWords * 2 + (size_t)kPBias * 4 + (size_t)kPWaves * 16 * kPTileRow * 2;

struct PredictArgs {
    const uint16_t* p16;  // BF16 parameters (layer by layer: W[in][out] row-major, then b[out])
    const float* p;       // FP32 parameters (the biases are read from here)
    const float* X;       // [M][16]
    float* Y;             // [M][32]
    int M;
};

// One layer of a 16-row tile: out = act(A W + b) for K = KC * 16 inputs and NT * 16 outputs.
// a[KC]: this lane's operand fragments (row l % 16, inputs 16 kc + 4 (l / 16) + 0..3).
template <int KC, int NT, bool Last>
__device__ __forceinline__ void predict_layer(const bf16x4 (&a)[KC], const uint16_t* Wt, const float* bias,
                                              uint16_t* tile, float* Y, int row0, int M) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, q = (lane >> 4) << 2;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* w = Wt + (size_t)(nt * 16 + r) * (KC * 16 + kPad) + q;
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
            const bf16x4 b = *(const bf16x4*)(w + kc * 16);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kc], b, acc, 0, 0, 0);
        }
        const int col = nt * 16 + r;
        const float bc = bias[col];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float v = acc[j] + bc;
            if constexpr (Last) {
                const int row = row0 + q + j;
                if (row < M) Y[(size_t)row * kPOut + col] = v;
            } else {
                tile[(q + j) * kPTileRow + col] = f2bf(tanhf(v));
            }
        }
    }
}

__global__ void __launch_bounds__(64 * kPWaves) mlp_predict_kernel(PredictArgs g) {
    extern __shared__ __align__(16) unsigned char s_pred[];
    uint16_t* const sw = (uint16_t*)s_pred;
    float* const sb = (float*)(s_pred + (size_t)kPWords * 2);
    uint16_t* const tiles = (uint16_t*)(s_pred + (size_t)kPWords * 2 + (size_t)kPBias * 4);
    // stage the weights transposed ([out][in]) and the biases
    {
        const int dims[5] = {kPIn, kPH, kPH, kPH, kPOut};
        const int dst[4] = {kPW1, kPW2, kPW3, kPW4};
        size_t src = 0;
        int boff = 0;
        for (int l = 0; l < 4; l++) {
            const int K = dims[l], N = dims[l + 1];
            for (int i = threadIdx.x; i < K * N; i += blockDim.x) {
                const int k = i / N, n = i - k * N;
                sw[dst[l] + n * (K + kPad) + k] = g.p16[src + i];
            }
            src += (size_t)K * N;
            for (int i = threadIdx.x; i < N; i += blockDim.x) sb[boff + i] = g.p[src + i];
            src += N;
            boff += N;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, q = (lane >> 4) << 2;
    uint16_t* const tile = tiles + (size_t)wave * 16 * kPTileRow;
    const int ntiles = (g.M + 15) >> 4;
    for (int t = blockIdx.x * kPWaves + wave; t < ntiles; t += gridDim.x * kPWaves) {
        const int row0 = t << 4;
        bf16x4 a1[1];
        {
            const int row = row0 + r;
#pragma unroll
            for (int j = 0; j < 4; j++) a1[0][j] = (short)(row < g.M ? f2bf(g.X[(size_t)row * kPIn + q + j]) : 0);
        }
        predict_layer<1, kPH / 16, false>(a1, sw + kPW1, sb, tile, nullptr, row0, g.M);
        bf16x4 a[kPH / 16];
        // (LDS operations of one wave complete in order; the fence keeps the compiler's order)
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
        for (int kc = 0; kc < kPH / 16; kc++) a[kc] = *(const bf16x4*)(tile + r * kPTileRow + kc * 16 + q);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        predict_layer<kPH / 16, kPH / 16, false>(a, sw + kPW2, sb + kPH, tile, nullptr, row0, g.M);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
        for (int kc = 0; kc < kPH / 16; kc++) a[kc] = *(const bf16x4*)(tile + r * kPTileRow + kc * 16 + q);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        predict_layer<kPH / 16, kPH / 16, false>(a, sw + kPW3, sb + 2 * kPH, tile, nullptr, row0, g.M);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
        for (int kc = 0; kc < kPH / 16; kc++) a[kc] = *(const bf16x4*)(tile + r * kPTileRow + kc * 16 + q);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        predict_layer<kPH / 16, kPOut / 16, true>(a, sw + kPW4, sb + 3 * kPH, tile, g.Y, row0, g.M);
    }
}

}  // namespace assist_gpu

using namespace assist_gpu;

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
namespace {

constexpr int kLayers = 4;
constexpr int kDims[kLayers + 1] = {ASSIST_MLP_IN, ASSIST_MLP_HIDDEN, ASSIST_MLP_HIDDEN, ASSIST_MLP_HIDDEN, ASSIST_MLP_OUT};

size_t w_off(int l) {
    size_t o = 0;
    for (int i = 0; i < l; i++) o += (size_t)kDims[i] * kDims[i + 1] + kDims[i + 1];
    return o;
}
size_t b_off(int l) { return w_off(l) + (size_t)kDims[l] * kDims[l + 1]; }

bool ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "(ASSIST) HIP error in %s: %s\n", what, hipGetErrorString(e));
    return false;
}

template <typename TAe, typename TBe, bool TA, bool TB>
hipError_t gemm(const GemmArgs& g, int zsplit, hipStream_t s) {
    const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
    dim3 grid((tiles + 3) / 4, 1, zsplit);
    hipLaunchKernelGGL((gemm_bf16_kernel<TAe, TBe, TA, TB>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

unsigned grid1(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct assist_mlp {
    int device = 0;
    int max_batch = 0;
    float* p = nullptr;      // FP32 master parameters
    uint16_t* p16 = nullptr; // BF16 weights for the matrix cores
    float* g = nullptr;      // gradients
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    float* act = nullptr;    // H1, H2, H3, Y: max_batch x (3 * hidden + out)
    float* dact = nullptr;   // dH (max_batch x hidden) x 2 ping-pong + dY
    float* wout = nullptr;   // output weights of the loss [out]
    float* loss = nullptr;   // [1]
};

static void mlp_release(assist_mlp* n) {
    if (!n) return;
    for (void* q : {(void*)n->p, (void*)n->p16, (void*)n->g, (void*)n->m, (void*)n->v, (void*)n->act, (void*)n->dact,
                    (void*)n->wout, (void*)n->loss})
        if (q) (void)hipFree(q);
    delete n;
}

static size_t act_stride(int max_batch) { return (size_t)max_batch; }

extern "C" {

int assist_mlp_num_params(void) { return (int)w_off(kLayers); }

struct assist_mlp* assist_mlp_create(int max_batch, unsigned long long seed) {
    if (max_batch <= 0) return nullptr;
    auto* n = new assist_mlp();
    n->max_batch = max_batch;
    const size_t np = w_off(kLayers);
    const size_t H = ASSIST_MLP_HIDDEN, O = ASSIST_MLP_OUT;
    const size_t B = act_stride(max_batch);
    bool good = ok(hipGetDevice(&n->device), "hipGetDevice") && ok(hipMalloc(&n->p, np * 4), "hipMalloc") &&
                ok(hipMalloc(&n->p16, np * 2), "hipMalloc") && ok(hipMalloc(&n->g, np * 4), "hipMalloc") &&
                ok(hipMalloc(&n->m, np * 4), "hipMalloc") && ok(hipMalloc(&n->v, np * 4), "hipMalloc") &&
                ok(hipMalloc(&n->act, B * (3 * H + O) * 4), "hipMalloc") &&
                ok(hipMalloc(&n->dact, B * (2 * H + O) * 4), "hipMalloc") && ok(hipMalloc(&n->wout, O * 4), "hipMalloc") &&
                ok(hipMalloc(&n->loss, 4), "hipMalloc");
    if (!good) {
        mlp_release(n);
        return nullptr;
    }
    // Glorot-uniform weights from a 64-bit LCG, zero biases.
    std::vector<float> h(np, 0.f);
    unsigned long long st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    for (int l = 0; l < kLayers; l++) {
        const double lim = sqrt(6.0 / (kDims[l] + kDims[l + 1]));
        for (size_t i = 0; i < (size_t)kDims[l] * kDims[l + 1]; i++) {
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            const double u = (double)(st >> 11) * (1.0 / 9007199254740992.0);
            h[w_off(l) + i] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    std::vector<float> ones(O, 1.f);
    if (assist_mlp_set_params(n, h.data()) != 0 || !ok(hipMemcpy(n->wout, ones.data(), O * 4, hipMemcpyHostToDevice), "hipMemcpy")) {
        mlp_release(n);
        return nullptr;
    }
    return n;
}

void assist_mlp_free(struct assist_mlp* n) { mlp_release(n); }

int assist_mlp_max_batch(const struct assist_mlp* n) { return n ? n->max_batch : 0; }

int assist_mlp_set_params(struct assist_mlp* n, const float* host) {
    if (!n || !host) return ASSIST_BATCH_ERROR_ARG;
    const size_t np = w_off(kLayers);
    if (!ok(hipMemcpy(n->p, host, np * 4, hipMemcpyHostToDevice), "hipMemcpy") ||
        !ok(hipMemset(n->m, 0, np * 4), "hipMemset") || !ok(hipMemset(n->v, 0, np * 4), "hipMemset"))
        return ASSIST_BATCH_ERROR_DEVICE;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1(np)), dim3(256), 0, 0, n->p, n->p16, np);
    if (!ok(hipGetLastError(), "to_bf16") || !ok(hipDeviceSynchronize(), "to_bf16")) return ASSIST_BATCH_ERROR_DEVICE;
    return 0;
}

int assist_mlp_get_params(const struct assist_mlp* n, float* host) {
    if (!n || !host) return ASSIST_BATCH_ERROR_ARG;
    return ok(hipMemcpy(host, n->p, w_off(kLayers) * 4, hipMemcpyDeviceToHost), "hipMemcpy") ? 0 : ASSIST_BATCH_ERROR_DEVICE;
}

int assist_mlp_set_output_weights(struct assist_mlp* n, const float* host_w) {
    if (!n || !host_w) return ASSIST_BATCH_ERROR_ARG;
    return ok(hipMemcpy(n->wout, host_w, ASSIST_MLP_OUT * 4, hipMemcpyHostToDevice), "hipMemcpy") ? 0
                                                                                              : ASSIST_BATCH_ERROR_DEVICE;
}

// Forward pass of M rows (M <= max_batch): d_X [M][16] -> d_Y [M][32] (device pointers).  The
// hidden activations stay in the network's workspace for a following backward pass.
int assist_mlp_forward(struct assist_mlp* n, const float* d_X, int M, float* d_Y, void* stream) {
    if (!n || !d_X || M < 0 || M > n->max_batch) return ASSIST_BATCH_ERROR_ARG;
    if (M == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int H = ASSIST_MLP_HIDDEN;
    const size_t B = act_stride(n->max_batch);
    const float* in = d_X;
    int ld_in = ASSIST_MLP_IN;
    for (int l = 0; l < kLayers; l++) {
        GemmArgs g;
        memset(&g, 0, sizeof g);
        g.M = M;
        g.N = kDims[l + 1];
        g.K = kDims[l];
        g.A = in;
        g.lda = ld_in;
        g.B = n->p16 + w_off(l);
        g.ldb = kDims[l + 1];
        float* out = (l < kLayers - 1) ? n->act + (size_t)l * B * H : (d_Y ? d_Y : n->act + 3 * B * H);
        g.C = out;
        g.ldc = kDims[l + 1];
        g.bias = n->p + b_off(l);
        g.act = (l < kLayers - 1) ? kActTanh : kActNone;
        g.kchunk = ((g.K + 15) / 16) * 16;
        if (!ok(gemm<float, uint16_t, false, false>(g, 1, s), "mlp forward")) return ASSIST_BATCH_ERROR_DEVICE;
        in = out;
        ld_in = kDims[l + 1];
    }
    return 0;
}

// Prediction forward pass of any number of rows: d_X [M][16] -> d_Y [M][32], keeping no
// activations (the fused kernel above; outputs equal assist_mlp_forward's bit for bit).
int assist_mlp_predict(struct assist_mlp* n, const float* d_X, int M, float* d_Y, void* stream) {
    if (!n || !d_X || !d_Y || M < 0) return ASSIST_BATCH_ERROR_ARG;
    if (M == 0) return 0;
    if (!ok(hipFuncSetAttribute((const void*)mlp_predict_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kPLdsBytes),
            "hipFuncSetAttribute(mlp_predict)"))
        return ASSIST_BATCH_ERROR_DEVICE;
    int dev = 0, ncu = 0;
    if (!ok(hipGetDevice(&dev), "hipGetDevice") ||
        !ok(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute"))
        return ASSIST_BATCH_ERROR_DEVICE;
    PredictArgs g;
    g.p16 = n->p16;
    g.p = n->p;
    g.X = d_X;
    g.Y = d_Y;
    g.M = M;
    // one workgroup per CU (the staged weights and 16 activation tiles fill its LDS), or fewer
    // for small batches
    const int tiles = (M + 15) / 16;
    int blocks = (tiles + kPWaves - 1) / kPWaves;
    if (blocks > ncu) blocks = ncu;
    hipLaunchKernelGGL(mlp_predict_kernel, dim3(blocks), dim3(64 * kPWaves), kPLdsBytes, (hipStream_t)stream, g);
    return ok(hipGetLastError(), "mlp predict") ? 0 : ASSIST_BATCH_ERROR_DEVICE;
}

// Forward and backward pass of the weighted squared error against d_T [M][32]; leaves the
// gradients in the network (and copies them to host_grad when given) and returns the loss.
int assist_mlp_loss_grad(struct assist_mlp* n, const float* d_X, const float* d_T, int M, float* loss,
                         float* host_grad, void* stream) {
    if (!n || !d_X || !d_T || M <= 0 || M > n->max_batch) return ASSIST_BATCH_ERROR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int H = ASSIST_MLP_HIDDEN, O = ASSIST_MLP_OUT;
    const size_t B = act_stride(n->max_batch);
    const size_t np = w_off(kLayers);
    float* Y = n->act + 3 * B * H;
    int rc = assist_mlp_forward(n, d_X, M, nullptr, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(fill_kernel, dim3(grid1(np)), dim3(256), 0, s, n->g, np, 0.f);
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(256), 0, s, n->loss, (size_t)1, 0.f);
    float* dY = n->dact + 2 * B * H;
    hipLaunchKernelGGL(mse_kernel, dim3(256), dim3(256), 0, s, (const float*)Y, d_T, (const float*)n->wout, M, O, dY,
                       n->loss);
    if (!ok(hipGetLastError(), "mse")) return ASSIST_BATCH_ERROR_DEVICE;
    // walk the layers backwards; dcur = gradient w.r.t. the pre-activation of layer l's output
    float* dcur = dY;
    for (int l = kLayers - 1; l >= 0; l--) {
        const int K = kDims[l], N = kDims[l + 1];
        const float* Hin = (l == 0) ? d_X : n->act + (size_t)(l - 1) * B * H;
        const int ld_in = K;
        // dW_l = Hin^T dcur  (reduction over the batch, split over blockIdx.z)
        {
            GemmArgs g;
            memset(&g, 0, sizeof g);
            g.M = K;
            g.N = N;
            g.K = M;
            g.A = Hin;
            g.lda = ld_in;
            g.B = dcur;
            g.ldb = N;
            g.C = n->g + w_off(l);
            g.ldc = N;
            g.atomic = 1;
            const int zs = (M + 1023) / 1024;
            g.kchunk = ((M + zs - 1) / zs + 15) / 16 * 16;
            const int zsplit = (M + g.kchunk - 1) / g.kchunk;
            if (!ok(gemm<float, float, true, false>(g, zsplit, s), "mlp dW")) return ASSIST_BATCH_ERROR_DEVICE;
        }
        // db_l = column sums of dcur
        hipLaunchKernelGGL(colsum_kernel, dim3((M + 1023) / 1024, N), dim3(256), 0, s, (const float*)dcur, M, N,
                           n->g + b_off(l));
        if (!ok(hipGetLastError(), "colsum")) return ASSIST_BATCH_ERROR_DEVICE;
        if (l == 0) break;
        // dHin = dcur W_l^T, then through tanh
        float* dnext = n->dact + (size_t)(l % 2) * B * H;
        GemmArgs g;
        memset(&g, 0, sizeof g);
        g.M = M;
        g.N = K;
        g.K = N;
        g.A = dcur;
        g.lda = N;
        g.B = n->p16 + w_off(l);
        g.ldb = N;  // W stored [K][N]; element (k = n_idx, j = k_idx) of W^T is W[j * N + k]
        g.C = dnext;
        g.ldc = K;
        g.kchunk = ((N + 15) / 16) * 16;
        if (!ok(gemm<float, uint16_t, false, true>(g, 1, s), "mlp dX")) return ASSIST_BATCH_ERROR_DEVICE;
        const size_t cnt = (size_t)M * K;
        hipLaunchKernelGGL(tanh_backward_kernel, dim3(grid1(cnt)), dim3(256), 0, s, dnext,
                           (const float*)(n->act + (size_t)(l - 1) * B * H), cnt);
        if (!ok(hipGetLastError(), "tanh backward")) return ASSIST_BATCH_ERROR_DEVICE;
        dcur = dnext;
    }
    if (loss && !ok(hipMemcpyAsync(loss, n->loss, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) return ASSIST_BATCH_ERROR_DEVICE;
    if (host_grad && !ok(hipMemcpyAsync(host_grad, n->g, np * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
        return ASSIST_BATCH_ERROR_DEVICE;
    if (!ok(hipStreamSynchronize(s), "mlp backward")) return ASSIST_BATCH_ERROR_DEVICE;
    return 0;
}

int assist_mlp_adam_step(struct assist_mlp* n, float lr, float beta1, float beta2, float eps, int step, void* stream) {
    if (!n || step < 1) return ASSIST_BATCH_ERROR_ARG;
    const size_t np = w_off(kLayers);
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid1(np)), dim3(256), 0, (hipStream_t)stream, n->p, n->p16, (const float*)n->g,
                       n->m, n->v, np, lr, beta1, beta2, eps, c1, c2);
    return ok(hipGetLastError(), "adam") ? 0 : ASSIST_BATCH_ERROR_DEVICE;
}

}  // extern "C"
