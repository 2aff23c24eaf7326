// SPDX-License-Identifier: 0BSD
//
// The surrogate of log|M|^2: a multilayer perceptron (inputs -> hidden x layers, SiLU -> 1) with
// BF16 weights and activations in the matrix products, FP32 accumulation, FP32 master weights,
// trained by Adam, all on the GPU. No reference counterpart (the reference has no learned part).
//
// Matrix products run on the CDNA matrix cores: v_mfma_f32_16x16x16_bf16 through
// __builtin_amdgcn_mfma_f32_16x16x16bf16_1k. For a 16x16x16 block, lane l supplies row l%16 of
// A and column l%16 of B at k = 4*(l/16) .. 4*(l/16)+3 (four bf16 each) and receives D rows
// 4*(l/16) .. +3 of column l%16. Every product here has both operands contiguous along k
// ("NT": C[m][n] = sum_k A[m][k] B[n][k]), so each lane's operand is one 8-byte load.
//
// Forward: one workgroup (4 wavefronts) takes 64 points through every layer; the activations
// stay in LDS (bf16, two buffers), the weights come from L2. In training mode it also keeps,
// per layer, the pre-activations Z (fp32, [point][unit]) and the activations transposed
// (bf16, [unit][point]) for the backward products.
//
// Backward (loss = mean of (y - t)^2): the output layer by reductions, then per hidden layer
// dW_l = dZ_l^T H_(l-1) (K = batch) and dH_(l-1) = dZ_l W_l (K = hidden) as NT products with
// the SiLU derivative fused into the second one's epilogue.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pacman_recurrence.h"
#include "pacuda_amd.h"
#include "runtime.hpp"

using namespace pacuda;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

constexpr int TILE = 64;  // points per workgroup tile and the side of a wavefront's output tile

__device__ __forceinline__ bf16 to_bf16(float f) {  // round to nearest even (finite values)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
    float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct Net {
    uint32_t inputs, in_pad, hidden, layers;
    // fp32 master parameters: per hidden layer W [hidden][fan_in] then b [hidden]; then
    // W_out [hidden], b_out [1]; fan_in = inputs for layer 0 (padded copies live in bf16 only)
    float *params, *grads, *m, *v;
    uint64_t count;
    std::vector<uint64_t> w_off, b_off;  // offsets per hidden layer
    uint64_t wout_off, bout_off;
    bf16 *wb;                            // bf16 weights per hidden layer, [hidden][fan_in_pad]
    bf16 *wtb;                           // bf16 transposed weights of layers >= 1, [fan_in][hidden]
    std::vector<uint64_t> wb_off, wtb_off;
    // training workspace (max_batch points)
    uint32_t max_batch;
    float *z;        // [layers][batch][hidden]
    bf16 *ht;        // [layers + 1][width][batch]: x^T (in_pad rows) then H_l^T (hidden rows)
    bf16 *dz[2], *dzt;  // [batch][hidden] (two: a layer's dZ is read while the next is written), [hidden][batch]
    float *delta, *loss;
};

// ---- forward -------------------------------------------------------------------------------

struct FwdArgs {
    const float *x;
    float *y;
    uint32_t batch, inputs, in_pad, hidden, layers;
    const bf16 *wb;
    const uint64_t *wb_off;  // device copy of per-layer offsets
    const float *params;
    const uint64_t *b_off;
    uint64_t wout_off, bout_off;
    float *z;
    bf16 *ht;
    uint32_t train;
};

// LDS position of activation (r, c) in a row of `stride` bf16: the 4-element granule index is
// XORed with r % 16, so the 16 rows one MFMA operand load touches fall on distinct banks
__device__ __forceinline__ uint32_t lds_at(uint32_t r, uint32_t c, uint32_t stride) {
    return r * stride + ((((c >> 2) ^ (r & 15))) << 2) + (c & 3);
}

// ROWS blocks of 16 points per workgroup (FT = 16 ROWS points): small batches take ROWS = 1 so
// that they still spread over the compute units
template <int ROWS>
__global__ void __launch_bounds__(256) mlp_forward(FwdArgs a) {
    constexpr uint32_t FT = 16 * ROWS;
    extern __shared__ bf16 lds[];
    const uint32_t H = a.hidden, stride = H;  // two FT x H bf16 buffers: 64 KB at FT = 64, H = 256
    bf16 *buf[2] = {lds, lds + FT * stride};
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t row0 = blockIdx.x * FT;
    // the input tile, bf16, padded with zeros to in_pad columns
    for (uint32_t idx = tid; idx < FT * a.in_pad; idx += blockDim.x) {
        uint32_t r = idx / a.in_pad, c = idx % a.in_pad;
        float v = c < a.inputs ? a.x[(size_t)(row0 + r) * a.inputs + c] : 0.0f;
        bf16 h = to_bf16(v);
        buf[0][lds_at(r, c, stride)] = h;
        if (a.train) a.ht[(size_t)c * a.batch + row0 + r] = h;
    }
    __syncthreads();
    const uint32_t cols_per_wave = H / 4;  // 16, 32 or 64
    const uint32_t col_blocks = cols_per_wave / 16;
    uint32_t cur = 0;
    for (uint32_t l = 0; l < a.layers; l++) {
        const uint32_t K = l == 0 ? a.in_pad : H;
        const bf16 *W = a.wb + a.wb_off[l];
        const float *bias = a.params + a.b_off[l];
        const uint32_t colbase = wave * cols_per_wave;
        // constant loop bounds with predicates: the accumulators stay in registers
        f32x4 acc[ROWS][4];
#pragma unroll
        for (int i = 0; i < ROWS; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (uint32_t k0 = 0; k0 < K; k0 += 16) {
            const uint32_t kk = k0 + 4 * (lane >> 4);
            bf16x4 av[ROWS], bv[4];
#pragma unroll
            for (int i = 0; i < ROWS; i++)
                av[i] = *reinterpret_cast<const bf16x4 *>(&buf[cur][lds_at(16 * i + (lane & 15), kk, stride)]);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((uint32_t)j < col_blocks)
                    bv[j] = *reinterpret_cast<const bf16x4 *>(&W[(size_t)(colbase + 16 * j + (lane & 15)) * K + kk]);
#pragma unroll
            for (int i = 0; i < ROWS; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if ((uint32_t)j < col_blocks) acc[i][j] = mfma(av[i], bv[j], acc[i][j]);
        }
        const uint32_t nxt = cur ^ 1;
#pragma unroll
        for (int i = 0; i < ROWS; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((uint32_t)j < col_blocks)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint32_t r = 16 * i + 4 * (lane >> 4) + e, c = colbase + 16 * j + (lane & 15);
                    const float zv = acc[i][j][e] + bias[c];
                    const bf16 h = to_bf16(silu(zv));
                    buf[nxt][lds_at(r, c, stride)] = h;
                    if (a.train) {
                        a.z[((size_t)l * a.batch + row0 + r) * H + c] = zv;
                        a.ht[(size_t)(a.in_pad + (size_t)l * H + c) * a.batch + row0 + r] = h;
                    }
                }
        __syncthreads();
        cur = nxt;
    }
    // output unit: y = sum_k h[k] w_out[k] + b_out, TPP threads per point
    constexpr uint32_t TPP = 256 / FT;
    const uint32_t r = tid / TPP, part = tid % TPP, span = H / TPP;
    const float *wout = a.params + a.wout_off;
    float s = 0.0f;
    for (uint32_t k = part * span; k < (part + 1) * span; k++) s += from_bf16(buf[cur][lds_at(r, k, stride)]) * wout[k];
#pragma unroll
    for (uint32_t m = 1; m < TPP; m *= 2) s += __shfl_xor(s, m);
    if (part == 0) a.y[row0 + r] = s + a.params[a.bout_off];
}

// ---- inference at hidden = 256 -----------------------------------------------------------------
//
// 128 points per workgroup, every layer in one pass. The products are computed transposed,
// Z^T = W H^T (A = W: 16 output units per block, B = H: 16 points per block), so that a lane's
// four results are four consecutive units of one point: bias, SiLU and the bf16 rounding happen
// in registers and one 8-byte LDS store writes them. Wavefront w owns units 64w .. 64w+63 of all
// 128 points: its weight slice (64 x 256 bf16) sits in registers for the whole layer and the
// activations stream from LDS (two 128 x 256 bf16 buffers, 16-byte granules XOR-swizzled by
// row). Layers >= 1 use v_mfma_f32_16x16x32_bf16 (lane l: A row / B column l % 16, k = 8 (l / 16)
// .. +7, eight bf16 each); the first layer (K = in_pad = 16) uses the 16x16x16 form.

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr uint32_t ITILE = 128;  // points per inference workgroup
constexpr uint32_t IH = 256;     // hidden units handled by the inference kernel
constexpr uint32_t MAX_LAYERS = 16;
constexpr uint32_t MAX_SPLIT = 16;  // split-K ways of the weight-gradient products

struct InferArgs {
    const float *x;
    float *y;
    uint32_t batch, inputs, layers;
    const bf16 *wb;
    const float *params;
    uint64_t wb_off[MAX_LAYERS], b_off[MAX_LAYERS];
    uint64_t wout_off, bout_off;
};

__device__ __forceinline__ uint32_t swz(uint32_t r, uint32_t c) {
    return r * IH + ((((c >> 3) ^ (r & 15))) << 3) + (c & 7);
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// bias, SiLU and bf16 of acc[j][i] (units 64w + 16j + 4(l/16) + e, point 16i + l%16) into `out`
__device__ __forceinline__ void infer_epilogue(f32x4 (&acc)[4][8], const float *bias, bf16 *out, uint32_t wave,
                                               uint32_t lane) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t c = 64 * wave + 16 * j + 4 * (lane >> 4);
        const float4 b = *reinterpret_cast<const float4 *>(bias + c);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t r = 16 * i + (lane & 15);
            const uint32_t h0 = to_bf16(silu(acc[j][i][0] + b.x)), h1 = to_bf16(silu(acc[j][i][1] + b.y));
            const uint32_t h2 = to_bf16(silu(acc[j][i][2] + b.z)), h3 = to_bf16(silu(acc[j][i][3] + b.w));
            uint2 v;
            v.x = h0 | (h1 << 16);
            v.y = h2 | (h3 << 16);
            *reinterpret_cast<uint2 *>(&out[swz(r, c)]) = v;
        }
    }
}

__global__ void __launch_bounds__(256, 1) mlp_infer(InferArgs a) {
    extern __shared__ bf16 lds[];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t row0 = blockIdx.x * ITILE;
    bf16 *const base = lds;
    auto act = [base](uint32_t k) { return base + k * (ITILE * IH); };
    // the input tile in bf16, columns 0..15 (zero beyond the inputs and beyond the batch)
    for (uint32_t idx = tid; idx < ITILE * 16; idx += blockDim.x) {
        const uint32_t r = idx >> 4, c = idx & 15;
        const float v = (c < a.inputs && row0 + r < a.batch) ? a.x[(size_t)(row0 + r) * a.inputs + c] : 0.0f;
        act(0)[swz(r, c)] = to_bf16(v);
    }
    __syncthreads();
    f32x4 acc[4][8];
    {  // layer 0: K = 16
        const bf16 *W = a.wb + a.wb_off[0];
        bf16x4 w0[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            w0[j] = *reinterpret_cast<const bf16x4 *>(&W[(64 * wave + 16 * j + (lane & 15)) * 16 + 4 * (lane >> 4)]);
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int i = 0; i < 8; i++) acc[j][i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const bf16x4 h = *reinterpret_cast<const bf16x4 *>(&act(0)[swz(16 * i + (lane & 15), 4 * (lane >> 4))]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[j][i] = mfma(w0[j], h, acc[j][i]);
        }
    }
    bf16x8 wf[4][8];
    auto load_weights = [&](uint32_t l) {
        const bf16 *W = a.wb + a.wb_off[l];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int s = 0; s < 8; s++)
                wf[j][s] = *reinterpret_cast<const bf16x8 *>(&W[(size_t)(64 * wave + 16 * j + (lane & 15)) * IH + 32 * s +
                                                               8 * (lane >> 4)]);
    };
    if (a.layers > 1) load_weights(1);
    infer_epilogue(acc, a.params + a.b_off[0], act(1), wave, lane);
    __syncthreads();
    uint32_t cur = 1;
    for (uint32_t l = 1; l < a.layers; l++) {
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int i = 0; i < 8; i++) acc[j][i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 8; s++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const bf16x8 h =
                    *reinterpret_cast<const bf16x8 *>(&act(cur)[swz(16 * i + (lane & 15), 32 * s + 8 * (lane >> 4))]);
#pragma unroll
                for (int j = 0; j < 4; j++) acc[j][i] = mfma32(wf[j][s], h, acc[j][i]);
            }
        if (l + 1 < a.layers) load_weights(l + 1);
        infer_epilogue(acc, a.params + a.b_off[l], act(cur ^ 1), wave, lane);
        __syncthreads();
        cur ^= 1;
    }
    // output unit: two threads per point, 128 units each, then one shuffle
    const uint32_t r = tid >> 1, part = tid & 1;
    const float *wout = a.params + a.wout_off;
    float s = 0.0f;
    for (uint32_t k = part * 128; k < part * 128 + 128; k++) s += from_bf16(act(cur)[swz(r, k)]) * wout[k];
    s += __shfl_xor(s, 1);
    if (part == 0 && row0 + r < a.batch) a.y[row0 + r] = s + a.params[a.bout_off];
}

// ---- backward --------------------------------------------------------------------------------

// delta = 2 (y - t) / batch; loss += (y - t)^2 / batch; dZ_last = delta w_out silu'(Z_last).
// One workgroup per TILE points: dZ rows are written along the units, and their transpose goes
// through LDS so that dZ^T rows are written along the points; one atomic add of the tile's loss.
__global__ void __launch_bounds__(256) mlp_output_grad(const float *y, const float *t, uint32_t batch, uint32_t hidden,
                                                       const float *z_last, const float *wout, float *delta, float *loss,
                                                       bf16 *dz, bf16 *dzt) {
    __shared__ bf16 tile[TILE][256 + 2];  // hidden <= 256; the pad spreads a column over the banks
    __shared__ float g_s[TILE], l_s[TILE];
    const uint32_t b0 = blockIdx.x * TILE;
    if (threadIdx.x < TILE) {
        const uint32_t b = b0 + threadIdx.x;
        const float d = y[b] - t[b];
        const float g = 2.0f * d / (float)batch;
        delta[b] = g;
        g_s[threadIdx.x] = g;
        l_s[threadIdx.x] = d * d / (float)batch;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float sum = 0.0f;
        for (uint32_t r = 0; r < TILE; r++) sum += l_s[r];
        atomicAdd(loss, sum);
    }
    for (uint32_t idx = threadIdx.x; idx < TILE * hidden; idx += blockDim.x) {
        const uint32_t r = idx / hidden, n = idx % hidden;
        const float v = g_s[r] * wout[n] * silu_grad(z_last[(size_t)(b0 + r) * hidden + n]);
        const bf16 h = to_bf16(v);
        dz[(size_t)(b0 + r) * hidden + n] = h;
        tile[r][n] = h;
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < TILE * hidden; idx += blockDim.x) {
        const uint32_t n = idx / TILE, r = idx % TILE;
        dzt[(size_t)n * batch + b0 + r] = tile[r][n];
    }
}

// out[n] = sum_b a[n][b] * w[b] (rows of length batch), or sum_b a[n][b] when w is NULL; one
// workgroup per n (a fixed-order tree over the batch)
template <class T>
__global__ void __launch_bounds__(256) row_dot(const T *a, const float *w, uint32_t batch, float *out) {
    __shared__ float part[256];
    const uint32_t n = blockIdx.x;
    float s = 0.0f;
    for (uint32_t b = threadIdx.x; b < batch; b += blockDim.x) {
        float v;
        if constexpr (__is_same(T, bf16)) v = from_bf16(a[(size_t)n * batch + b]);
        else v = a[(size_t)n * batch + b];
        s += w ? v * w[b] : v;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = part[0];
}

// C[m][n] = sum_k A[m][k] B[n][k], one wavefront per 64x64 tile of C.
// mode 0: C fp32 to out (ld = N).
// mode 1: the SiLU derivative: v = C * silu'(z[m][n]) -> dz (bf16 [m][n]) and dzt ([n][m]).
// blockIdx.z takes k in [z * k_split, (z + 1) * k_split) (a multiple of 16); mode 0 writes its
// partial product to out + z M N (summed in order by sum_splits)
__global__ void __launch_bounds__(64) gemm_nt(const bf16 *A, const bf16 *B, uint32_t M, uint32_t N, uint32_t K,
                                              uint32_t k_split, uint32_t mode, float *out, const float *z, bf16 *dz,
                                              bf16 *dzt) {
    const uint32_t lane = threadIdx.x;
    const uint32_t tm = blockIdx.x * TILE, tn = blockIdx.y * TILE;
    const uint32_t nb = std::min<uint32_t>(4, (N - tn) / 16);
    const uint32_t k_begin = blockIdx.z * k_split, k_end = std::min(K, k_begin + k_split);
    if (mode == 0) out += (size_t)blockIdx.z * M * N;
    // constant loop bounds with predicates (the accumulators stay in registers); the operands of
    // four k-steps are loaded before their products when the range allows
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    uint32_t k0 = k_begin;
    for (; k0 + 64 <= k_end; k0 += 64) {
        bf16x4 av[4][4], bv[4][4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t kk = k0 + 16 * q + 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 4; i++)
                av[q][i] = *reinterpret_cast<const bf16x4 *>(&A[(size_t)(tm + 16 * i + (lane & 15)) * K + kk]);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((uint32_t)j < nb)
                    bv[q][j] = *reinterpret_cast<const bf16x4 *>(&B[(size_t)(tn + 16 * j + (lane & 15)) * K + kk]);
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if ((uint32_t)j < nb) acc[i][j] = mfma(av[q][i], bv[q][j], acc[i][j]);
    }
    for (; k0 < k_end; k0 += 16) {
        const uint32_t kk = k0 + 4 * (lane >> 4);
        bf16x4 av[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) av[i] = *reinterpret_cast<const bf16x4 *>(&A[(size_t)(tm + 16 * i + (lane & 15)) * K + kk]);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((uint32_t)j < nb) bv[j] = *reinterpret_cast<const bf16x4 *>(&B[(size_t)(tn + 16 * j + (lane & 15)) * K + kk]);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((uint32_t)j < nb) acc[i][j] = mfma(av[i], bv[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((uint32_t)j < nb)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint32_t r = tm + 16 * i + 4 * (lane >> 4) + e, c = tn + 16 * j + (lane & 15);
                const float v = acc[i][j][e];
                if (mode == 0) {
                    out[(size_t)r * N + c] = v;
                } else {
                    const bf16 h = to_bf16(v * silu_grad(z[(size_t)r * N + c]));
                    dz[(size_t)r * N + c] = h;
                    dzt[(size_t)c * M + r] = h;
                }
            }
}

// out[i] = sum over z = 0 .. splits-1 of partial[z][i], in that order
__global__ void sum_splits(const float *partial, uint32_t splits, uint32_t count, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float v = partial[i];
    for (uint32_t z = 1; z < splits; z++) v += partial[(size_t)z * count + i];
    out[i] = v;
}

// the first layer's weight gradient drops the padded input columns: dst[n][c] = src[n][c], c < inputs
__global__ void unpad_columns(const float *src, uint32_t rows, uint32_t in_pad, uint32_t inputs, float *dst) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * inputs) return;
    const uint32_t n = idx / inputs, c = idx % inputs;
    dst[idx] = src[(size_t)n * in_pad + c];
}

__global__ void adam(float *p, const float *g, float *m, float *v, uint64_t count, float lr, float b1, float b2, float eps,
                     float c1, float c2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of a layer's weights: wb [out][in_pad] (zero padded) and, if wtb, wtb [in][out]
__global__ void refresh_layer(const float *w, uint32_t out, uint32_t in, uint32_t in_pad, bf16 *wb, bf16 *wtb) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= out * in_pad) return;
    const uint32_t o = idx / in_pad, c = idx % in_pad;
    const bf16 h = to_bf16(c < in ? w[(size_t)o * in + c] : 0.0f);
    wb[idx] = h;
    if (wtb && c < in) wtb[(size_t)c * out + o] = h;
}

}  // namespace

struct pacuda_mlp {
    Net n;
    uint64_t *d_wb_off = nullptr, *d_b_off = nullptr;
    float *scratch_w = nullptr;  // the first layer's padded weight gradient
    float *partial = nullptr;    // split-K partial weight gradients, [MAX_SPLIT][hidden][fan_in_pad]
};

namespace {

hipStream_t mlp_stream() { return runner_stream(); }

bool refresh(pacuda_mlp *m) {
    Net &n = m->n;
    hipStream_t s = mlp_stream();
    for (uint32_t l = 0; l < n.layers; l++) {
        const uint32_t in = l == 0 ? n.inputs : n.hidden, in_pad = l == 0 ? n.in_pad : n.hidden;
        const uint32_t total = n.hidden * in_pad;
        refresh_layer<<<(total + 255) / 256, 256, 0, s>>>(n.params + n.w_off[l], n.hidden, in, in_pad, n.wb + n.wb_off[l],
                                                          l == 0 ? nullptr : n.wtb + n.wtb_off[l]);
    }
    return !hip_failed(hipGetLastError(), "mlp refresh");
}

}  // namespace

extern "C" int pacuda_mlp_create(uint32_t inputs, uint32_t hidden, uint32_t layers, uint32_t max_batch, pacuda_mlp **out) {
    if (!out) return PACMAN_RECURRENCE_INVALID;
    *out = nullptr;
    if (inputs == 0 || inputs > 256 || (hidden != 64 && hidden != 128 && hidden != 256) || layers == 0 || layers > 16 ||
        max_batch == 0 || max_batch % TILE)
        return PACMAN_RECURRENCE_INVALID;
    pacuda_mlp *m = new pacuda_mlp();
    Net &n = m->n;
    n.inputs = inputs;
    n.in_pad = (inputs + 15) / 16 * 16;
    n.hidden = hidden;
    n.layers = layers;
    n.max_batch = max_batch;
    uint64_t off = 0, wb = 0, wtb = 0;
    for (uint32_t l = 0; l < layers; l++) {
        const uint32_t in = l == 0 ? inputs : hidden, in_pad = l == 0 ? n.in_pad : hidden;
        n.w_off.push_back(off);
        off += (uint64_t)hidden * in;
        n.b_off.push_back(off);
        off += hidden;
        n.wb_off.push_back(wb);
        wb += (uint64_t)hidden * in_pad;
        n.wtb_off.push_back(wtb);
        if (l > 0) wtb += (uint64_t)hidden * hidden;
    }
    n.wout_off = off;
    off += hidden;
    n.bout_off = off;
    off += 1;
    n.count = off;
    const size_t B = max_batch;
    const size_t ht_rows = n.in_pad + (size_t)layers * hidden;
    bool ok = !hip_failed(hipMalloc((void **)&n.params, off * 4), "mlp") && !hip_failed(hipMalloc((void **)&n.grads, off * 4), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.m, off * 4), "mlp") && !hip_failed(hipMalloc((void **)&n.v, off * 4), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.wb, std::max<uint64_t>(wb, 1) * 2), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.wtb, std::max<uint64_t>(wtb, 1) * 2), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.z, (size_t)layers * B * hidden * 4), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.ht, ht_rows * B * 2), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.dz[0], B * hidden * 2), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.dz[1], B * hidden * 2), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.dzt, B * hidden * 2), "mlp") &&
              !hip_failed(hipMalloc((void **)&n.delta, B * 4), "mlp") && !hip_failed(hipMalloc((void **)&n.loss, 64), "mlp") &&
              !hip_failed(hipMalloc((void **)&m->d_wb_off, layers * 8), "mlp") &&
              !hip_failed(hipMalloc((void **)&m->d_b_off, layers * 8), "mlp") &&
              !hip_failed(hipMalloc((void **)&m->scratch_w, (size_t)hidden * n.in_pad * 4), "mlp") &&
              !hip_failed(hipMalloc((void **)&m->partial, (size_t)MAX_SPLIT * hidden * std::max(hidden, n.in_pad) * 4),
                          "mlp") &&
              !hip_failed(hipMemcpy(m->d_wb_off, n.wb_off.data(), layers * 8, hipMemcpyHostToDevice), "mlp") &&
              !hip_failed(hipMemcpy(m->d_b_off, n.b_off.data(), layers * 8, hipMemcpyHostToDevice), "mlp") &&
              !hip_failed(hipMemset(n.params, 0, off * 4), "mlp") && !hip_failed(hipMemset(n.m, 0, off * 4), "mlp") &&
              !hip_failed(hipMemset(n.v, 0, off * 4), "mlp") && !hip_failed(hipMemset(n.grads, 0, off * 4), "mlp");
    if (!ok) {
        pacuda_mlp_destroy(m);
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    }
    *out = m;
    return PACMAN_RECURRENCE_OK;
}

extern "C" void pacuda_mlp_destroy(pacuda_mlp *m) {
    if (!m) return;
    (void)hipDeviceSynchronize();
    Net &n = m->n;
    for (void *p : {(void *)n.params, (void *)n.grads, (void *)n.m, (void *)n.v, (void *)n.wb, (void *)n.wtb, (void *)n.z,
                    (void *)n.ht, (void *)n.dz[0], (void *)n.dz[1], (void *)n.dzt, (void *)n.delta, (void *)n.loss, (void *)m->d_wb_off,
                    (void *)m->d_b_off, (void *)m->scratch_w, (void *)m->partial})
        if (p) (void)hipFree(p);
    delete m;
}

extern "C" uint64_t pacuda_mlp_param_count(const pacuda_mlp *m) { return m ? m->n.count : 0; }

extern "C" int pacuda_mlp_set_params(pacuda_mlp *m, const float *host, uint64_t count) {
    if (!m || !host || count != m->n.count) return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    if (hip_failed(hipMemcpy(n.params, host, count * 4, hipMemcpyHostToDevice), "mlp") ||
        hip_failed(hipMemset(n.m, 0, count * 4), "mlp") || hip_failed(hipMemset(n.v, 0, count * 4), "mlp") || !refresh(m) ||
        hip_failed(hipStreamSynchronize(mlp_stream()), "mlp"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}

static int copy_out(const float *src, float *host, uint64_t count) {
    if (hip_failed(hipStreamSynchronize(mlp_stream()), "mlp") || hip_failed(hipMemcpy(host, src, count * 4, hipMemcpyDeviceToHost), "mlp"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_mlp_get_params(const pacuda_mlp *m, float *host, uint64_t count) {
    if (!m || !host || count != m->n.count) return PACMAN_RECURRENCE_INVALID;
    return copy_out(m->n.params, host, count);
}

extern "C" int pacuda_mlp_get_grads(const pacuda_mlp *m, float *host, uint64_t count) {
    if (!m || !host || count != m->n.count) return PACMAN_RECURRENCE_INVALID;
    return copy_out(m->n.grads, host, count);
}

extern "C" int pacuda_mlp_forward(pacuda_mlp *m, const float *x, uint32_t batch, float *y, uint32_t train) {
    if (!m || !x || !y || batch == 0 || batch % TILE || (train && batch > m->n.max_batch)) return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    FwdArgs a;
    a.x = x; a.y = y; a.batch = batch; a.inputs = n.inputs; a.in_pad = n.in_pad; a.hidden = n.hidden; a.layers = n.layers;
    a.wb = n.wb; a.wb_off = m->d_wb_off; a.params = n.params; a.b_off = m->d_b_off;
    a.wout_off = n.wout_off; a.bout_off = n.bout_off;
    a.z = n.z; a.ht = n.ht; a.train = train ? 1 : 0;
    const size_t lds = 2 * (size_t)TILE * n.hidden * sizeof(bf16);
    if (n.in_pad > n.hidden) return PACMAN_RECURRENCE_INVALID;
    if (!train && n.hidden == IH && n.in_pad == 16 && n.layers <= MAX_LAYERS) {
        InferArgs ia;
        ia.x = x; ia.y = y; ia.batch = batch; ia.inputs = n.inputs; ia.layers = n.layers;
        ia.wb = n.wb; ia.params = n.params;
        for (uint32_t l = 0; l < n.layers; l++) {
            ia.wb_off[l] = n.wb_off[l];
            ia.b_off[l] = n.b_off[l];
        }
        ia.wout_off = n.wout_off; ia.bout_off = n.bout_off;
        static const bool lds_ok = !hip_failed(
            hipFuncSetAttribute(reinterpret_cast<const void *>(mlp_infer), hipFuncAttributeMaxDynamicSharedMemorySize,
                                2 * ITILE * IH * sizeof(bf16)),
            "mlp inference");
        if (!lds_ok) return PACMAN_RECURRENCE_DEVICE_ERROR;
        mlp_infer<<<(batch + ITILE - 1) / ITILE, 256, 2 * ITILE * IH * sizeof(bf16), mlp_stream()>>>(ia);
        return hip_failed(hipGetLastError(), "mlp inference") ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK;
    }
    // workgroups of 64 points when that gives two per compute unit, else of 16
    if (batch / TILE >= 512) {
        mlp_forward<4><<<batch / TILE, 256, lds, mlp_stream()>>>(a);
    } else {
        mlp_forward<1><<<batch / 16, 256, lds / 4, mlp_stream()>>>(a);
    }
    if (hip_failed(hipGetLastError(), "mlp forward")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_mlp_backward(pacuda_mlp *m, const float *y, const float *target, uint32_t batch, float *loss) {
    if (!m || !y || !target || batch == 0 || batch % TILE || batch > m->n.max_batch) return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    hipStream_t s = mlp_stream();
    const uint32_t H = n.hidden, L = n.layers;
    const size_t B = batch;
    auto htrow = [&](int l) { return n.ht + (l < 0 ? 0 : (n.in_pad + (size_t)l * H) * B); };  // H_l^T rows (l = -1: x^T)
    if (hip_failed(hipMemsetAsync(n.loss, 0, 4, s), "mlp")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    mlp_output_grad<<<batch / TILE, 256, 0, s>>>(y, target, batch, H, n.z + (size_t)(L - 1) * B * H, n.params + n.wout_off,
                                          n.delta, n.loss, n.dz[0], n.dzt);
    // output layer: dW_out[k] = sum_b delta[b] H_last[b][k]; db_out = sum_b delta[b]
    row_dot<bf16><<<H, 256, 0, s>>>(htrow(L - 1), n.delta, batch, n.grads + n.wout_off);
    row_dot<float><<<1, 256, 0, s>>>(n.delta, nullptr, batch, n.grads + n.bout_off);
    int cur = 0;
    for (int l = (int)L - 1; l >= 0; l--) {
        const uint32_t K_in = l == 0 ? n.in_pad : H;
        // dW_l[n][k] = sum_b dZ_l^T[n][b] H_(l-1)^T[k][b]
        float *dw = l == 0 ? m->scratch_w : n.grads + n.w_off[l];
        // split K (= batch) so that the few output tiles still fill the GPU; partials summed in order
        const uint32_t splits = std::max(1u, std::min(MAX_SPLIT, batch / 256));
        const uint32_t k_split = (batch / splits + 15) / 16 * 16;
        const uint32_t used = (batch + k_split - 1) / k_split;
        dim3 g1(H / TILE, (K_in + TILE - 1) / TILE, used);
        gemm_nt<<<g1, 64, 0, s>>>(n.dzt, htrow(l - 1), H, K_in, batch, k_split, 0, m->partial, nullptr, nullptr, nullptr);
        sum_splits<<<(H * K_in + 255) / 256, 256, 0, s>>>(m->partial, used, H * K_in, dw);
        if (l == 0) {
            const uint32_t total = H * n.inputs;
            unpad_columns<<<(total + 255) / 256, 256, 0, s>>>(m->scratch_w, H, n.in_pad, n.inputs, n.grads + n.w_off[0]);
        }
        // db_l[n] = sum_b dZ_l[b][n]
        row_dot<bf16><<<H, 256, 0, s>>>(n.dzt, nullptr, batch, n.grads + n.b_off[l]);
        if (l > 0) {
            // dZ_(l-1) = (dZ_l W_l) * silu'(Z_(l-1)): A = dZ_l [b][n], B = W_l^T [k][n]
            dim3 g2(batch / TILE, H / TILE);
            gemm_nt<<<g2, 64, 0, s>>>(n.dz[cur], n.wtb + n.wtb_off[l], batch, H, H, H, 1, nullptr,
                                      n.z + (size_t)(l - 1) * B * H, n.dz[cur ^ 1], n.dzt);
            cur ^= 1;
        }
    }
    if (hip_failed(hipGetLastError(), "mlp backward")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (loss && (hip_failed(hipMemcpyAsync(loss, n.loss, 4, hipMemcpyDeviceToHost, s), "mlp") ||
                 hip_failed(hipStreamSynchronize(s), "mlp")))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_mlp_adam(pacuda_mlp *m, float lr, float beta1, float beta2, float eps, uint32_t step) {
    if (!m || step == 0) return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    adam<<<(unsigned)((n.count + 255) / 256), 256, 0, mlp_stream()>>>(n.params, n.grads, n.m, n.v, n.count, lr, beta1, beta2,
                                                                      eps, c1, c2);
    if (hip_failed(hipGetLastError(), "mlp adam") || !refresh(m)) return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}
