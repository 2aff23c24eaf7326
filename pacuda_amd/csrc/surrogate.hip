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
// Training runs layer by layer on one product kernel (gemm_tile: NT products, 64 x 64 tiles of C
// per workgroup of four wavefronts, K staged through LDS, v_mfma_f32_16x16x32_bf16) with the
// layer's elementwise work fused into its epilogue: forward Z_l = H_(l-1) W_l^T with bias, SiLU
// and the bf16 activations stored row-major (the next layer's operand) and transposed (the
// weight gradient's); backward (loss = mean of (y - t)^2) the output layer by per-tile
// reductions, then per hidden layer dW_l = dZ_l^T H_(l-1) (K = batch, split 16 ways, partial
// products and the bias gradient's row sums kept per split) and dZ_(l-1) = (dZ_l W_l)
// silu'(Z_(l-1)); one kernel sums every split in a fixed order into the gradients, and one does
// Adam and the bf16 weight copies. Inference keeps the activations in LDS (mlp_infer at hidden
// 256, mlp_forward otherwise).
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
    bf16 *hb;        // the same row-major: x [batch][in_pad], then H_l [batch][hidden]
    bf16 *dz[2], *dzt[2];  // [batch][hidden] and [hidden][batch] (two: a layer's dZ is read while the next is written)
    float *loss, *delta;
    // per-split partial gradients: dW_l [layer][split][hidden][kmax], db_l [layer][split][hidden];
    // per 64-point tile: the loss, and dW_out (hidden values) then db_out
    float *part, *dbpart, *loss_part, *head_part;
    uint32_t kmax;
    // the per-tile path (hidden 256, at most 16 inputs): dZ_l^T of every layer [layers][hidden][batch],
    // the training loop's y
    bf16 *dztl;
    float *ytrain, *head_total;
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
};

// LDS position of activation (r, c) in a row of `stride` bf16: the 4-element granule index is
// XORed with r % 16, so the 16 rows one MFMA operand load touches fall on distinct banks
__device__ __forceinline__ uint32_t lds_at(uint32_t r, uint32_t c, uint32_t stride) {
    return r * stride + ((((c >> 2) ^ (r & 15))) << 2) + (c & 3);
}

// Inference at hidden 64 and 128: ROWS blocks of 16 points per workgroup (FT = 16 ROWS points)
// through every layer, the activations in LDS; small batches take ROWS = 1 so that they still
// spread over the compute units
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
        buf[0][lds_at(r, c, stride)] = to_bf16(v);
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
                    buf[nxt][lds_at(r, c, stride)] = to_bf16(silu(acc[i][j][e] + bias[c]));
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

// ---- training -----------------------------------------------------------------------------
//
// gemm_tile: C[m][n] = sum_k A[m][k] B[n][k] for row-major A (rows of lda) and B (rows of ldb),
// both contiguous along k; M a multiple of 64, N and K multiples of 16. A workgroup of four
// wavefronts takes a 64 x 64 tile of C, each wavefront a 32 x 32 quarter (2 x 2 blocks of
// v_mfma_f32_16x16x32_bf16). K goes through LDS in chunks of up to 256: both 64-row operand
// tiles (32 KB each) are loaded with coalesced 16-byte loads and stored with their 16-byte
// granules XOR-swizzled by row (the 16 rows one operand read touches fall on distinct banks);
// granules past K are zero. blockIdx.z takes k in [z k_split, (z + 1) k_split).
// Epilogues: E_PARTIAL stores C (fp32) to out + z M N; E_FORWARD stores Z = C + bias (fp32, to
// z) and H = bf16(SiLU(Z)) row-major (h, [M][N]) and transposed (ht, [N][M]); E_DZ stores
// bf16(C silu'(z[m][n])) to h and ht. With row_sums, the workgroups of the first column of
// tiles also store sum_k A[m][k] over their k range to row_sums[z M + m] (each row in a fixed
// order: four threads over interleaved granules, then two shuffles).
constexpr uint32_t GT = 64;      // tile side
constexpr uint32_t GK = 256;     // k per LDS chunk
constexpr uint32_t GG = GK / 8;  // 16-byte granules per LDS row
constexpr uint32_t GD = 4;       // chunks whose loads are issued together
enum { E_PARTIAL, E_FORWARD, E_DZ };

struct GemmArgs {
    const bf16 *A, *B;
    uint32_t M, N, K, lda, ldb, k_split, mode;
    float *out;
    const float *bias;
    float *z;
    bf16 *h, *ht;
    float *row_sums;
};

__device__ __forceinline__ uint32_t gran(uint32_t r, uint32_t g) { return r * GG + (g ^ (r & 15)); }

__device__ __forceinline__ float granule_sum(uint4 v) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        s += __uint_as_float(w[k] << 16);
        s += __uint_as_float(w[k] & 0xffff0000u);
    }
    return s;
}

// Two independent products may share a launch (the workgroups from `first1` on take the second),
// so that more workgroups overlap their latencies; each product's workgroups are numbered
// m-tile fastest, then n-tile, then k split.
struct GemmPair {
    GemmArgs g[2];
    uint32_t first1, tiles_m[2], tiles_n[2];
};

// (the products of one launch as a device table: product p takes workgroups first[p] ..
// first[p + 1] - 1)
// Workgroups from first[count] on reduce the output unit's per-tile partials instead: value v
// (v < hidden: dW_out[v], v = hidden: db_out, v = hidden + 1: the loss) by 16 threads, thread q
// adding tiles q, q + 16, ... in order, then a fixed four-step shuffle tree, into head_total[v].
struct GemmTable {
    GemmArgs g[MAX_LAYERS];
    uint32_t first[MAX_LAYERS + 1], tiles_m[MAX_LAYERS], tiles_n[MAX_LAYERS], count;
    const float *head_part, *loss_part;
    float *head_total;
    uint32_t tiles, hidden;
};

__device__ __forceinline__ void head_reduce(const GemmTable *t, uint32_t block) {
    const uint32_t v = block * 16 + (threadIdx.x >> 4), q = threadIdx.x & 15, H = t->hidden;
    float s = 0.0f;
    // (eight loads issued before their eight additions, which keep the order tile q, q + 16, ...:
    // the loads' latencies overlap instead of adding up)
    if (v <= H + 1)
        for (uint32_t base = q; base < t->tiles; base += 8 * 16) {
            float x[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) {
                const uint32_t tile = base + 16 * j;
                x[j] = tile >= t->tiles ? 0.0f
                       : v == H + 1    ? t->loss_part[tile]
                                       : t->head_part[(size_t)tile * (H + 1) + v];
            }
#pragma unroll
            for (uint32_t j = 0; j < 8; j++)
                if (base + 16 * j < t->tiles) s += x[j];
        }
#pragma unroll
    for (uint32_t m = 1; m < 16; m *= 2) s += __shfl_xor(s, m);
    if (q == 0 && v <= H + 1) t->head_total[v] = s;
}

__device__ __forceinline__ void gemm_body(const GemmArgs &a, uint32_t bx, uint32_t by, uint32_t bz, uint4 *sa,
                                          uint4 *sb);

// XCD-aware numbering: the hardware hands consecutive workgroups to the eight XCDs in turn, each
// with its own L2, so workgroup b takes logical index `xcd_major(b)`: the workgroups of one XCD
// get consecutive logical indices. The products number their workgroups m-tile fastest, then
// n-tile, then k split, so the tiles that share one k split's operand rows (a few hundred KB per
// layer) run on one XCD and read them once from HBM instead of once per XCD.
constexpr uint32_t XCDS = 8;
__device__ __forceinline__ uint32_t xcd_major(uint32_t b, uint32_t n) {
    const uint32_t q = n / XCDS, rem = n % XCDS, x = b % XCDS, local = b / XCDS;
    return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + local;
}

__global__ void __launch_bounds__(256) gemm_tile(GemmPair pair) {
    __shared__ uint4 sa[GT * GG], sb[GT * GG];
    const uint32_t block = xcd_major(blockIdx.x, gridDim.x);
    const bool second = block >= pair.first1;
    const GemmArgs a = second ? pair.g[1] : pair.g[0];
    const uint32_t local = block - (second ? pair.first1 : 0), tiles_m = pair.tiles_m[second],
                   tiles_n = pair.tiles_n[second];
    gemm_body(a, local % tiles_m, (local / tiles_m) % tiles_n, local / (tiles_m * tiles_n), sa, sb);
}

__global__ void __launch_bounds__(256) gemm_table(const GemmTable *table) {
    __shared__ uint4 sa[GT * GG], sb[GT * GG];
    const GemmTable *t = table;
    const uint32_t block = xcd_major(blockIdx.x, gridDim.x);
    if (block >= t->first[t->count]) {
        head_reduce(t, block - t->first[t->count]);
        return;
    }
    uint32_t p = 0;
    while (p + 1 < t->count && block >= t->first[p + 1]) p++;
    const GemmArgs a = t->g[p];
    const uint32_t local = block - t->first[p], tiles_m = t->tiles_m[p], tiles_n = t->tiles_n[p];
    gemm_body(a, local % tiles_m, (local / tiles_m) % tiles_n, local / (tiles_m * tiles_n), sa, sb);
}

__device__ __forceinline__ void gemm_body(const GemmArgs &a, uint32_t bx, uint32_t by, uint32_t bz, uint4 *sa,
                                          uint4 *sb) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t tm = bx * GT, tn = by * GT, wm = wave & 1, wn = wave >> 1;
    const uint32_t k_begin = bz * a.k_split, k_end = min(a.K, k_begin + a.k_split);
    const bool sums = a.row_sums && by == 0;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    float rs = 0.0f;
    // staging: thread t loads row t / 4, granules t % 4 + 4 q (q < gs / 4), every load of a chunk
    // issued before the first LDS store
    const uint32_t sr = tid >> 2, sg = tid & 3;
    const bool b_row = tn + sr < a.N;
    const bf16 *ap = a.A + (size_t)(tm + sr) * a.lda, *bp = a.B + (size_t)(tn + sr) * a.ldb;
    // The loads of up to GD chunks are issued together, into registers, before the first is stored
    // to LDS: a workgroup waits one round trip to memory per GD chunks rather than per chunk (the
    // products of a chunk take far less time than a load's latency)
    uint4 ra[GD][GG / 4], rb[GD][GG / 4];
    for (uint32_t kg = k_begin; kg < k_end; kg += GD * GK) {
#pragma unroll
        for (uint32_t d = 0; d < GD; d++) {
            const uint32_t k0 = kg + d * GK;
            const uint32_t gs = k0 < k_end ? (min(GK, k_end - k0) + 31) / 32 * 4 : 0;
#pragma unroll
            for (uint32_t q = 0; q < GG / 4; q++) {
                const uint32_t g = sg + 4 * q, k = k0 + 8 * g;
                ra[d][q] = rb[d][q] = uint4{0u, 0u, 0u, 0u};
                if (4 * q < gs && k < k_end) {
                    ra[d][q] = *reinterpret_cast<const uint4 *>(ap + k);
                    if (b_row) rb[d][q] = *reinterpret_cast<const uint4 *>(bp + k);
                }
            }
        }
#pragma unroll
        for (uint32_t d = 0; d < GD; d++) {
        const uint32_t k0 = kg + d * GK;
        if (k0 >= k_end) break;
        const uint32_t kc = min(GK, k_end - k0), gs = (kc + 31) / 32 * 4;  // granules staged per row
        __syncthreads();  // (the previous chunk's reads are done)
#pragma unroll
        for (uint32_t q = 0; q < GG / 4; q++)
            if (4 * q < gs) {
                sa[gran(sr, sg + 4 * q)] = ra[d][q];
                sb[gran(sr, sg + 4 * q)] = rb[d][q];
            }
        __syncthreads();
        if (sums)
            for (uint32_t g = tid & 3; g < gs; g += 4) rs += granule_sum(sa[gran(tid >> 2, g)]);
        for (uint32_t st = 0; st < gs / 4; st++) {
            bf16x8 av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
                av[i] = __builtin_bit_cast(bf16x8, sa[gran(32 * wm + 16 * i + (lane & 15), 4 * st + (lane >> 4))]);
#pragma unroll
            for (int j = 0; j < 2; j++)
                bv[j] = __builtin_bit_cast(bf16x8, sb[gran(32 * wn + 16 * j + (lane & 15), 4 * st + (lane >> 4))]);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = mfma32(av[i], bv[j], acc[i][j]);
        }
    }
    }
    if (sums) {
        rs += __shfl_xor(rs, 1);
        rs += __shfl_xor(rs, 2);
        if ((tid & 3) == 0) a.row_sums[(size_t)bz * a.M + tm + (tid >> 2)] = rs;
    }
    // the epilogue through LDS: the tile of C (fp32) into the first operand buffer, then rows of
    // four values per thread (16-byte stores along n); E_FORWARD and E_DZ also keep their bf16
    // results in the second buffer and store its columns, eight values per thread, along m
    __syncthreads();  // (every wavefront's operand reads are done)
    float *ct = reinterpret_cast<float *>(sa);           // [64][CT_PITCH]
    bf16 *hs = reinterpret_cast<bf16 *>(sb);             // [64][HS_PITCH], transposed: hs[n][m]
    constexpr uint32_t CT_PITCH = GT + 4, HS_PITCH = GT + 8;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                ct[(32 * wm + 16 * i + 4 * (lane >> 4) + e) * CT_PITCH + 32 * wn + 16 * j + (lane & 15)] = acc[i][j][e];
    __syncthreads();
    const uint32_t n_valid = min(GT, a.N - tn);
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
        const uint32_t idx = tid + 256 * q, r = idx >> 4, c4 = 4 * (idx & 15);
        if (c4 >= n_valid) continue;
        const float4 v = *reinterpret_cast<const float4 *>(&ct[r * CT_PITCH + c4]);
        const size_t at = (size_t)(tm + r) * a.N + tn + c4;
        if (a.mode == E_PARTIAL) {
            *reinterpret_cast<float4 *>(a.out + (size_t)bz * a.M * a.N + at) = v;
            continue;
        }
        float o[4] = {v.x, v.y, v.z, v.w};
        if (a.mode == E_FORWARD) {
            const float4 b = *reinterpret_cast<const float4 *>(a.bias + tn + c4);
            const float zv[4] = {v.x + b.x, v.y + b.y, v.z + b.z, v.w + b.w};
            *reinterpret_cast<float4 *>(a.z + at) = float4{zv[0], zv[1], zv[2], zv[3]};
#pragma unroll
            for (int e = 0; e < 4; e++) o[e] = silu(zv[e]);
        } else {
            const float4 zz = *reinterpret_cast<const float4 *>(a.z + at);
            const float zv[4] = {zz.x, zz.y, zz.z, zz.w};
#pragma unroll
            for (int e = 0; e < 4; e++) o[e] = o[e] * silu_grad(zv[e]);
        }
        uint32_t h[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            h[e] = to_bf16(o[e]);
            hs[(c4 + e) * HS_PITCH + r] = (bf16)h[e];
        }
        *reinterpret_cast<uint2 *>(a.h + at) = uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
    }
    if (a.mode == E_PARTIAL) return;
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < 2; q++) {
        const uint32_t idx = tid + 256 * q, c = idx >> 3, r8 = 8 * (idx & 7);
        if (c >= n_valid) continue;
        *reinterpret_cast<uint4 *>(a.ht + (size_t)(tn + c) * a.M + tm + r8) =
            *reinterpret_cast<const uint4 *>(&hs[c * HS_PITCH + r8]);
    }
}

// the training input: x (fp32 [batch][inputs]) as bf16, zero padded to in_pad columns, row-major
// (hx) and transposed (htx)
__global__ void mlp_input(const float *x, uint32_t batch, uint32_t inputs, uint32_t in_pad, bf16 *hx, bf16 *htx) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= batch * in_pad) return;
    const uint32_t r = idx / in_pad, c = idx % in_pad;
    const bf16 h = to_bf16(c < inputs ? x[(size_t)r * inputs + c] : 0.0f);
    hx[idx] = h;
    htx[(size_t)c * batch + r] = h;
}

// y = H_last w_out + b_out of point p by four threads (q = 0..3 each over a quarter of the units,
// in 8-unit granules), then two shuffles; every thread returns y
__device__ __forceinline__ float output_point(const bf16 *h, uint32_t hidden, const float *wout, const float *bout,
                                              uint32_t p, uint32_t q) {
    const uint32_t span = hidden / 4;
    const uint4 *row = reinterpret_cast<const uint4 *>(h + (size_t)p * hidden + q * span);
    uint4 v[8];  // (span <= 64: eight granules)
#pragma unroll
    for (uint32_t g = 0; g < 8; g++)
        if (8 * g < span) v[g] = row[g];
    float s = 0.0f;
#pragma unroll
    for (uint32_t g = 0; g < 8; g++)
        if (8 * g < span) {
            const uint32_t w[4] = {v[g].x, v[g].y, v[g].z, v[g].w};
            const float *wo = wout + q * span + 8 * g;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                s += __uint_as_float(w[e] << 16) * wo[2 * e];
                s += __uint_as_float(w[e] & 0xffff0000u) * wo[2 * e + 1];
            }
        }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    return s + bout[0];
}

// y of 64 points per workgroup
__global__ void __launch_bounds__(256) mlp_output(const bf16 *h, uint32_t hidden, const float *wout, const float *bout,
                                                  float *y) {
    const uint32_t p = blockIdx.x * 64 + (threadIdx.x >> 2);
    const float v = output_point(h, hidden, wout, bout, p, threadIdx.x & 3);
    if ((threadIdx.x & 3) == 0) y[p] = v;
}

// The output layer's backward per 64-point tile: delta = 2 (y - t) / batch (to delta), the
// tile's loss sum((y - t)^2 / batch), dW_out[k] = sum delta H_last[.][k] and db_out = sum delta
// (into head_part[tile][hidden + 1]); every sum over the tile's points in point order.
// (fused: with y NULL the workgroup computes its points' y first, as mlp_output does)
__global__ void __launch_bounds__(256) mlp_head_grad(const float *y, const float *t, uint32_t batch, uint32_t hidden,
                                                     const bf16 *h_last, const float *wout, const float *bout,
                                                     float *delta, float *loss_part, float *head_part) {
    __shared__ float g_s[TILE], l_s[TILE], y_s[TILE];
    const uint32_t b0 = blockIdx.x * TILE;
    if (!y) {
        const uint32_t r = threadIdx.x >> 2;
        const float v = output_point(h_last, hidden, wout, bout, b0 + r, threadIdx.x & 3);
        if ((threadIdx.x & 3) == 0) y_s[r] = v;
        __syncthreads();
    }
    if (threadIdx.x < TILE) {
        const uint32_t b = b0 + threadIdx.x;
        const float d = (y ? y[b] : y_s[threadIdx.x]) - t[b];
        const float g = 2.0f * d / (float)batch;
        delta[b] = g;
        g_s[threadIdx.x] = g;
        l_s[threadIdx.x] = d * d / (float)batch;
    }
    __syncthreads();
    float *hp = head_part + (size_t)blockIdx.x * (hidden + 1);
    if (threadIdx.x == 0) {
        float sl = 0.0f, sg = 0.0f;
        for (uint32_t r = 0; r < TILE; r++) {
            sl += l_s[r];
            sg += g_s[r];
        }
        loss_part[blockIdx.x] = sl;
        hp[hidden] = sg;
    }
    for (uint32_t k = threadIdx.x; k < hidden; k += blockDim.x) {
        float hv[TILE];
#pragma unroll
        for (uint32_t r = 0; r < TILE; r++) hv[r] = from_bf16(h_last[(size_t)(b0 + r) * hidden + k]);
        float s = 0.0f;
#pragma unroll
        for (uint32_t r = 0; r < TILE; r++) s += g_s[r] * hv[r];
        hp[k] = s;
    }
}

// dZ_last = delta w_out silu'(Z_last) per 64 x 64 tile (points x units): row-major to dz and,
// through LDS, transposed to dzt
__global__ void __launch_bounds__(256) mlp_head_dz(const float *delta, uint32_t batch, uint32_t hidden,
                                                   const float *z_last, const float *wout, bf16 *dz, bf16 *dzt) {
    __shared__ bf16 tile[64][64 + 2];
    const uint32_t b0 = blockIdx.x * 64, n0 = blockIdx.y * 64, c = threadIdx.x & 63, r4 = threadIdx.x >> 6;
    const float w = wout[n0 + c];
#pragma unroll
    for (uint32_t q = 0; q < 16; q++) {
        const uint32_t r = r4 + 4 * q;
        const size_t at = (size_t)(b0 + r) * hidden + n0 + c;
        const bf16 h = to_bf16(delta[b0 + r] * w * silu_grad(z_last[at]));
        dz[at] = h;
        tile[r][c] = h;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < 16; q++) {
        const uint32_t n = r4 + 4 * q;
        dzt[(size_t)(n0 + n) * batch + b0 + c] = tile[c][n];
    }
}

// ---- training at hidden 256: per-tile forward and backward chains ---------------------------
//
// A point's forward pass and its backward chain dZ_(l-1) = (dZ_l W_l) silu'(Z_(l-1)) involve that
// point alone, so one workgroup takes PT points through every layer, the activations (or dZ) in
// LDS, as the inference kernel does: the products transposed (Z^T = W H^T, dZ_(l-1)^T = W_l^T
// dZ_l^T), wavefront w owning units 64w .. 64w+63, its slice of the layer's weights in registers
// (the next layer's slice loaded while the current epilogue runs), v_mfma_f32_16x16x32_bf16 with
// the first layer's 16 inputs zero-padded to 32. Only the weight gradients, sums over the batch,
// are left to gemm_tile.
constexpr uint32_t PT = 16;  // points per training tile (one workgroup per compute unit at batch 4096)

struct TileArgs {
    const float *x, *t;
    float *y;
    uint32_t batch, inputs, layers;
    const bf16 *wb, *wtb;
    const float *params;
    uint64_t wb_off[MAX_LAYERS], wtb_off[MAX_LAYERS], b_off[MAX_LAYERS];
    uint64_t wout_off, bout_off;
    float *z;        // [layers][batch][IH]
    bf16 *ht;        // x^T (16 rows) then H_l^T, rows of batch
    bf16 *dzt;       // dZ_l^T [layers][IH][batch]
    float *loss_part, *head_part;  // per tile
};

// wavefront w's slice of a 256 x 256 bf16 matrix (rows 64w + 16j + lane % 16, k = 32s + 8 (lane / 16))
__device__ __forceinline__ void load_slice(const bf16 *W, uint32_t wave, uint32_t lane, bf16x8 (&wf)[4][8]) {
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int s = 0; s < 8; s++)
            wf[j][s] = *reinterpret_cast<const bf16x8 *>(&W[(size_t)(64 * wave + 16 * j + (lane & 15)) * IH + 32 * s +
                                                           8 * (lane >> 4)]);
}

// acc[j][i] = (W slice) x (LDS rows 16i .. of `act`), k over `steps` 32-wide steps
__device__ __forceinline__ void tile_product(const bf16x8 (&wf)[4][8], const bf16 *act, uint32_t lane, int steps,
                                             f32x4 (&acc)[4][PT / 16]) {
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < (int)(PT / 16); i++) acc[j][i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; s++) {
        if (s >= steps) break;
#pragma unroll
        for (int i = 0; i < (int)(PT / 16); i++) {
            const bf16x8 h = *reinterpret_cast<const bf16x8 *>(&act[swz(16 * i + (lane & 15), 32 * s + 8 * (lane >> 4))]);
#pragma unroll
            for (int j = 0; j < 4; j++) acc[j][i] = mfma32(wf[j][s], h, acc[j][i]);
        }
    }
}

__global__ void __launch_bounds__(256) mlp_tile_forward(TileArgs a) {
    __shared__ bf16 act[2][PT * IH];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b0 = blockIdx.x * PT, B = a.batch;
    // the input tile in bf16, columns 0..31 (zero beyond the inputs), and x^T (16 rows)
    for (uint32_t idx = tid; idx < PT * 32; idx += blockDim.x) {
        const uint32_t r = idx >> 5, c = idx & 31;
        const bf16 h = to_bf16(c < a.inputs ? a.x[(size_t)(b0 + r) * a.inputs + c] : 0.0f);
        act[0][swz(r, c)] = h;
        if (c < 16) a.ht[(size_t)c * B + b0 + r] = h;
    }
    bf16x8 wf[4][8];
    {  // the first layer's weights, [unit][16]: k 16..31 are zero
        const bf16 *W = a.wb + a.wb_off[0];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            bf16x8 w;
#pragma unroll
            for (int e = 0; e < 8; e++) w[e] = (__bf16)0.0f;
            if ((lane >> 4) < 2) w = *reinterpret_cast<const bf16x8 *>(&W[(64 * wave + 16 * j + (lane & 15)) * 16 + 8 * (lane >> 4)]);
            wf[j][0] = w;
        }
    }
    __syncthreads();
    f32x4 acc[4][PT / 16];
    uint32_t cur = 0;
    for (uint32_t l = 0; l < a.layers; l++) {
        tile_product(wf, act[cur], lane, l == 0 ? 1 : 8, acc);
        if (l + 1 < a.layers) load_slice(a.wb + a.wb_off[l + 1], wave, lane, wf);
        // Z = acc + bias (fp32, [point][unit]), H = bf16(SiLU(Z)) into LDS (row-major) and H^T
        const float *bias = a.params + a.b_off[l];
        float *Z = a.z + (size_t)l * B * IH;
        bf16 *HT = a.ht + (size_t)(16 + l * IH) * B;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t c = 64 * wave + 16 * j + 4 * (lane >> 4);
            const float4 bv = *reinterpret_cast<const float4 *>(bias + c);
#pragma unroll
            for (int i = 0; i < (int)(PT / 16); i++) {
                const uint32_t r = 16 * i + (lane & 15);
                const float zv[4] = {acc[j][i][0] + bv.x, acc[j][i][1] + bv.y, acc[j][i][2] + bv.z, acc[j][i][3] + bv.w};
                *reinterpret_cast<float4 *>(&Z[(size_t)(b0 + r) * IH + c]) = float4{zv[0], zv[1], zv[2], zv[3]};
                uint32_t h[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    h[e] = to_bf16(silu(zv[e]));
                    HT[(size_t)(c + e) * B + b0 + r] = (bf16)h[e];
                }
                *reinterpret_cast<uint2 *>(&act[cur ^ 1][swz(r, c)]) = uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    // y = H_last w_out + b_out: TPP threads per point over IH / TPP units each, then shuffles
    {
        constexpr uint32_t TPP = 256 / PT, SPAN = IH / TPP;
        const uint32_t r = tid / TPP, part = tid % TPP;
        const float *wout = a.params + a.wout_off;
        float sum = 0.0f;
        for (uint32_t k = SPAN * part; k < SPAN * part + SPAN; k++) sum += from_bf16(act[cur][swz(r, k)]) * wout[k];
#pragma unroll
        for (uint32_t m = 1; m < TPP; m *= 2) sum += __shfl_xor(sum, m);
        if (part == 0) a.y[b0 + r] = sum + a.params[a.bout_off];
    }
}

// The backward chain of a tile: delta = 2 (y - t) / batch, the tile's loss and output-unit
// gradient partials (dW_out[k] = sum delta H_last[.][k] from H_last^T, db_out = sum delta, in
// point order), dZ_last = delta w_out silu'(Z_last), then dZ_(l-1) layer by layer; every dZ_l^T
// to dzt for the weight gradients.
__global__ void __launch_bounds__(256) mlp_tile_backward(TileArgs a) {
    __shared__ bf16 act[2][PT * IH];
    __shared__ float g_s[PT], l_s[PT];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, b0 = blockIdx.x * PT, B = a.batch;
    const uint32_t L = a.layers;
    if (tid < PT) {
        const float d = a.y[b0 + tid] - a.t[b0 + tid];
        g_s[tid] = 2.0f * d / (float)B;
        l_s[tid] = d * d / (float)B;
    }
    bf16x8 wf[4][8];
    if (L > 1) load_slice(a.wtb + a.wtb_off[L - 1], wave, lane, wf);
    __syncthreads();
    float *hp = a.head_part + (size_t)blockIdx.x * (IH + 1);
    if (tid == 0) {
        float sl = 0.0f, sg = 0.0f;
        for (uint32_t r = 0; r < PT; r++) {
            sl += l_s[r];
            sg += g_s[r];
        }
        a.loss_part[blockIdx.x] = sl;
        hp[IH] = sg;
    }
    {  // thread k: the output unit's weight gradient over the tile, and dZ_last[.][k]
        const uint32_t k = tid;
        const bf16 *hl = a.ht + (size_t)(16 + (L - 1) * IH + k) * B + b0;
        uint4 hv[PT / 8];
#pragma unroll
        for (uint32_t q = 0; q < PT / 8; q++) hv[q] = reinterpret_cast<const uint4 *>(hl)[q];
        float s = 0.0f;
#pragma unroll
        for (uint32_t q = 0; q < PT / 8; q++) {
            const uint32_t w[4] = {hv[q].x, hv[q].y, hv[q].z, hv[q].w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                s += g_s[8 * q + 2 * e] * __uint_as_float(w[e] << 16);
                s += g_s[8 * q + 2 * e + 1] * __uint_as_float(w[e] & 0xffff0000u);
            }
        }
        hp[k] = s;
        const float wk = a.params[a.wout_off + k];
        const float *Z = a.z + (size_t)(L - 1) * B * IH;
        bf16 *D = a.dzt + (size_t)(L - 1) * IH * B + (size_t)k * B + b0;
#pragma unroll 8
        for (uint32_t r = 0; r < PT; r++) {
            const bf16 h = to_bf16(g_s[r] * wk * silu_grad(Z[(size_t)(b0 + r) * IH + k]));
            act[0][swz(r, k)] = h;
            D[r] = h;
        }
    }
    __syncthreads();
    f32x4 acc[4][PT / 16];
    uint32_t cur = 0;
    for (uint32_t l = L - 1; l >= 1; l--) {
        tile_product(wf, act[cur], lane, 8, acc);
        if (l > 1) load_slice(a.wtb + a.wtb_off[l - 1], wave, lane, wf);
        const float *Z = a.z + (size_t)(l - 1) * B * IH;
        bf16 *D = a.dzt + (size_t)(l - 1) * IH * B;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t c = 64 * wave + 16 * j + 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < (int)(PT / 16); i++) {
                const uint32_t r = 16 * i + (lane & 15);
                const float4 zv = *reinterpret_cast<const float4 *>(&Z[(size_t)(b0 + r) * IH + c]);
                const float zz[4] = {zv.x, zv.y, zv.z, zv.w};
                uint32_t h[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    h[e] = to_bf16(acc[j][i][e] * silu_grad(zz[e]));
                    D[(size_t)(c + e) * B + b0 + r] = (bf16)h[e];
                }
                *reinterpret_cast<uint2 *>(&act[cur ^ 1][swz(r, c)]) = uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
            }
        }
        __syncthreads();
        cur ^= 1;
    }
}

// the parameter layout, by value (at most MAX_LAYERS hidden layers)
struct Layout {
    uint32_t inputs, in_pad, hidden, layers;
    uint64_t w_off[MAX_LAYERS], b_off[MAX_LAYERS], wb_off[MAX_LAYERS], wtb_off[MAX_LAYERS];
    uint64_t wout_off, bout_off, count;
};

// which parameter i is: layer (or -1 for the output unit), and for a weight its (unit, input)
__device__ __forceinline__ int locate(const Layout &L, uint64_t i, uint32_t &n, uint32_t &k, bool &bias) {
    if (i >= L.wout_off) {
        bias = i == L.bout_off;
        k = bias ? 0 : (uint32_t)(i - L.wout_off);
        n = 0;
        return -1;
    }
    int l = 0;
    while (l + 1 < (int)L.layers && i >= L.w_off[l + 1]) l++;
    const uint32_t fan = l == 0 ? L.inputs : L.hidden;
    bias = i >= L.b_off[l];
    const uint64_t j = bias ? i - L.b_off[l] : i - L.w_off[l];
    n = bias ? (uint32_t)j : (uint32_t)(j / fan);
    k = bias ? 0 : (uint32_t)(j % fan);
    return l;
}

// Where the gradients' partial sums are: dW_l per split ([split][hidden][kin] at part + l
// part_stride), db_l per split (dbpart + l MAX_SPLIT hidden), the output unit's per tile
// (head_part [tile][hidden + 1]), the loss per tile
struct ReduceArgs {
    const float *part, *dbpart, *head_part, *loss_part;
    const float *head_total;  // not NULL: the output unit's gradients and the loss, already summed
    uint64_t part_stride;
    uint32_t splits, tiles;
    float *grads, *loss;
};

// sum of n terms p[0], p[stride], ... in order (loaded MAX_SPLIT at a time, then added)
__device__ __forceinline__ float ordered_sum(const float *p, size_t stride, uint32_t n) {
    float g = 0.0f, v[MAX_SPLIT];
    for (uint32_t t0 = 0; t0 < n; t0 += MAX_SPLIT) {
#pragma unroll
        for (uint32_t z = 0; z < MAX_SPLIT; z++) v[z] = t0 + z < n ? p[(size_t)(t0 + z) * stride] : 0.0f;
#pragma unroll
        for (uint32_t z = 0; z < MAX_SPLIT; z++)
            if (t0 + z < n) g += v[z];
    }
    return g;
}

// parameter i's gradient from the partials (thread 0 also sums the loss)
__device__ __forceinline__ float reduced_grad(const Layout &L, const ReduceArgs &r, uint64_t i, int l, uint32_t n,
                                              uint32_t k, bool bias) {
    if (i == 0) r.loss[0] = r.head_total ? r.head_total[L.hidden + 1] : ordered_sum(r.loss_part, 1, r.tiles);
    if (l < 0)
        return r.head_total ? r.head_total[bias ? L.hidden : k]
                            : ordered_sum(r.head_part + (bias ? L.hidden : k), L.hidden + 1, r.tiles);
    if (bias) return ordered_sum(r.dbpart + (size_t)l * MAX_SPLIT * L.hidden + n, L.hidden, r.splits);
    const uint32_t kin = l == 0 ? L.in_pad : L.hidden;
    return ordered_sum(r.part + (size_t)l * r.part_stride + (size_t)n * kin + k, (size_t)L.hidden * kin, r.splits);
}

struct AdamArgs {
    float *p, *m, *v;
    float lr, b1, b2, eps, c1, c2;
    bf16 *wb, *wtb;
};

// Adam on parameter i, then the bf16 copies of a weight: wb [n][k] (padded rows) and, for
// layers >= 1, wtb [k][n]
__device__ __forceinline__ void adam_update(const Layout &L, const AdamArgs &a, uint64_t i, int l, uint32_t n,
                                            uint32_t k, bool bias, float gi) {
    const float mi = a.b1 * a.m[i] + (1.0f - a.b1) * gi;
    const float vi = a.b2 * a.v[i] + (1.0f - a.b2) * gi * gi;
    a.m[i] = mi;
    a.v[i] = vi;
    const float pi = a.p[i] - a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
    a.p[i] = pi;
    if (l < 0 || bias) return;
    const bf16 h = to_bf16(pi);
    a.wb[L.wb_off[l] + (size_t)n * (l == 0 ? L.in_pad : L.hidden) + k] = h;
    if (l > 0) a.wtb[L.wtb_off[l] + (size_t)k * L.hidden + n] = h;
}

// the gradients (and the loss) from the partials
__global__ void mlp_reduce(Layout L, ReduceArgs r) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.count) return;
    uint32_t n, k;
    bool bias;
    const int l = locate(L, i, n, k, bias);
    r.grads[i] = reduced_grad(L, r, i, l, n, k, bias);
}

// one Adam step on the gradients
__global__ void mlp_adam_step(Layout L, AdamArgs a, const float *grads) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.count) return;
    uint32_t n, k;
    bool bias;
    const int l = locate(L, i, n, k, bias);
    adam_update(L, a, i, l, n, k, bias, grads[i]);
}

// both in one pass (the training loop's last kernel of a step)
__global__ void mlp_reduce_adam(Layout L, ReduceArgs r, AdamArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.count) return;
    uint32_t n, k;
    bool bias;
    const int l = locate(L, i, n, k, bias);
    const float g = reduced_grad(L, r, i, l, n, k, bias);
    r.grads[i] = g;
    adam_update(L, a, i, l, n, k, bias, g);
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
    Layout layout;
    GemmTable *d_table = nullptr;  // the weight-gradient products of the per-tile path
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
    if (inputs == 0 || inputs > 256 || (hidden != 64 && hidden != 128 && hidden != 256) || layers == 0 ||
        layers > MAX_LAYERS || max_batch == 0 || max_batch % TILE)
        return PACMAN_RECURRENCE_INVALID;
    pacuda_mlp *m = new pacuda_mlp();
    Net &n = m->n;
    n.inputs = inputs;
    n.in_pad = (inputs + 15) / 16 * 16;
    n.hidden = hidden;
    n.layers = layers;
    n.max_batch = max_batch;
    n.kmax = std::max(hidden, n.in_pad);
    Layout &L = m->layout;
    L = Layout{};
    L.inputs = inputs; L.in_pad = n.in_pad; L.hidden = hidden; L.layers = layers;
    uint64_t off = 0, wb = 0, wtb = 0;
    for (uint32_t l = 0; l < layers; l++) {
        const uint32_t in = l == 0 ? inputs : hidden, in_pad = l == 0 ? n.in_pad : hidden;
        n.w_off.push_back(off);
        L.w_off[l] = off;
        off += (uint64_t)hidden * in;
        n.b_off.push_back(off);
        L.b_off[l] = off;
        off += hidden;
        n.wb_off.push_back(wb);
        L.wb_off[l] = wb;
        wb += (uint64_t)hidden * in_pad;
        n.wtb_off.push_back(wtb);
        L.wtb_off[l] = wtb;
        if (l > 0) wtb += (uint64_t)hidden * hidden;
    }
    n.wout_off = L.wout_off = off;
    off += hidden;
    n.bout_off = L.bout_off = off;
    off += 1;
    n.count = L.count = off;
    const size_t B = max_batch, tiles = max_batch / 16;  // (tiles of 64 or 32 points)
    const size_t ht_rows = n.in_pad + (size_t)layers * hidden;
    struct Alloc { void **p; size_t bytes; };
    const Alloc allocs[] = {
        {(void **)&n.params, off * 4}, {(void **)&n.grads, off * 4}, {(void **)&n.m, off * 4}, {(void **)&n.v, off * 4},
        {(void **)&n.wb, std::max<uint64_t>(wb, 1) * 2}, {(void **)&n.wtb, std::max<uint64_t>(wtb, 1) * 2},
        {(void **)&n.z, (size_t)layers * B * hidden * 4}, {(void **)&n.ht, ht_rows * B * 2},
        {(void **)&n.hb, ht_rows * B * 2}, {(void **)&n.dz[0], B * hidden * 2}, {(void **)&n.dz[1], B * hidden * 2},
        {(void **)&n.dzt[0], B * hidden * 2}, {(void **)&n.dzt[1], B * hidden * 2}, {(void **)&n.loss, 64}, {(void **)&n.delta, B * 4},
        {(void **)&n.part, (size_t)layers * MAX_SPLIT * hidden * n.kmax * 4},
        {(void **)&n.dbpart, (size_t)layers * MAX_SPLIT * hidden * 4}, {(void **)&n.loss_part, tiles * 4},
        {(void **)&n.head_part, tiles * (hidden + 1) * 4}, {(void **)&m->d_wb_off, layers * 8},
        {(void **)&m->d_b_off, layers * 8}, {(void **)&n.dztl, (size_t)layers * hidden * B * 2},
        {(void **)&n.ytrain, B * 4}, {(void **)&n.head_total, (hidden + 2) * 4},
        {(void **)&m->d_table, sizeof(GemmTable)}};
    bool ok = true;
    for (const Alloc &x : allocs) ok = ok && !hip_failed(hipMalloc(x.p, x.bytes), "mlp");
    ok = ok && !hip_failed(hipMemcpy(m->d_wb_off, n.wb_off.data(), layers * 8, hipMemcpyHostToDevice), "mlp") &&
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
                    (void *)n.ht, (void *)n.hb, (void *)n.dz[0], (void *)n.dz[1], (void *)n.dzt[0], (void *)n.dzt[1], (void *)n.loss,
                    (void *)n.delta, (void *)n.dztl, (void *)n.ytrain, (void *)n.head_total, (void *)m->d_table,
                    (void *)n.part, (void *)n.dbpart, (void *)n.loss_part, (void *)n.head_part, (void *)m->d_wb_off,
                    (void *)m->d_b_off})
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

namespace {

GemmArgs product(const bf16 *A, uint32_t lda, const bf16 *B, uint32_t ldb, uint32_t M, uint32_t N, uint32_t K,
                 uint32_t k_split, uint32_t mode, float *out, const float *bias, float *z, bf16 *h, bf16 *ht,
                 float *row_sums) {
    GemmArgs g;
    g.A = A; g.B = B; g.M = M; g.N = N; g.K = K; g.lda = lda; g.ldb = ldb; g.k_split = k_split; g.mode = mode;
    g.out = out; g.bias = bias; g.z = z; g.h = h; g.ht = ht; g.row_sums = row_sums;
    return g;
}

ReduceArgs reduce_args(pacuda_mlp *m, uint32_t splits, uint32_t tiles, bool head_summed) {
    Net &n = m->n;
    ReduceArgs r;
    r.part = n.part; r.dbpart = n.dbpart; r.head_part = n.head_part; r.loss_part = n.loss_part;
    r.head_total = head_summed ? n.head_total : nullptr;
    r.part_stride = (uint64_t)MAX_SPLIT * n.hidden * n.kmax;
    r.splits = splits; r.tiles = tiles; r.grads = n.grads; r.loss = n.loss;
    return r;
}

AdamArgs adam_args(pacuda_mlp *m, float lr, float beta1, float beta2, float eps, uint32_t step) {
    Net &n = m->n;
    AdamArgs a;
    a.p = n.params; a.m = n.m; a.v = n.v;
    a.lr = lr; a.b1 = beta1; a.b2 = beta2; a.eps = eps;
    a.c1 = 1.0f - powf(beta1, (float)step);
    a.c2 = 1.0f - powf(beta2, (float)step);
    a.wb = n.wb; a.wtb = n.wtb;
    return a;
}

// one or two NT products (gemm_tile) in one launch on the network's stream
void gemm(const GemmArgs &g0, const GemmArgs *g1 = nullptr) {
    GemmPair p;
    uint32_t blocks = 0;
    for (int k = 0; k < (g1 ? 2 : 1); k++) {
        const GemmArgs &g = k ? *g1 : g0;
        p.g[k] = g;
        p.tiles_m[k] = g.M / GT;
        p.tiles_n[k] = (g.N + GT - 1) / GT;
        if (k == 1) p.first1 = blocks;
        blocks += p.tiles_m[k] * p.tiles_n[k] * ((g.K + g.k_split - 1) / g.k_split);
    }
    if (!g1) {
        p.g[1] = g0;
        p.tiles_m[1] = p.tiles_m[0];
        p.tiles_n[1] = p.tiles_n[0];
        p.first1 = blocks;
    }
    gemm_tile<<<blocks, 256, 0, mlp_stream()>>>(p);
}

}  // namespace

namespace {

bf16 *hrow(Net &n, int l, uint32_t batch) { return n.hb + (l < 0 ? 0 : (n.in_pad + (size_t)l * n.hidden) * batch); }
bf16 *htrow(Net &n, int l, uint32_t batch) { return n.ht + (l < 0 ? 0 : (n.in_pad + (size_t)l * n.hidden) * batch); }

// the training forward: x in bf16, then Z_l = H_(l-1) W_l^T (+ bias, SiLU) layer by layer, then
// y (unless y is NULL: the backward's first kernel then computes it)
void forward_launches(pacuda_mlp *m, const float *x, uint32_t batch, float *y) {
    Net &n = m->n;
    const uint32_t H = n.hidden, L = n.layers;
    const size_t B = batch;
    const uint32_t total = batch * n.in_pad;
    mlp_input<<<(total + 255) / 256, 256, 0, mlp_stream()>>>(x, batch, n.inputs, n.in_pad, hrow(n, -1, batch),
                                                             htrow(n, -1, batch));
    for (uint32_t l = 0; l < L; l++) {
        const uint32_t K = l == 0 ? n.in_pad : H;
        gemm(product(hrow(n, (int)l - 1, batch), K, n.wb + n.wb_off[l], K, batch, H, K, K, E_FORWARD, nullptr,
                     n.params + n.b_off[l], n.z + (size_t)l * B * H, hrow(n, (int)l, batch), htrow(n, (int)l, batch),
                     nullptr));
    }
    if (y)
        mlp_output<<<batch / 64, 256, 0, mlp_stream()>>>(hrow(n, (int)L - 1, batch), H, n.params + n.wout_off,
                                                        n.params + n.bout_off, y);
}

// the K splits of the weight-gradient products (K = batch): a workgroup takes at least
// SPLIT_K_MIN of K in LDS chunks of GK whose loads overlap the previous chunk's products, and the
// few output tiles (hidden/64 x fan_in/64 per layer) times the splits still give every CU work
constexpr uint32_t SPLIT_K_MIN = 1024;
uint32_t weight_splits(uint32_t batch) { return std::max(1u, std::min(MAX_SPLIT, batch / SPLIT_K_MIN)); }

// the backward down to the partial sums (y NULL: computed from the kept activations); returns
// the dW splits used
uint32_t backward_launches(pacuda_mlp *m, const float *y, const float *target, uint32_t batch) {
    Net &n = m->n;
    hipStream_t s = mlp_stream();
    const uint32_t H = n.hidden, L = n.layers, tiles = batch / TILE;
    const size_t B = batch;
    mlp_head_grad<<<tiles, 256, 0, s>>>(y, target, batch, H, hrow(n, (int)L - 1, batch), n.params + n.wout_off,
                                        n.params + n.bout_off, n.delta, n.loss_part, n.head_part);
    mlp_head_dz<<<dim3(batch / 64, H / 64), 256, 0, s>>>(n.delta, batch, H, n.z + (size_t)(L - 1) * B * H,
                                                        n.params + n.wout_off, n.dz[0], n.dzt[0]);
    // dW_l splits K (= batch) so that the few output tiles still fill the GPU; the partials and the
    // bias gradient's row sums are summed in order afterwards
    const uint32_t splits = weight_splits(batch);
    const uint32_t k_split = (batch / splits + 15) / 16 * 16;
    const uint32_t used = (batch + k_split - 1) / k_split;
    const uint64_t part_stride = (uint64_t)MAX_SPLIT * H * n.kmax;
    int cur = 0;
    for (int l = (int)L - 1; l >= 0; l--) {
        const uint32_t K_in = l == 0 ? n.in_pad : H;
        // dW_l[n][k] = sum_b dZ_l^T[n][b] H_(l-1)^T[k][b]; db_l[n] = sum_b dZ_l^T[n][b]; and in the
        // same launch dZ_(l-1) = (dZ_l W_l) * silu'(Z_(l-1)) (A = dZ_l [b][n], B = W_l^T [k][n]),
        // which writes dZ^T into the other buffer
        const GemmArgs dw = product(n.dzt[cur], batch, htrow(n, l - 1, batch), batch, H, K_in, batch, k_split, E_PARTIAL,
                                    n.part + (size_t)l * part_stride, nullptr, nullptr, nullptr, nullptr,
                                    n.dbpart + (size_t)l * MAX_SPLIT * H);
        if (l > 0) {
            const GemmArgs dzp = product(n.dz[cur], H, n.wtb + n.wtb_off[l], H, batch, H, H, H, E_DZ, nullptr, nullptr,
                                         n.z + (size_t)(l - 1) * B * H, n.dz[cur ^ 1], n.dzt[cur ^ 1], nullptr);
            gemm(dw, &dzp);
            cur ^= 1;
        } else {
            gemm(dw);
        }
    }
    return used;
}

// the per-tile kernels take hidden 256 and at most 16 inputs
bool tile_path(const Net &n) { return n.hidden == IH && n.in_pad == 16; }

TileArgs tile_args(pacuda_mlp *m, const float *x, const float *t, float *y, uint32_t batch) {
    Net &n = m->n;
    TileArgs a;
    a.x = x; a.t = t; a.y = y; a.batch = batch; a.inputs = n.inputs; a.layers = n.layers;
    a.wb = n.wb; a.wtb = n.wtb; a.params = n.params;
    for (uint32_t l = 0; l < n.layers; l++) {
        a.wb_off[l] = n.wb_off[l];
        a.wtb_off[l] = n.wtb_off[l];
        a.b_off[l] = n.b_off[l];
    }
    a.wout_off = n.wout_off; a.bout_off = n.bout_off;
    a.z = n.z; a.ht = n.ht; a.dzt = n.dztl; a.loss_part = n.loss_part; a.head_part = n.head_part;
    return a;
}

// dW_l = dZ_l^T H_(l-1) (and db_l's row sums) of every layer in one launch, the products' table
// uploaded first when `upload` (stream-ordered; it depends only on the batch); returns the K
// splits used
uint32_t weight_gradients(pacuda_mlp *m, uint32_t batch, bool upload = true) {
    Net &n = m->n;
    const uint32_t H = n.hidden, L = n.layers;
    const uint32_t splits = weight_splits(batch);
    const uint32_t k_split = (batch / splits + 15) / 16 * 16;
    const uint32_t used = (batch + k_split - 1) / k_split;
    const uint64_t part_stride = (uint64_t)MAX_SPLIT * H * n.kmax;
    GemmTable t{};
    uint32_t blocks = 0;
    for (uint32_t l = 0; l < L; l++) {
        const uint32_t K_in = l == 0 ? n.in_pad : H;
        t.g[l] = product(n.dztl + (size_t)l * H * batch, batch, htrow(n, (int)l - 1, batch), batch, H, K_in, batch,
                         k_split, E_PARTIAL, n.part + (size_t)l * part_stride, nullptr, nullptr, nullptr, nullptr,
                         n.dbpart + (size_t)l * MAX_SPLIT * H);
        t.first[l] = blocks;
        t.tiles_m[l] = H / GT;
        t.tiles_n[l] = (K_in + GT - 1) / GT;
        blocks += t.tiles_m[l] * t.tiles_n[l] * used;
    }
    t.first[L] = blocks;
    t.count = L;
    t.head_part = n.head_part;
    t.loss_part = n.loss_part;
    t.head_total = n.head_total;
    t.tiles = batch / PT;
    t.hidden = H;
    blocks += (H + 2 + 15) / 16;
    if (upload) (void)hipMemcpyAsync(m->d_table, &t, sizeof t, hipMemcpyHostToDevice, mlp_stream());
    gemm_table<<<blocks, 256, 0, mlp_stream()>>>(m->d_table);
    return used;
}

int launched(const char *what) { return hip_failed(hipGetLastError(), what) ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK; }

int copy_loss(pacuda_mlp *m, float *loss) {
    hipStream_t s = mlp_stream();
    if (loss && (hip_failed(hipMemcpyAsync(loss, m->n.loss, 4, hipMemcpyDeviceToHost, s), "mlp") ||
                 hip_failed(hipStreamSynchronize(s), "mlp")))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}

}  // namespace

extern "C" int pacuda_mlp_forward(pacuda_mlp *m, const float *x, uint32_t batch, float *y, uint32_t train) {
    if (!m || !x || !y || batch == 0 || batch % TILE || (train && batch > m->n.max_batch)) return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    if (n.in_pad > n.hidden) return PACMAN_RECURRENCE_INVALID;
    const uint32_t H = n.hidden, L = n.layers;
    if (train) {
        if (tile_path(n))
            mlp_tile_forward<<<batch / PT, 256, 0, mlp_stream()>>>(tile_args(m, x, nullptr, y, batch));
        else
            forward_launches(m, x, batch, y);
        return launched("mlp forward");
    }
    if (H == IH && n.in_pad == 16) {
        InferArgs ia;
        ia.x = x; ia.y = y; ia.batch = batch; ia.inputs = n.inputs; ia.layers = L;
        ia.wb = n.wb; ia.params = n.params;
        for (uint32_t l = 0; l < L; l++) {
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
        return launched("mlp inference");
    }
    FwdArgs a;
    a.x = x; a.y = y; a.batch = batch; a.inputs = n.inputs; a.in_pad = n.in_pad; a.hidden = H; a.layers = L;
    a.wb = n.wb; a.wb_off = m->d_wb_off; a.params = n.params; a.b_off = m->d_b_off;
    a.wout_off = n.wout_off; a.bout_off = n.bout_off;
    const size_t lds = 2 * (size_t)TILE * H * sizeof(bf16);
    // workgroups of 64 points when that gives two per compute unit, else of 16
    if (batch / TILE >= 512) {
        mlp_forward<4><<<batch / TILE, 256, lds, mlp_stream()>>>(a);
    } else {
        mlp_forward<1><<<batch / 16, 256, lds / 4, mlp_stream()>>>(a);
    }
    return launched("mlp forward");
}

extern "C" int pacuda_mlp_backward(pacuda_mlp *m, const float *y, const float *target, uint32_t batch, float *loss) {
    if (!m || !y || !target || batch == 0 || batch % TILE || batch > m->n.max_batch) return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    uint32_t used, tiles;
    if (tile_path(n)) {
        mlp_tile_backward<<<batch / PT, 256, 0, mlp_stream()>>>(tile_args(m, nullptr, target, (float *)y, batch));
        used = weight_gradients(m, batch);
        tiles = batch / PT;
    } else {
        used = backward_launches(m, y, target, batch);
        tiles = batch / TILE;
    }
    mlp_reduce<<<(unsigned)((n.count + 255) / 256), 256, 0, mlp_stream()>>>(m->layout, reduce_args(m, used, tiles, tile_path(n)));
    if (launched("mlp backward")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    return copy_loss(m, loss);
}

extern "C" int pacuda_mlp_adam(pacuda_mlp *m, float lr, float beta1, float beta2, float eps, uint32_t step) {
    if (!m || step == 0) return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    mlp_adam_step<<<(unsigned)((n.count + 255) / 256), 256, 0, mlp_stream()>>>(
        m->layout, adam_args(m, lr, beta1, beta2, eps, step), n.grads);
    return launched("mlp adam");
}

extern "C" int pacuda_mlp_last_loss(pacuda_mlp *m, float *loss) {
    if (!m || !loss) return PACMAN_RECURRENCE_INVALID;
    return copy_loss(m, loss);
}

extern "C" int pacuda_mlp_train_steps(pacuda_mlp *m, const float *x, const float *target, uint32_t batch, uint32_t steps,
                                      double lr, uint32_t first_step, uint32_t total_steps, float beta1, float beta2,
                                      float eps, float *loss) {
    if (!m || !x || !target || batch == 0 || batch % TILE || batch > m->n.max_batch || total_steps == 0 ||
        first_step + steps > total_steps || m->n.in_pad > m->n.hidden)
        return PACMAN_RECURRENCE_INVALID;
    Net &n = m->n;
    for (uint32_t s = 0; s < steps; s++) {
        const uint32_t g = first_step + s;
        // the cosine decay of the learning rate over the whole run, as the caller's loop applies it
        const float lr_g = (float)(lr * 0.5 * (1.0 + cos(M_PI * (double)g / (double)total_steps)));
        const float *xs = x + (size_t)s * batch * n.inputs, *ts = target + (size_t)s * batch;
        uint32_t used, tiles;
        if (tile_path(n)) {
            // (the products' table is the same for every step: uploaded with the first)
            mlp_tile_forward<<<batch / PT, 256, 0, mlp_stream()>>>(tile_args(m, xs, nullptr, n.ytrain, batch));
            mlp_tile_backward<<<batch / PT, 256, 0, mlp_stream()>>>(tile_args(m, nullptr, ts, n.ytrain, batch));
            used = weight_gradients(m, batch, s == 0);
            tiles = batch / PT;
        } else {
            forward_launches(m, xs, batch, nullptr);
            used = backward_launches(m, nullptr, ts, batch);
            tiles = batch / TILE;
        }
        mlp_reduce_adam<<<(unsigned)((n.count + 255) / 256), 256, 0, mlp_stream()>>>(
            m->layout, reduce_args(m, used, tiles, tile_path(n)), adam_args(m, lr_g, beta1, beta2, eps, g + 1));
    }
    if (launched("mlp training")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    return copy_loss(m, loss);
}
