// SPDX-License-Identifier: 0BSD
//
// Device side of unweighting: surrogate weights, weight statistics (a log histogram for
// quantile thresholds, and max / sum / sum of squares per block for a fixed-order reduction on
// the host) and ordered selection of the points that pass a flag or a hit-or-miss test against a
// threshold. Everything stays on the device; the host reads a few kilobytes per call.
//
// Hit-or-miss test of point i: keep when w[i] > 0 and u * threshold < w[i], with u the uniform
// number (id + 1) of the counter-based stream `seed` (((mix(mix(seed + g) + (id + 1) g) >> 11) +
// 0.5) / 2^53, mix the splitmix64 finaliser, g = 0x9E3779B97F4A7C15) and id = first + ids[i]
// (or first + i): a point keeps its random number wherever it sits in a batch, and the host can
// draw the same numbers.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "pacman_recurrence.h"
#include "pacuda_amd.h"
#include "runtime.hpp"

using namespace pacuda;

namespace {

constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;
constexpr uint32_t SELECT_TILE = 4096;  // points per workgroup of the selection (256 threads x 16)

__host__ __device__ __forceinline__ uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) surrogate_weight_kernel(const float *y, const double *log_prior, const uint8_t *pass,
                                                               uint32_t count, double mean, double scale, double *s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double l = mean;
    if (y) l += scale * (double)y[i];
    if (log_prior) l += log_prior[i];
    s[i] = (!pass || pass[i]) ? exp(l) : 0.0;
}

// log histogram (bins over [log_lo, log_hi), clamped at both ends; w <= 0 not counted) and, per
// workgroup of a grid of STATS_BLOCKS, the max, sum and sum of squares of its contiguous range
__global__ void __launch_bounds__(256) weight_stats_kernel(const double *w, uint32_t count, double log_lo, double log_hi,
                                                           uint32_t bins, uint32_t *hist, double *stats) {
    __shared__ double red[3][256];
    extern __shared__ uint32_t local[];  // this workgroup's histogram
    if (hist)
        for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) local[b] = 0;
    __syncthreads();
    const uint32_t per = (count + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = blockIdx.x * per, hi = min(count, lo + per);
    double mx = 0.0, s1 = 0.0, s2 = 0.0;
    const double inv = bins / (log_hi - log_lo);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const double v = w[i];
        mx = fmax(mx, v);
        s1 += v;
        s2 += v * v;
        if (hist && v > 0.0) {
            double b = floor((log(v) - log_lo) * inv);
            b = fmin(fmax(b, 0.0), (double)(bins - 1));
            atomicAdd(&local[(uint32_t)b], 1u);
        }
    }
    __syncthreads();
    if (hist)
        for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x)
            if (local[b]) atomicAdd(&hist[b], local[b]);
    red[0][threadIdx.x] = mx;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + h]);
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
            red[2][threadIdx.x] += red[2][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[3 * blockIdx.x] = red[0][0];
        stats[3 * blockIdx.x + 1] = red[1][0];
        stats[3 * blockIdx.x + 2] = red[2][0];
    }
}

struct Keep {
    const double *w;
    const uint8_t *flags;
    const uint32_t *ids;
    uint64_t first, state;
    double threshold;
    uint32_t count;
    __device__ bool operator()(uint32_t i) const {
        if (i >= count) return false;
        if (flags && !flags[i]) return false;
        if (!w) return true;
        const double v = w[i];
        if (!(v > 0.0)) return false;
        const uint64_t id = first + (ids ? ids[i] : i);
        const double u = ((double)(mix(state + (id + 1) * GAMMA) >> 11) + 0.5) / 9007199254740992.0;
        return u * threshold < v;
    }
};

__device__ __forceinline__ uint32_t block_count(const Keep &keep, uint32_t base) {
    __shared__ uint32_t part[4];
    uint32_t n = 0;
    for (uint32_t j = 0; j < SELECT_TILE / 256; j++) n += keep(base + j * 256 + threadIdx.x) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    return part[0] + part[1] + part[2] + part[3];
}

__global__ void __launch_bounds__(256) select_count_kernel(Keep keep, uint32_t *counts) {
    const uint32_t n = block_count(keep, blockIdx.x * SELECT_TILE);
    if (threadIdx.x == 0) counts[blockIdx.x] = n;
}

// exclusive scan of counts[0..blocks) in place; total into *selected
__global__ void __launch_bounds__(1024) select_scan_kernel(uint32_t *counts, uint32_t blocks, uint32_t *selected) {
    __shared__ uint32_t buf[1024];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < blocks; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < blocks ? counts[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
            const uint32_t t = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < blocks) counts[i] = carry + buf[threadIdx.x] - v;
        carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *selected = carry;
}

__global__ void __launch_bounds__(256) select_scatter_kernel(Keep keep, const uint32_t *offsets, uint32_t *index,
                                                             double *values) {
    __shared__ uint32_t wave_n[4];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t at = offsets[blockIdx.x];
    const uint32_t base = blockIdx.x * SELECT_TILE;
    for (uint32_t j = 0; j < SELECT_TILE / 256; j++) {
        const uint32_t i = base + j * 256 + threadIdx.x;
        const bool k = keep(i);
        const uint64_t ballot = __ballot(k);
        const uint32_t before = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0) wave_n[wave] = __popcll(ballot);
        __syncthreads();
        uint32_t off = at;
        for (uint32_t v = 0; v < wave; v++) off += wave_n[v];
        if (k) {
            index[off + before] = i;
            if (values) values[off + before] = keep.w[i];
        }
        at += wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
        __syncthreads();
    }
}

// grow-only workspace for the per-block counts of the selection
std::mutex work_lock;
uint32_t *work_buf = nullptr;
size_t work_len = 0;

uint32_t *select_work(size_t n) {
    if (n <= work_len) return work_buf;
    if (work_buf) {  // an earlier selection's kernels may still read it
        hip_failed(hipStreamSynchronize(runner_stream()), "select");
        hip_failed(hipFree(work_buf), "select");
    }
    work_buf = nullptr;
    work_len = 0;
    if (hip_failed(hipMalloc((void **)&work_buf, n * 4), "select")) return nullptr;
    work_len = n;
    return work_buf;
}

}  // namespace

extern "C" int pacuda_surrogate_weight(const float *y, const double *log_prior, const uint8_t *pass, uint32_t count,
                                       double mean, double scale, double *s) {
    if (!s) return PACMAN_RECURRENCE_INVALID;
    if (count == 0) return PACMAN_RECURRENCE_OK;
    surrogate_weight_kernel<<<(count + 255) / 256, 256, 0, runner_stream()>>>(y, log_prior, pass, count, mean, scale, s);
    return hip_failed(hipGetLastError(), "surrogate weight") ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_weight_stats(const double *w, uint32_t count, double log_lo, double log_hi, uint32_t bins,
                                   uint32_t *hist, double *stats) {
    if (!w || !stats || (hist && (bins == 0 || bins > PACUDA_MAX_BINS || !(log_hi > log_lo)))) return PACMAN_RECURRENCE_INVALID;
    const size_t lds = hist ? bins * sizeof(uint32_t) : 0;
    weight_stats_kernel<<<PACUDA_STATS_BLOCKS, 256, lds, runner_stream()>>>(w, count, log_lo, log_hi, bins, hist, stats);
    return hip_failed(hipGetLastError(), "weight stats") ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_select(const double *w, const uint8_t *flags, const uint32_t *ids, uint64_t first, uint32_t count,
                             double threshold, uint64_t seed, uint32_t *index, double *values, uint32_t *selected) {
    if (!index || !selected || (values && !w) || (w && !(threshold > 0.0))) return PACMAN_RECURRENCE_INVALID;
    std::lock_guard<std::mutex> g(work_lock);
    const uint32_t blocks = (count + SELECT_TILE - 1) / SELECT_TILE;
    hipStream_t s = runner_stream();
    if (blocks == 0) {
        return hip_failed(hipMemsetAsync(selected, 0, 4, s), "select") ? PACMAN_RECURRENCE_DEVICE_ERROR
                                                                       : PACMAN_RECURRENCE_OK;
    }
    uint32_t *counts = select_work(blocks);
    if (!counts) return PACMAN_RECURRENCE_DEVICE_ERROR;
    Keep keep;
    keep.w = w;
    keep.flags = flags;
    keep.ids = ids;
    keep.first = first;
    keep.state = mix(seed + GAMMA);
    keep.threshold = threshold;
    keep.count = count;
    select_count_kernel<<<blocks, 256, 0, s>>>(keep, counts);
    select_scan_kernel<<<1, 1024, 0, s>>>(counts, blocks, selected);
    select_scatter_kernel<<<blocks, 256, 0, s>>>(keep, counts, index, values);
    return hip_failed(hipGetLastError(), "select") ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK;
}
