// SPDX-License-Identifier: 0BSD
//
// librecurrence.so for MI355X: the v1 C ABI of the recurrence runner (include/pacman_recurrence.h)
// on gfx950.
//
// One lane per phase-space point, 64 points per wavefront, the plan's instructions in SGPRs
// (scalar loads, uniform dispatch), the arena plane-major (structure of arrays) so each value
// access is one coalesced 512-byte wavefront load. The arena and the compiled programs persist
// between calls; only the named plane runs move; df64/mixed layouts are converted in place; all
// device work of a call is queued on one stream (default priority) and synchronised once.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <iterator>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "pacman_recurrence.h"
#include "pacuda_amd.h"
#include "plan_compiler.hpp"
#include "program.hpp"
#include "runtime.hpp"

using namespace pacuda;

// ---- kernels ---------------------------------------------------------------------------------

// momentum forms from the external momenta: 0.0 + c0 p0, then + c p in term order, f64
// (runner/recurrence_runner.cu:745-764)
__global__ void __launch_bounds__(256) fill_forms_kernel(const double *__restrict__ external, uint32_t points,
                                                         const uint32_t *__restrict__ form_terms, uint32_t form_count,
                                                         const uint32_t *__restrict__ slots,
                                                         const double *__restrict__ coefficients, double *mom,
                                                         size_t ld) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    for (uint32_t f = 0; f < form_count; f++) {
        const uint32_t first = cnst(form_terms)[2 * f], n = cnst(form_terms)[2 * f + 1];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double v = 0.0;
            for (uint32_t t = 0; t < n; t++) {
                double term = cnst(coefficients)[first + t] * external[((size_t)cnst(slots)[first + t] * 4 + k) * points + i];
                v = t == 0 ? 0.0 + term : v + term;
            }
            mom[(f * 4 + k) * ld + i] = v;
        }
    }
}

template <class R>
__global__ void __launch_bounds__(256) sources_kernel(const Instr *__restrict__ prog, uint32_t count, Arena a,
                                                      uint32_t points) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    run_sources_point<R>(prog, count, PlaneArena{a, i});
}

template <class K, bool SYMJIT>
__global__ void __launch_bounds__(1024) program_kernel(const Instr *__restrict__ prog, uint32_t count, Arena a,
                                                       uint32_t points) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    interpret<K, SYMJIT>(prog, count, PlaneArena{a, i});
}

// in-place f64 <-> df64 (runner/recurrence_runner.cu:840-880)
__global__ void to_df64_kernel(double *p, size_t count) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count; k += (size_t)gridDim.x * blockDim.x)
        str(p, k, narrow<D>(p[k]));
}
__global__ void to_df64_planes_kernel(double *p, size_t planes, size_t stride, size_t points) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < planes * points; k += (size_t)gridDim.x * blockDim.x) {
        size_t at = k / points * stride + k % points;
        str(p, at, narrow<D>(p[at]));
    }
}
__global__ void to_f64_planes_kernel(double *p, size_t planes, size_t stride, size_t points) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < planes * points; k += (size_t)gridDim.x * blockDim.x) {
        size_t at = k / points * stride + k % points;
        float2 v = reinterpret_cast<float2 *>(p)[at];
        p[at] = (double)v.x + (double)v.y;
    }
}
__global__ void f64_to_f32_planes_kernel(const double *src, float *dst, size_t planes, size_t stride, size_t points) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < planes * points; k += (size_t)gridDim.x * blockDim.x) {
        size_t at = k / points * stride + k % points;
        dst[at] = (float)src[at];
    }
}
__global__ void f32_to_f64_planes_kernel(const float *src, double *dst, size_t planes, size_t stride, size_t points) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < planes * points; k += (size_t)gridDim.x * blockDim.x) {
        size_t at = k / points * stride + k % points;
        dst[at] = (double)src[at];
    }
}

// a kept amplitude as the double it stands for (df64: hi + lo, as the download converts it)
template <bool DF> PCD double amp_at(const double *p, size_t at) {
    if constexpr (DF) {
        float2 v = reinterpret_cast<const float2 *>(p)[at];
        return (double)v.x + (double)v.y;
    } else {
        return p[at];
    }
}

// acc + w[k] * fma(re, re, im * im), k = first .. end-1 in order (runner/recurrence_runner.cu:1496-1545)
template <bool DF>
PCD double reduce_terms(const double *are, const double *aim, size_t ld, const uint32_t *__restrict__ planes,
                        const double *__restrict__ weights, uint32_t first, uint32_t end, size_t i, double acc) {
    for (uint32_t k = first; k < end; k++) {
        size_t at = (size_t)cnst(planes)[k] * ld + i;
        double re = amp_at<DF>(are, at), im = amp_at<DF>(aim, at);
        acc = acc + cnst(weights)[k] * __builtin_fma(re, re, im * im);
    }
    return acc;
}

template <bool DF>
__global__ void reduce_kernel(const double *are, const double *aim, size_t ld, const uint32_t *__restrict__ planes,
                              const double *__restrict__ weights, uint32_t count, double *output, uint32_t points) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    output[i] = reduce_terms<DF>(are, aim, ld, planes, weights, 0, count, i, output[i]);
}

// one lane per point, every output in order (from zero when no two outputs share a place, else
// from the output's value)
template <bool DF>
__global__ void reduce_resolved_kernel(const double *are, const double *aim, size_t ld, const uint32_t *__restrict__ planes,
                                       const double *__restrict__ weights, const uint32_t *__restrict__ offsets,
                                       uint32_t outputs, double *output, uint32_t points, size_t output_stride,
                                       size_t point_stride, bool from_zero) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    for (uint32_t o = 0; o < outputs; o++) {
        size_t out = o * output_stride + i * point_stride;
        double acc = from_zero ? 0.0 : output[out];
        output[out] = reduce_terms<DF>(are, aim, ld, planes, weights, cnst(offsets)[o], cnst(offsets)[o + 1], i, acc);
    }
}

// The same from zero into rows, output[i * outputs + o]: each wavefront's 64 points' values of
// up to ROW_GROUP outputs at a time go through LDS, so that its stores write whole rows (a lane's
// outputs are a row: stored straight from the lanes, a wavefront's store would touch 64 rows)
constexpr uint32_t ROW_GROUP = 16;
template <bool DF>
__global__ void __launch_bounds__(256)
reduce_rows_kernel(const double *are, const double *aim, size_t ld, const uint32_t *__restrict__ planes,
                   const double *__restrict__ weights, const uint32_t *__restrict__ offsets, uint32_t outputs,
                   double *output, uint32_t points) {
    constexpr uint32_t PITCH = ROW_GROUP + 1;  // (an odd row pitch spreads the lanes' rows over the banks)
    __shared__ double tile[4][64 * PITCH];
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const size_t p0 = ((size_t)blockIdx.x * 4 + wave) * 64, i = p0 + lane;
    const uint32_t here = p0 >= points ? 0u : points - p0 < 64 ? (uint32_t)(points - p0) : 64u;
    double *t = tile[wave];
    for (uint32_t g0 = 0; g0 < outputs; g0 += ROW_GROUP) {
        const uint32_t gn = outputs - g0 < ROW_GROUP ? outputs - g0 : ROW_GROUP;
        if (i < points)
            for (uint32_t o = 0; o < gn; o++)
                t[lane * PITCH + o] =
                    reduce_terms<DF>(are, aim, ld, planes, weights, cnst(offsets)[g0 + o], cnst(offsets)[g0 + o + 1], i, 0.0);
        __syncthreads();
        for (uint32_t e = lane; e < here * gn; e += 64) {
            const uint32_t q = e / gn, o = e - q * gn;
            output[(p0 + q) * outputs + g0 + o] = t[q * PITCH + o];
        }
        __syncthreads();
    }
}

template <bool DF>
PCD double sample_value(const double *are, const double *aim, size_t ld, const uint32_t *planes, const double *weights,
                        uint32_t first, uint32_t end, size_t i, bool factor, double num, double den, double s) {
    double acc = reduce_terms<DF>(are, aim, ld, planes, weights, first, end, i, 0.0);
    if (factor) acc = (acc * num) / den;
    if (s != 1.0) acc = acc * s;
    return acc;
}

// (runner/recurrence_runner.cu:1547-1572): the total from 0 in output order, then the first
// output whose cumulative positive value exceeds random * total (else the last positive one)
template <bool DF>
__global__ void reduce_sample_kernel(const double *are, const double *aim, size_t ld, const uint32_t *__restrict__ planes,
                                     const double *__restrict__ weights, const uint32_t *__restrict__ offsets,
                                     const uint32_t *__restrict__ helicity, uint32_t outputs, uint32_t points,
                                     const double *random, const double *scale, bool factor, double num, double den,
                                     double *total, uint32_t *chosen) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    double s = scale ? scale[i] : 1.0;
    double t = 0.0;
    for (uint32_t o = 0; o < outputs; o++)
        t = t + sample_value<DF>(are, aim, ld, planes, weights, cnst(offsets)[o], cnst(offsets)[o + 1], i, factor, num, den, s);
    double target = random[i] * t, cumulative = 0.0;
    uint32_t pick = 0;
    for (uint32_t o = 0; o < outputs; o++) {
        double v = sample_value<DF>(are, aim, ld, planes, weights, cnst(offsets)[o], cnst(offsets)[o + 1], i, factor, num, den, s);
        if (v <= 0.0) continue;
        pick = cnst(helicity)[o];
        cumulative = cumulative + v;
        if (cumulative > target) break;
    }
    total[i] = t;
    chosen[i] = pick;
}

__global__ void test_sources_kernel(const uint32_t *codes, const double *masses, const double *momenta, uint32_t count,
                                    double *out_re, double *out_im) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    W v[4] = {c64(0.0, 0.0), c64(0.0, 0.0), c64(0.0, 0.0), c64(0.0, 0.0)};
    double p[4] = {momenta[4 * i], momenta[4 * i + 1], momenta[4 * i + 2], momenta[4 * i + 3]};
    source_values(codes[i], masses[i], p, v);
    for (int k = 0; k < 4; k++) { out_re[4 * i + k] = v[k].re; out_im[4 * i + k] = v[k].im; }
}


// ---- host ------------------------------------------------------------------------------------
//
// The v1 call contract (runner/pacman_recurrence.h, runner/recurrence_runner.cu:1290-1865):
// status codes, which calls are refused, what a call moves between host and device, the kept
// amplitudes and the lock rules. The structure here:
//   Runner       the process's one device state: arenas, compiled programs, kept amplitudes,
//                configuration, statistics; created on first use and never destroyed (the HIP
//                runtime may be gone when static destructors run)
//   RunPlan      one run call: every check on the host first, then the device work queued on
//                the library's stream in the order the data needs, then one synchronisation
//   Reduction    the three reductions of a kept run, over one upload of their tables
//   RunnerLock   one lock for all entry points; a kept run leaves it with its thread until that
//                thread's reduction

namespace pacuda {

bool hip_failed(hipError_t e, const char *step) {
    if (e == hipSuccess) return false;
    fprintf(stderr, "pacman recurrence runner: %s: %s\n", step, hipGetErrorString(e));
    return true;
}

hipStream_t runner_stream() {
    static hipStream_t stream = [] {
        hipStream_t s = nullptr;
        return hip_failed(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") ? nullptr : s;
    }();
    return stream;
}

}  // namespace pacuda

namespace {

constexpr int OK = PACMAN_RECURRENCE_OK, INVALID = PACMAN_RECURRENCE_INVALID, DEVICE_ERROR = PACMAN_RECURRENCE_DEVICE_ERROR;
constexpr KernelLimits GENERATED_KERNELS = {PACUDA_KERNEL_COUNT, PACUDA_MAX_IN, PACUDA_MAX_OUT};

// A device array that grows on demand; growing drops the contents (so does the reference's arena:
// the caller uploads what a program reads).
template <class T> class DeviceArray {
  public:
    DeviceArray() = default;
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    ~DeviceArray() { release(); }
    bool fit(size_t n, const char *what) {
        n = std::max<size_t>(n, 1);
        if (n <= size_) return true;
        release();
        if (hip_failed(hipMalloc((void **)&p_, n * sizeof(T)), what)) {
            p_ = nullptr;
            return false;
        }
        size_ = n;
        return true;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        size_ = 0;
    }
    T *data() const { return p_; }

  private:
    T *p_ = nullptr;
    size_t size_ = 0;
};

// pacman_recurrence_options as the runner uses them
struct Config {
    uint32_t kernel = PACMAN_RECURRENCE_KERNEL_DF64;
    uint32_t block = 128;
    bool sort = true;
    // kernels other than f64 keep the uploaded values, the constants and (df64, mixed) the forms
    // as df64 pairs; the mixed kernels keep the currents as separate f32 planes
    bool df64_layout() const { return kernel != PACMAN_RECURRENCE_KERNEL_F64; }
    bool df64_forms() const { return kernel == PACMAN_RECURRENCE_KERNEL_DF64 || kernel == PACMAN_RECURRENCE_KERNEL_MIXED; }
    bool f32_currents() const {
        return kernel == PACMAN_RECURRENCE_KERNEL_MIXED || kernel == PACMAN_RECURRENCE_KERNEL_MIXED64;
    }
    bool operator==(const Config &o) const { return kernel == o.kernel && block == o.block && sort == o.sort; }
};

struct Stats {
    bool enabled = false;
    uint64_t calls = 0, points = 0, reductions = 0, copies = 0, bytes = 0;
    double seconds = 0;
    void copied(uint64_t count, uint64_t b) {
        copies += count;
        bytes += b;
    }
    void print() const {
        if (!enabled || calls == 0) return;
        fprintf(stderr,
                "pacman recurrence runner: %llu calls, %llu points, %llu reductions, %llu copies (%.1f MB), "
                "%.3f s (%.2f us/point)\n",
                (unsigned long long)calls, (unsigned long long)points, (unsigned long long)reductions,
                (unsigned long long)copies, bytes / 1e6, seconds, points ? seconds * 1e6 / points : 0.0);
    }
};

// wall time of a call, taken only when statistics are on
class CallTimer {
  public:
    explicit CallTimer(bool on) : on_(on) {
        if (on_) t0_ = std::chrono::steady_clock::now();
    }
    double seconds() const {
        return on_ ? std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count() : 0.0;
    }

  private:
    bool on_;
    std::chrono::steady_clock::time_point t0_;
};

// A compiled program on the device.
struct DeviceProgram {
    DeviceArray<Instr> ops;
    DeviceArray<uint32_t> bindings;
    uint32_t count = 0, sources = 0;
    uint64_t need[N_COUNT] = {0, 0, 0, 0, 0};
    bool symjit = false;  // the program holds symjit rows
};

// Programs by words key (0: not kept -- the next uncached call replaces it).
class ProgramCache {
  public:
    const DeviceProgram *get(const pacman_recurrence_call &c, bool sort, hipStream_t s) {
        if (c.words_key) {
            auto it = keyed_.find(c.words_key);
            if (it != keyed_.end()) return it->second.get();
        }
        CompiledProgram cp = compile_program(c.words, c.word_count, c.width, sort, GENERATED_KERNELS);
        if (cp.error) {
            fprintf(stderr, "pacman recurrence runner: %s\n", cp.error);
            return nullptr;
        }
        std::unique_ptr<DeviceProgram> p(new DeviceProgram);
        p->count = (uint32_t)cp.ops.size();
        p->sources = cp.sources;
        p->symjit = cp.symjit;
        std::copy(cp.need, cp.need + N_COUNT, p->need);
        // synchronous copies: the host vectors end with this function
        if (!p->ops.fit(cp.ops.size(), "program") || !p->bindings.fit(cp.bindings.size(), "bindings") ||
            (!cp.ops.empty() &&
             hip_failed(hipMemcpy(p->ops.data(), cp.ops.data(), cp.ops.size() * sizeof(Instr), hipMemcpyHostToDevice),
                        "program")) ||
            (!cp.bindings.empty() &&
             hip_failed(hipMemcpy(p->bindings.data(), cp.bindings.data(), cp.bindings.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice),
                        "bindings")))
            return nullptr;
        if (c.words_key) return (keyed_[c.words_key] = std::move(p)).get();
        // the replaced program may still be in use by queued work
        if (unkeyed_ && hip_failed(hipStreamSynchronize(s), "program")) return nullptr;
        unkeyed_ = std::move(p);
        return unkeyed_.get();
    }

  private:
    std::unordered_map<uint64_t, std::unique_ptr<DeviceProgram>> keyed_;
    std::unique_ptr<DeviceProgram> unkeyed_;
};

// Fused plans of keyed calls (words_key nonzero), each with the call shape and tables it was
// built from: a call under the same key with other tables (bit for bit) builds it again.
class FusedCache {
  public:
    struct Entry {
        pacuda_plan *plan = nullptr;       // null: the key's program does not run on the fused kernels
        std::vector<uint32_t> amplitudes;  // the amplitude planes its closures write, ascending
        std::vector<uint64_t> shape;       // the counts, widths and sort it was built for
        std::vector<uint64_t> numbers;     // its scales, factors and form coefficients (bits)
        std::vector<uint32_t> indices;     // its form terms and term slots
        bool operator==(const Entry &o) const { return shape == o.shape && numbers == o.numbers && indices == o.indices; }
    };

    const Entry &get(const pacman_recurrence_call &c, bool sort) {
        Entry e;
        describe(c, sort, e);
        auto it = entries_.find(c.words_key);
        if (it != entries_.end()) {
            if (it->second == e) return it->second;
            pacuda_plan_destroy(it->second.plan);  // (waits for the device: no launch still reads it)
            entries_.erase(it);
        }
        e.plan = v1_fused_plan(c, sort, e.amplitudes);
        return entries_[c.words_key] = std::move(e);
    }

  private:
    static void bits(std::vector<uint64_t> &out, const double *v, uint32_t n) {
        const size_t at = out.size();
        out.resize(at + n);
        if (n) memcpy(out.data() + at, v, n * sizeof(double));
    }
    static void describe(const pacman_recurrence_call &c, bool sort, Entry &e) {
        uint64_t width = 0;
        memcpy(&width, c.width, sizeof width);
        e.shape = {c.word_count, c.current_planes, c.amplitude_planes, c.momentum_planes, c.scale_count,
                   c.factor_count, c.external_count, c.term_count, width, sort ? 1u : 0u};
        bits(e.numbers, c.scale_re, c.scale_count);
        bits(e.numbers, c.scale_im, c.scale_count);
        bits(e.numbers, c.factor_re, c.factor_count);
        bits(e.numbers, c.factor_im, c.factor_count);
        bits(e.numbers, c.term_coefficients, c.term_count);
        e.indices.assign(c.form_terms, c.form_terms + 2 * (size_t)(c.momentum_planes / 4));
        e.indices.insert(e.indices.end(), c.term_slots, c.term_slots + c.term_count);
    }
    std::unordered_map<uint64_t, Entry> entries_;
};

// The amplitudes a KEEP run left on the device for the next reduction.
struct Kept {
    bool valid = false;
    bool df64 = false;  // stored as df64 pairs
    size_t stride = 0;
    uint32_t points = 0, planes = 0;
};

struct Runner {
    Config config;
    bool started = false;  // a run was attempted: the configuration is fixed
    Stats stats;
    Kept kept;
    ProgramCache programs;
    FusedCache fused;
    uint64_t fused_runs = 0;  // runs that took the fused kernels
    // a run on the fused kernels uploads its external momenta in chunks on this stream, each
    // chunk's launch waiting for its event, so later chunks' copies overlap earlier launches
    hipStream_t copies = nullptr;
    std::vector<hipEvent_t> chunk_done;
    bool chunk_events(size_t n) {
        if (!copies && hip_failed(hipStreamCreateWithFlags(&copies, hipStreamNonBlocking), "copy stream")) return false;
        while (chunk_done.size() < n) {
            hipEvent_t e = nullptr;
            if (hip_failed(hipEventCreateWithFlags(&e, hipEventDisableTiming), "copy event")) return false;
            chunk_done.push_back(e);
        }
        return true;
    }
    // the arena: currents (f64 or df64 pairs; f32 planes beside them for the mixed kernels),
    // amplitudes, momentum forms, the constants (scales re, im, factors re, im)
    DeviceArray<double> cre, cim, are, aim, forms, constants;
    DeviceArray<float> fcre, fcim;
    // FORMS_ON_DEVICE inputs
    DeviceArray<double> external, coefficients;
    DeviceArray<uint32_t> form_terms, term_slots;
    // reduction inputs and outputs
    DeviceArray<uint32_t> rplanes, roffsets, rhelicity, rchosen;
    DeviceArray<double> rweights, routput, rrandom, rscale, rtotal;

    void account(size_t points, double seconds) {
        stats.calls++;
        stats.points += points;
        stats.seconds += seconds;
    }
};

Runner &runner() {
    static Runner *r = [] {
        Runner *n = new Runner;  // never destroyed: see above
        atexit([] { runner().stats.print(); });
        return n;
    }();
    return *r;
}

unsigned grid_stride_blocks(size_t count) { return (unsigned)std::min<size_t>(4096, std::max<size_t>(1, (count + 255) / 256)); }

// A plane run: planes [first, first + count) of a table of `planes` planes.
struct PlaneRun {
    uint32_t first, count;
};
bool runs_in_range(const uint32_t *runs, uint32_t n, uint32_t planes) {
    for (uint32_t r = 0; r < n; r++)
        if (runs[2 * r] > planes || runs[2 * r + 1] > planes - runs[2 * r]) return false;
    return true;
}

// One run call.
class RunPlan {
  public:
    RunPlan(Runner &r, const pacman_recurrence_call &c) : r_(r), c_(c), s_(runner_stream()) {}

    int execute() {
        const CallTimer timer(r_.stats.enabled);
        int status = check();
        if (status == OK) choose_kernels();
        if (status == OK) status = reserve();
        if (status == OK) status = stage_inputs();
        if (status == OK) status = prepare();
        if (status == OK) status = compute();
        if (status == OK) status = deliver();
        // (a failed fused run may have left chunk copies from the caller's memory in flight)
        if (status != OK && fused_ && r_.copies) (void)hipStreamSynchronize(r_.copies);
        if (status != OK) return status;
        if (keep_) r_.kept = Kept{true, r_.config.df64_layout(), c_.stride, c_.points, c_.amplitude_planes};
        if (fused_) r_.fused_runs++;
        r_.account(c_.points, timer.seconds());
        return OK;
    }

  private:
    Runner &r_;
    const pacman_recurrence_call &c_;
    hipStream_t s_;
    bool keep_ = false, forms_on_device_ = false;
    const DeviceProgram *program_ = nullptr;
    Arena arena_{};
    // the fused kernels' plan when this run takes them (null: the v1 kernels), and the cleared
    // amplitude runs they do not write
    const FusedCache::Entry *fused_ = nullptr;
    std::vector<uint32_t> zero_runs_;
    std::vector<PlaneRun> chunks_;  // (first point, points) of each chunk of a fused run

    // Chunks of a fused run's points: 1 below 2^18 points, else up to 8 of at least 2^17 (whole
    // 64-point units). With PACUDA_SIMD_LANES a symjit row's operation order depends on the
    // point's place among the call's points, so such builds run one chunk.
    void split_points() {
        chunks_.clear();
        uint32_t n = c_.points < (1u << 18) || PACUDA_SIMD_LANES > 0 ? 1u : std::min(8u, c_.points >> 17);
        const uint32_t size = ((c_.points + n - 1) / n + 63) & ~63u;
        for (uint32_t first = 0; first < c_.points; first += size) chunks_.push_back({first, std::min(size, c_.points - first)});
    }

    static bool no_runs(const uint32_t *runs, uint32_t n) {
        for (uint32_t k = 0; k < n; k++)
            if (runs[2 * k + 1]) return false;
        return true;
    }

    // A run takes the fused kernels (launches that run forms, sources and program and store the
    // amplitudes into the arena) when it gives the same results there: a keyed call (its program is built once) with
    // the forms built on the device, nothing uploaded but the external momenta, no current
    // downloaded, a program that writes every current it reads and closures that write cleared
    // amplitude planes only. Its currents then stay on chip: the arena's current planes are not
    // written. PACUDA_V1_FUSED=0 in the environment keeps every run on the v1 kernels.
    void choose_kernels() {
        zero_runs_.assign(c_.zero_amplitude, c_.zero_amplitude + 2 * (size_t)c_.zero_amplitude_runs);
        const char *env = getenv("PACUDA_V1_FUSED");
        if (!c_.words_key || !forms_on_device_ || (env && strcmp(env, "0") == 0) ||
            (c_.scale_count && (!c_.scale_re || !c_.scale_im)) || (c_.factor_count && (!c_.factor_re || !c_.factor_im)) ||
            !no_runs(c_.upload_current, c_.upload_current_runs) ||
            !no_runs(c_.upload_amplitude, c_.upload_amplitude_runs) ||
            !no_runs(c_.download_current, c_.download_current_runs))
            return;
        const FusedCache::Entry &e = r_.fused.get(c_, r_.config.sort);
        if (!e.plan) return;
        std::vector<char> cleared(c_.amplitude_planes, 0), written(c_.amplitude_planes, 0);
        for (uint32_t k = 0; k < c_.zero_amplitude_runs; k++)
            for (uint32_t q = 0; q < c_.zero_amplitude[2 * k + 1]; q++) cleared[c_.zero_amplitude[2 * k] + q] = 1;
        for (uint32_t q : e.amplitudes) {
            if (!cleared[q]) return;  // (a closure adds to what the plane held: the v1 kernels)
            written[q] = 1;
        }
        fused_ = &e;
        // the cleared planes the launch does not write, as runs
        zero_runs_.clear();
        for (uint32_t q = 0; q < c_.amplitude_planes; q++) {
            if (!cleared[q] || written[q]) continue;
            if (!zero_runs_.empty() && zero_runs_[zero_runs_.size() - 2] + zero_runs_.back() == q) zero_runs_.back()++;
            else zero_runs_.insert(zero_runs_.end(), {q, 1u});
        }
    }

    size_t plane_doubles(uint32_t planes) const { return (size_t)planes * c_.stride; }
    double *scale_re() const { return r_.constants.data(); }
    double *scale_im() const { return scale_re() + c_.scale_count; }
    double *factor_re() const { return scale_im() + c_.scale_count; }
    double *factor_im() const { return factor_re() + c_.factor_count; }
    size_t constant_count() const { return 2 * ((size_t)c_.scale_count + c_.factor_count); }

    // everything the host can check before the device is touched
    int check() {
        if (c_.points == 0 || c_.points > c_.stride || !c_.words || (c_.zero_amplitude_runs && !c_.zero_amplitude))
            return INVALID;
        r_.kept.valid = false;  // a run ends any kept one, whatever its outcome
        keep_ = c_.flags & PACMAN_RECURRENCE_KEEP_AMPLITUDES;
        forms_on_device_ = c_.flags & PACMAN_RECURRENCE_FORMS_ON_DEVICE;
        if (forms_on_device_) {
            if (!c_.external_momenta || !c_.form_terms || (c_.term_count && (!c_.term_slots || !c_.term_coefficients)))
                return INVALID;
            if (!runs_in_range(c_.form_terms, c_.momentum_planes / 4, c_.term_count)) return INVALID;
            for (uint32_t t = 0; t < c_.term_count; t++)
                if (c_.term_slots[t] >= c_.external_count) return INVALID;
        }
        if (!runs_in_range(c_.zero_amplitude, c_.zero_amplitude_runs, c_.amplitude_planes)) return INVALID;
        // a run list past its table cannot be copied: a device error, as the copy's failure
        if (!runs_in_range(c_.upload_current, c_.upload_current_runs, c_.current_planes) ||
            !runs_in_range(c_.upload_amplitude, c_.upload_amplitude_runs, c_.amplitude_planes) ||
            !runs_in_range(c_.download_current, c_.download_current_runs, c_.current_planes) ||
            !runs_in_range(c_.download_amplitude, c_.download_amplitude_runs, c_.amplitude_planes))
            return DEVICE_ERROR;
        return s_ ? OK : DEVICE_ERROR;
    }

    int reserve() {
        const size_t currents = plane_doubles(c_.current_planes), amplitudes = plane_doubles(c_.amplitude_planes);
        if (fused_)
            return r_.are.fit(amplitudes, "amplitudes") && r_.aim.fit(amplitudes, "amplitudes") &&
                           r_.external.fit((size_t)c_.external_count * 4 * c_.points, "external momenta")
                       ? OK
                       : DEVICE_ERROR;
        bool ok = r_.cre.fit(currents, "current arena") && r_.cim.fit(currents, "current arena") &&
                  r_.are.fit(amplitudes, "amplitudes") && r_.aim.fit(amplitudes, "amplitudes") &&
                  r_.forms.fit(plane_doubles(c_.momentum_planes), "momenta") &&
                  r_.constants.fit(constant_count(), "factors");
        if (ok && r_.config.f32_currents())
            ok = r_.fcre.fit(currents, "current arena (f32)") && r_.fcim.fit(currents, "current arena (f32)");
        if (ok && forms_on_device_)
            ok = r_.external.fit((size_t)c_.external_count * 4 * c_.points, "external momenta") &&
                 r_.form_terms.fit(2 * (size_t)(c_.momentum_planes / 4), "form terms") &&
                 r_.coefficients.fit(c_.term_count, "form terms") && r_.term_slots.fit(c_.term_count, "form terms");
        return ok ? OK : DEVICE_ERROR;
    }

    bool upload(void *dst, const void *src, size_t bytes, const char *what) {
        return !bytes || !hip_failed(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s_), what);
    }
    // the first `points` values of each plane of a run, between the caller's planes and the arena
    bool copy_run(double *device, double *host, PlaneRun run, bool to_device) {
        if (run.count == 0) return true;
        const size_t pitch = c_.stride * sizeof(double), at = (size_t)run.first * c_.stride;
        const hipError_t e = to_device ? hipMemcpy2DAsync(device + at, pitch, host + at, pitch, c_.points * sizeof(double),
                                                          run.count, hipMemcpyHostToDevice, s_)
                                       : hipMemcpy2DAsync(host + at, pitch, device + at, pitch, c_.points * sizeof(double),
                                                          run.count, hipMemcpyDeviceToHost, s_);
        if (hip_failed(e, "copy")) return false;
        r_.stats.copied(1, (uint64_t)run.count * c_.points * sizeof(double));
        return true;
    }
    bool copy_runs(double *dre, double *dim, double *hre, double *him, const uint32_t *runs, uint32_t n, bool to_device) {
        for (uint32_t k = 0; k < n; k++) {
            const PlaneRun run{runs[2 * k], runs[2 * k + 1]};
            if (!copy_run(dre, hre, run, to_device) || !copy_run(dim, him, run, to_device)) return false;
        }
        return true;
    }

    bool zero_amplitudes() {
        for (size_t k = 0; k < zero_runs_.size(); k += 2) {
            const PlaneRun run{zero_runs_[k], zero_runs_[k + 1]};
            if (run.count == 0) continue;
            const size_t pitch = c_.stride * sizeof(double), at = (size_t)run.first * c_.stride;
            for (double *plane : {r_.are.data(), r_.aim.data()})
                if (hip_failed(hipMemset2DAsync(plane + at, pitch, 0, c_.points * sizeof(double), run.count, s_), "zero"))
                    return false;
        }
        return true;
    }

    int stage_inputs() {
        if (fused_) {  // the external momenta by chunks (the plan holds the tables), the cleared planes
            split_points();
            const size_t external = (size_t)c_.external_count * 4 * c_.points, pitch = c_.points * sizeof(double);
            if (!r_.chunk_events(chunks_.size())) return DEVICE_ERROR;
            for (size_t k = 0; k < chunks_.size(); k++) {
                const PlaneRun ch = chunks_[k];
                if (hip_failed(hipMemcpy2DAsync(r_.external.data() + ch.first, pitch, c_.external_momenta + ch.first, pitch,
                                                ch.count * sizeof(double), (size_t)c_.external_count * 4,
                                                hipMemcpyHostToDevice, r_.copies),
                               "external momenta") ||
                    hip_failed(hipEventRecord(r_.chunk_done[k], r_.copies), "copy event"))
                    return DEVICE_ERROR;
            }
            if (!zero_amplitudes()) return DEVICE_ERROR;
            r_.stats.copied(chunks_.size(), external * sizeof(double));
            return OK;
        }
        // momentum forms: from the external momenta on the device, or the caller's planes
        const uint32_t form_count = c_.momentum_planes / 4;
        if (forms_on_device_) {
            const size_t external = (size_t)c_.external_count * 4 * c_.points;
            if (!upload(r_.external.data(), c_.external_momenta, external * sizeof(double), "external momenta") ||
                !upload(r_.form_terms.data(), c_.form_terms, 2 * (size_t)form_count * sizeof(uint32_t), "form terms") ||
                !upload(r_.term_slots.data(), c_.term_slots, c_.term_count * sizeof(uint32_t), "form terms") ||
                !upload(r_.coefficients.data(), c_.term_coefficients, c_.term_count * sizeof(double), "form terms"))
                return DEVICE_ERROR;
            r_.stats.copied(1, external * sizeof(double));
            if (form_count) {
                fill_forms_kernel<<<(c_.points + 255) / 256, 256, 0, s_>>>(r_.external.data(), c_.points,
                                                                           r_.form_terms.data(), form_count,
                                                                           r_.term_slots.data(), r_.coefficients.data(),
                                                                           r_.forms.data(), c_.stride);
                if (hip_failed(hipGetLastError(), "forms")) return DEVICE_ERROR;
            }
        } else if (c_.momentum_planes && !copy_run(r_.forms.data(), const_cast<double *>(c_.momenta), {0, c_.momentum_planes}, true)) {
            return DEVICE_ERROR;
        }
        // scales and factors, then the named current and amplitude runs, then the zeroed runs
        if (!upload(scale_re(), c_.scale_re, c_.scale_count * sizeof(double), "scales") ||
            !upload(scale_im(), c_.scale_im, c_.scale_count * sizeof(double), "scales") ||
            !upload(factor_re(), c_.factor_re, c_.factor_count * sizeof(double), "factors") ||
            !upload(factor_im(), c_.factor_im, c_.factor_count * sizeof(double), "factors") ||
            !copy_runs(r_.cre.data(), r_.cim.data(), c_.current_re, c_.current_im, c_.upload_current,
                       c_.upload_current_runs, true) ||
            !copy_runs(r_.are.data(), r_.aim.data(), c_.amplitude_re, c_.amplitude_im, c_.upload_amplitude,
                       c_.upload_amplitude_runs, true) ||
            !zero_amplitudes())
            return DEVICE_ERROR;
        return OK;
    }

    int prepare() {
        if (fused_) return OK;
        program_ = r_.programs.get(c_, r_.config.sort, s_);
        if (!program_) return DEVICE_ERROR;
        // a program indexing past this call's tables would read or write outside the arena
        if (!program_fits(program_->need, c_.current_planes, c_.amplitude_planes, c_.momentum_planes / 4, c_.scale_count,
                          c_.factor_count)) {
            fprintf(stderr, "pacman recurrence runner: the program indexes past the call's planes or tables\n");
            return DEVICE_ERROR;
        }
        arena_.cre = r_.cre.data();
        arena_.cim = r_.cim.data();
        arena_.are = r_.are.data();
        arena_.aim = r_.aim.data();
        arena_.fcre = r_.fcre.data();
        arena_.fcim = r_.fcim.data();
        arena_.mom = r_.forms.data();
        arena_.sre = scale_re();
        arena_.sim = scale_im();
        arena_.fre = factor_re();
        arena_.fim = factor_im();
        arena_.bind = program_->bindings.data();
        arena_.ld = c_.stride;
        arena_.points = c_.points;
        return OK;
    }

    // in-place layout changes of plane runs: f64 <-> df64 pairs, or f64 <-> the f32 planes
    bool to_df64(double *p, PlaneRun run, bool forward) {
        const size_t n = (size_t)run.count * c_.points;
        if (!n) return true;
        double *at = p + (size_t)run.first * c_.stride;
        if (forward) to_df64_planes_kernel<<<grid_stride_blocks(n), 256, 0, s_>>>(at, run.count, c_.stride, c_.points);
        else to_f64_planes_kernel<<<grid_stride_blocks(n), 256, 0, s_>>>(at, run.count, c_.stride, c_.points);
        return !hip_failed(hipGetLastError(), "conversion");
    }
    bool runs_to_df64(double *re, double *im, const uint32_t *runs, uint32_t n, bool forward) {
        for (uint32_t k = 0; k < n; k++) {
            const PlaneRun run{runs[2 * k], runs[2 * k + 1]};
            if (!to_df64(re, run, forward) || !to_df64(im, run, forward)) return false;
        }
        return true;
    }
    bool runs_to_f32(const uint32_t *runs, uint32_t n, bool forward) {
        for (uint32_t k = 0; k < n; k++) {
            const size_t planes = runs[2 * k + 1], at = (size_t)runs[2 * k] * c_.stride, count = planes * c_.points;
            if (!count) continue;
            const std::pair<double *, float *> parts[2] = {{r_.cre.data() + at, r_.fcre.data() + at},
                                                           {r_.cim.data() + at, r_.fcim.data() + at}};
            for (const auto &part : parts) {
                if (forward)
                    f64_to_f32_planes_kernel<<<grid_stride_blocks(count), 256, 0, s_>>>(part.first, part.second, planes,
                                                                                      c_.stride, c_.points);
                else
                    f32_to_f64_planes_kernel<<<grid_stride_blocks(count), 256, 0, s_>>>(part.second, part.first, planes,
                                                                                      c_.stride, c_.points);
            }
            if (hip_failed(hipGetLastError(), "conversion")) return false;
        }
        return true;
    }

    template <class K> bool launch_program() {
        const Instr *ops = program_->ops.data() + program_->sources;
        const uint32_t count = program_->count - program_->sources;
        const unsigned block = r_.config.block, grid = (c_.points + block - 1) / block;
        if (program_->symjit) program_kernel<K, true><<<grid, block, 0, s_>>>(ops, count, arena_, c_.points);
        else program_kernel<K, false><<<grid, block, 0, s_>>>(ops, count, arena_, c_.points);
        return !hip_failed(hipGetLastError(), "launch");
    }
    template <class R> bool launch_sources() {
        if (!program_->sources) return true;
        sources_kernel<R><<<(c_.points + 255) / 256, 256, 0, s_>>>(program_->ops.data(), program_->sources, arena_, c_.points);
        return !hip_failed(hipGetLastError(), "sources");
    }

    int compute() {
        const Config &cfg = r_.config;
        if (fused_) {
            for (size_t k = 0; k < chunks_.size(); k++) {
                const PlaneRun ch = chunks_[k];
                if (hip_failed(hipStreamWaitEvent(s_, r_.chunk_done[k], 0), "copy event") ||
                    v1_fused_run(fused_->plan, cfg.kernel, r_.external.data() + ch.first, ch.count, c_.points,
                                 r_.are.data() + ch.first, r_.aim.data() + ch.first, c_.stride, s_) != OK)
                    return DEVICE_ERROR;
            }
            return OK;
        }
        // the sources in f64, written in the kernel's current layout
        const bool sources = cfg.kernel == PACMAN_RECURRENCE_KERNEL_F64 ? launch_sources<double>()
                             : cfg.f32_currents()                      ? launch_sources<float>()
                                                                        : launch_sources<D>();
        if (!sources) return DEVICE_ERROR;
        // the uploaded values into the kernel's layout
        if (cfg.f32_currents() && !runs_to_f32(c_.upload_current, c_.upload_current_runs, true)) return DEVICE_ERROR;
        if (cfg.df64_layout()) {
            const size_t constants = constant_count();
            bool ok = (cfg.f32_currents() ||
                       runs_to_df64(r_.cre.data(), r_.cim.data(), c_.upload_current, c_.upload_current_runs, true)) &&
                      runs_to_df64(r_.are.data(), r_.aim.data(), c_.upload_amplitude, c_.upload_amplitude_runs, true) &&
                      (!cfg.df64_forms() || to_df64(r_.forms.data(), {0, c_.momentum_planes}, true));
            if (ok && constants) {
                to_df64_kernel<<<grid_stride_blocks(constants), 256, 0, s_>>>(r_.constants.data(), constants);
                ok = !hip_failed(hipGetLastError(), "conversion");
            }
            if (!ok) return DEVICE_ERROR;
        }
        bool launched;
        switch (cfg.kernel) {
        case PACMAN_RECURRENCE_KERNEL_F64: launched = launch_program<KF64>(); break;
        case PACMAN_RECURRENCE_KERNEL_MIXED: launched = launch_program<KMixed>(); break;
        case PACMAN_RECURRENCE_KERNEL_MIXED64: launched = launch_program<KMixed64>(); break;
        default: launched = launch_program<KDF64>(); break;
        }
        return launched ? OK : DEVICE_ERROR;
    }

    int deliver() {
        const Config &cfg = r_.config;
        // kept amplitudes stay on the device, untouched
        const uint32_t amplitude_runs = keep_ ? 0 : c_.download_amplitude_runs;
        const uint32_t *amp = c_.download_amplitude, *cur = c_.download_current;
        const uint32_t current_runs = c_.download_current_runs;
        // back to f64 for the copies, then the copies, then (df64 layouts) the arena as it was
        if (cfg.df64_layout()) {
            const bool ok = runs_to_df64(r_.are.data(), r_.aim.data(), amp, amplitude_runs, false) &&
                            (cfg.f32_currents() ? runs_to_f32(cur, current_runs, false)
                                                : runs_to_df64(r_.cre.data(), r_.cim.data(), cur, current_runs, false));
            if (!ok) return DEVICE_ERROR;
        }
        if (!copy_runs(r_.are.data(), r_.aim.data(), c_.amplitude_re, c_.amplitude_im, amp, amplitude_runs, false) ||
            !copy_runs(r_.cre.data(), r_.cim.data(), c_.current_re, c_.current_im, cur, current_runs, false))
            return DEVICE_ERROR;
        if (cfg.df64_layout()) {
            const bool ok = runs_to_df64(r_.are.data(), r_.aim.data(), amp, amplitude_runs, true) &&
                            (cfg.f32_currents() || runs_to_df64(r_.cre.data(), r_.cim.data(), cur, current_runs, true));
            if (!ok) return DEVICE_ERROR;
        }
        return hip_failed(hipStreamSynchronize(s_), "kernel") ? DEVICE_ERROR : OK;
    }
};

// The reductions of a kept run: |a|^2 sums over the kept amplitude planes, per point
// (runner/pacman_recurrence.h:53-118).
class Reduction {
  public:
    explicit Reduction(Runner &r) : r_(r), s_(runner_stream()) {}

    int plain(const pacman_recurrence_reduce *q) {
        if (!q || q->abi_version != PACMAN_RECURRENCE_ABI || q->struct_size != sizeof(*q) ||
            (q->count && (!q->planes || !q->weights)) || !q->output)
            return INVALID;
        if (!matches_kept(q->points) || !planes_kept(q->planes, q->count)) return INVALID;
        const CallTimer timer(r_.stats.enabled);
        const size_t bytes = (size_t)q->points * sizeof(double);
        if (!s_ || !upload_tables(q->planes, q->weights, q->count, nullptr, 0) || !r_.routput.fit(q->points, "reduction") ||
            !upload(r_.routput.data(), q->output, bytes, "reduction"))
            return DEVICE_ERROR;
        if (q->points) {
            const unsigned grid = (q->points + 255) / 256;
            if (r_.kept.df64)
                reduce_kernel<true><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride, r_.rplanes.data(),
                                                          r_.rweights.data(), q->count, r_.routput.data(), q->points);
            else
                reduce_kernel<false><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride, r_.rplanes.data(),
                                                           r_.rweights.data(), q->count, r_.routput.data(), q->points);
            if (hip_failed(hipGetLastError(), "reduction") || !download(q->output, r_.routput.data(), bytes, "reduction"))
                return DEVICE_ERROR;
        }
        return finish(2, 2 * bytes, timer);
    }

    int resolved(const pacman_recurrence_reduce_resolved *q) {
        if (!q || q->abi_version != PACMAN_RECURRENCE_ABI || q->struct_size != sizeof(*q) || !q->offsets || !q->output ||
            q->outputs == 0 || q->outputs > 65535)
            return INVALID;
        const int valid = check_outputs(q->offsets, q->outputs, q->planes, q->weights, q->points);
        if (valid != OK || q->points == 0) return valid;
        const CallTimer timer(r_.stats.enabled);
        // the output span: every (output, point) value and whatever lies between them
        const size_t span = (size_t)(q->outputs - 1) * q->output_stride + (size_t)(q->points - 1) * q->point_stride + 1;
        const bool zero = q->flags & PACMAN_RECURRENCE_OUTPUT_ZERO;
        // outputs that fill the span, as rows per point or planes per output: nothing between them
        const bool rows = q->output_stride == 1 && q->point_stride == q->outputs;
        const bool dense = rows || (q->point_stride == 1 && (q->output_stride == q->points || q->outputs == 1));
        // OUTPUT_ZERO clears the span on the device (the values between the outputs come back as
        // zeros too; a dense span is all outputs, which the kernel writes from zero); otherwise
        // the caller's span goes up first
        if (!s_ || !upload_tables(q->planes, q->weights, q->offsets[q->outputs], q->offsets, q->outputs) ||
            !r_.routput.fit(span, "reduction") ||
            (zero ? !dense && hip_failed(hipMemsetAsync(r_.routput.data(), 0, span * sizeof(double), s_), "reduction")
                  : !upload(r_.routput.data(), q->output, span * sizeof(double), "reduction")))
            return DEVICE_ERROR;
        const unsigned grid = (q->points + 255) / 256;
        if (zero && rows) {
            if (r_.kept.df64)
                reduce_rows_kernel<true><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride, r_.rplanes.data(),
                                                               r_.rweights.data(), r_.roffsets.data(), q->outputs,
                                                               r_.routput.data(), q->points);
            else
                reduce_rows_kernel<false><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride,
                                                                r_.rplanes.data(), r_.rweights.data(), r_.roffsets.data(),
                                                                q->outputs, r_.routput.data(), q->points);
        } else if (r_.kept.df64) {
            reduce_resolved_kernel<true><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride, r_.rplanes.data(),
                                                               r_.rweights.data(), r_.roffsets.data(), q->outputs,
                                                               r_.routput.data(), q->points, q->output_stride, q->point_stride,
                                                               zero && dense);
        } else {
            reduce_resolved_kernel<false><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride, r_.rplanes.data(),
                                                                r_.rweights.data(), r_.roffsets.data(), q->outputs,
                                                                r_.routput.data(), q->points, q->output_stride, q->point_stride,
                                                                zero && dense);
        }
        if (hip_failed(hipGetLastError(), "reduction") ||
            !download(q->output, r_.routput.data(), span * sizeof(double), "reduction"))
            return DEVICE_ERROR;
        return finish(zero ? 1 : 2, (zero ? 1 : 2) * span * sizeof(double), timer);
    }

    int sample(const pacman_recurrence_reduce_sample *q) {
        if (!q || q->abi_version != PACMAN_RECURRENCE_ABI || q->struct_size != sizeof(*q) || !q->offsets || !q->helicity ||
            !q->random || !q->total || !q->chosen || q->outputs == 0)
            return INVALID;
        const int valid = check_outputs(q->offsets, q->outputs, q->planes, q->weights, q->points);
        if (valid != OK || q->points == 0) return valid;
        const CallTimer timer(r_.stats.enabled);
        const size_t n = q->points;
        if (!s_ || !upload_tables(q->planes, q->weights, q->offsets[q->outputs], q->offsets, q->outputs) ||
            !r_.rhelicity.fit(q->outputs, "sampling") || !r_.rrandom.fit(n, "sampling") || !r_.rscale.fit(n, "sampling") ||
            !r_.rtotal.fit(n, "sampling") || !r_.rchosen.fit(n, "sampling") ||
            !upload(r_.rhelicity.data(), q->helicity, q->outputs * sizeof(uint32_t), "sampling") ||
            !upload(r_.rrandom.data(), q->random, n * sizeof(double), "sampling") ||
            (q->scale && !upload(r_.rscale.data(), q->scale, n * sizeof(double), "sampling")))
            return DEVICE_ERROR;
        const bool factor = q->numerator != q->denominator;
        const double *scale = q->scale ? r_.rscale.data() : nullptr;
        const unsigned grid = (unsigned)((n + 255) / 256);
        if (r_.kept.df64)
            reduce_sample_kernel<true><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride, r_.rplanes.data(),
                                                             r_.rweights.data(), r_.roffsets.data(), r_.rhelicity.data(),
                                                             q->outputs, q->points, r_.rrandom.data(), scale, factor,
                                                             q->numerator, q->denominator, r_.rtotal.data(), r_.rchosen.data());
        else
            reduce_sample_kernel<false><<<grid, 256, 0, s_>>>(r_.are.data(), r_.aim.data(), r_.kept.stride, r_.rplanes.data(),
                                                              r_.rweights.data(), r_.roffsets.data(), r_.rhelicity.data(),
                                                              q->outputs, q->points, r_.rrandom.data(), scale, factor,
                                                              q->numerator, q->denominator, r_.rtotal.data(), r_.rchosen.data());
        if (hip_failed(hipGetLastError(), "sampling") || !download(q->total, r_.rtotal.data(), n * sizeof(double), "sampling") ||
            !download(q->chosen, r_.rchosen.data(), n * sizeof(uint32_t), "sampling"))
            return DEVICE_ERROR;
        return finish(q->scale ? 4 : 3, n * (sizeof(double) * (q->scale ? 3 : 2) + sizeof(uint32_t)), timer);
    }

  private:
    Runner &r_;
    hipStream_t s_;

    bool matches_kept(uint32_t points) const { return r_.kept.valid && points == r_.kept.points; }
    bool planes_kept(const uint32_t *planes, uint32_t count) const {
        for (uint32_t k = 0; k < count; k++)
            if (planes[k] >= r_.kept.planes) return false;
        return true;
    }
    // the offsets table of reduce_resolved / reduce_sample: from 0, non-decreasing, over kept planes
    int check_outputs(const uint32_t *offsets, uint32_t outputs, const uint32_t *planes, const double *weights,
                      uint32_t points) const {
        if (!matches_kept(points)) return INVALID;
        const uint32_t count = offsets[outputs];
        if (offsets[0] != 0 || (count && (!planes || !weights))) return INVALID;
        for (uint32_t o = 0; o < outputs; o++)
            if (offsets[o] > offsets[o + 1]) return INVALID;
        return planes_kept(planes, count) ? OK : INVALID;
    }
    bool upload(void *dst, const void *src, size_t bytes, const char *what) {
        return !bytes || !hip_failed(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s_), what);
    }
    bool download(void *dst, const void *src, size_t bytes, const char *what) {
        return !bytes || !hip_failed(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s_), what);
    }
    bool upload_tables(const uint32_t *planes, const double *weights, uint32_t count, const uint32_t *offsets,
                       uint32_t outputs) {
        return r_.rplanes.fit(count, "reduction") && r_.rweights.fit(count, "reduction") &&
               (!offsets || r_.roffsets.fit((size_t)outputs + 1, "reduction")) &&
               upload(r_.rplanes.data(), planes, count * sizeof(uint32_t), "reduction") &&
               upload(r_.rweights.data(), weights, count * sizeof(double), "reduction") &&
               (!offsets || upload(r_.roffsets.data(), offsets, ((size_t)outputs + 1) * sizeof(uint32_t), "reduction"));
    }
    int finish(uint64_t copies, uint64_t bytes, const CallTimer &timer) {
        if (hip_failed(hipStreamSynchronize(s_), "reduction")) return DEVICE_ERROR;
        r_.stats.copied(copies, bytes);
        r_.stats.reductions++;
        r_.stats.seconds += timer.seconds();
        return OK;
    }
};

// One lock serialises the entry points. A run that keeps its amplitudes returns with the lock
// still held by its thread; that thread's next reduction (or failed call) lets it go. The owner
// is a thread id, so a thread that already holds the lock passes straight through.
class RunnerLock {
  public:
    void enter() {
        if (owner_.load(std::memory_order_acquire) == std::this_thread::get_id()) return;
        mutex_.lock();
        owner_.store(std::this_thread::get_id(), std::memory_order_release);
    }
    void leave() {
        if (owner_.load(std::memory_order_acquire) != std::this_thread::get_id()) return;
        owner_.store(std::thread::id(), std::memory_order_release);
        mutex_.unlock();
    }

  private:
    std::mutex mutex_;
    std::atomic<std::thread::id> owner_{};
};

RunnerLock &runner_lock() {
    static RunnerLock *l = new RunnerLock;  // never destroyed, like the runner
    return *l;
}

// the call as the current layout: an older, shorter struct reads as zeros past its end
int normalise_call(const pacman_recurrence_call *in, pacman_recurrence_call *out) {
    static constexpr uint32_t SIZES[] = {sizeof(pacman_recurrence_call), offsetof(pacman_recurrence_call, external_momenta),
                                         offsetof(pacman_recurrence_call, zero_amplitude)};
    if (!in || in->abi_version != PACMAN_RECURRENCE_ABI ||
        std::find(std::begin(SIZES), std::end(SIZES), in->struct_size) == std::end(SIZES))
        return INVALID;
    memset(out, 0, sizeof *out);
    memcpy(out, in, in->struct_size);
    return OK;
}

}  // namespace

extern "C" int pacman_recurrence_v1_run(const pacman_recurrence_call *in) {
    RunnerLock &lock = runner_lock();
    lock.enter();
    Runner &r = runner();
    r.started = true;  // even a refused run fixes the configuration
    pacman_recurrence_call call;
    int status = normalise_call(in, &call);
    if (status == OK) status = RunPlan(r, call).execute();
    if (!(status == OK && r.kept.valid)) lock.leave();
    return status;
}

extern "C" int pacman_recurrence_v1_reduce(const pacman_recurrence_reduce *q) {
    runner_lock().enter();
    const int status = Reduction(runner()).plain(q);
    runner_lock().leave();
    return status;
}

extern "C" int pacman_recurrence_v1_reduce_resolved(const pacman_recurrence_reduce_resolved *q) {
    runner_lock().enter();
    const int status = Reduction(runner()).resolved(q);
    runner_lock().leave();
    return status;
}

extern "C" int pacman_recurrence_v1_reduce_sample(const pacman_recurrence_reduce_sample *q) {
    runner_lock().enter();
    const int status = Reduction(runner()).sample(q);
    runner_lock().leave();
    return status;
}

// The kernel, block and sort are set once: the same configuration again is accepted (several
// callers of one process share the runner), another one only before the first run (the kernel
// fixes the arena's layout). Statistics, once asked for, stay on.
extern "C" int pacman_recurrence_v1_configure(const pacman_recurrence_options *o) {
    if (!o || o->abi_version != PACMAN_RECURRENCE_ABI || o->struct_size < sizeof(pacman_recurrence_options) ||
        o->kernel > PACMAN_RECURRENCE_KERNEL_MIXED64 || (o->block && (o->block < 32 || o->block > 1024)))
        return INVALID;
    Config wanted;
    wanted.kernel = o->kernel;
    wanted.block = o->block ? o->block : 128;
    wanted.sort = !o->no_sort;
    runner_lock().enter();
    Runner &r = runner();
    int status = OK;
    if (r.started && !(wanted == r.config)) {
        fprintf(stderr, "pacman recurrence runner: another configuration after the first run\n");
        status = INVALID;
    } else {
        r.config = wanted;
        r.stats.enabled = r.stats.enabled || o->stats;
    }
    runner_lock().leave();
    return status;
}

extern "C" uint32_t pacman_recurrence_v1_capabilities(void) {
    return PACMAN_RECURRENCE_CAP_SOURCES | PACMAN_RECURRENCE_CAP_FORMS |
           (PACUDA_KERNEL_COUNT > 0 ? PACMAN_RECURRENCE_CAP_SYMJIT : 0u) | PACUDA_CAP_FUSED;
}

extern "C" uint64_t pacuda_v1_fused_runs(void) {
    runner_lock().enter();
    const uint64_t n = runner().fused_runs;
    runner_lock().leave();
    return n;
}

// The index of the generated kernel whose plane file has this SHA-256 (hex, either case), or -1.
extern "C" int pacman_recurrence_v1_symjit_kernel(const char *sha256_hex) {
    if (!sha256_hex) return -1;
    for (int k = 0; k < PACUDA_KERNEL_COUNT; k++)
        if (strcasecmp(pacuda_kernel_sha256[k], sha256_hex) == 0) return k;
    return -1;
}

// Test entry: the source wavefunctions of `count` (template, mass, momentum) items on the device.
extern "C" int pacman_recurrence_v1_test_sources(const uint32_t *codes, const double *masses, const double *momenta,
                                                 uint32_t count, double *out_re, double *out_im) {
    runner_lock().enter();
    int status = OK;
    {
        hipStream_t s = runner_stream();
        DeviceArray<uint32_t> dc;
        DeviceArray<double> dm, dp, dre, dim;
        const size_t n = count;
        bool ok = s && dc.fit(n, "test") && dm.fit(n, "test") && dp.fit(4 * n, "test") && dre.fit(4 * n, "test") &&
                  dim.fit(4 * n, "test");
        if (ok && n) {
            ok = !hip_failed(hipMemcpyAsync(dc.data(), codes, n * 4, hipMemcpyHostToDevice, s), "test") &&
                 !hip_failed(hipMemcpyAsync(dm.data(), masses, n * 8, hipMemcpyHostToDevice, s), "test") &&
                 !hip_failed(hipMemcpyAsync(dp.data(), momenta, 4 * n * 8, hipMemcpyHostToDevice, s), "test");
            if (ok) {
                test_sources_kernel<<<(count + 127) / 128, 128, 0, s>>>(dc.data(), dm.data(), dp.data(), count, dre.data(),
                                                                        dim.data());
                ok = !hip_failed(hipGetLastError(), "test") &&
                     !hip_failed(hipMemcpyAsync(out_re, dre.data(), 4 * n * 8, hipMemcpyDeviceToHost, s), "test") &&
                     !hip_failed(hipMemcpyAsync(out_im, dim.data(), 4 * n * 8, hipMemcpyDeviceToHost, s), "test");
            }
        }
        // (the buffers are freed at the end of the block, after the stream has finished with them)
        if (s && hip_failed(hipStreamSynchronize(s), "test")) ok = false;
        status = ok ? OK : DEVICE_ERROR;
    }
    runner_lock().leave();
    return status;
}

extern "C" void *pacman_recurrence_v1_host_alloc(size_t bytes) {
    void *p = nullptr;
    return hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}

extern "C" void pacman_recurrence_v1_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}
