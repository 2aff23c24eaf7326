// SPDX-License-Identifier: 0BSD
//
// librecurrence.so for MI355X: the v1 C ABI of the recurrence runner (include/pacman_recurrence.h)
// on gfx950.
//
// One lane per phase-space point, 64 points per wavefront, the plan's instructions in SGPRs
// (scalar loads, uniform dispatch), the arena plane-major (structure of arrays) so each value
// access is one coalesced 512-byte wavefront load. The host side keeps the reference's call
// semantics (runner/recurrence_runner.cu:1290-1865): the arena and the compiled programs persist
// between calls, only the named plane runs move, df64/mixed layouts are converted in place,
// KEEP holds the lock until the same thread's reduction. All device work of a call is queued on
// one stream (default priority) and synchronised once.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "pacman_recurrence.h"
#include "pacuda_amd.h"
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

template <class K>
__global__ void __launch_bounds__(1024) program_kernel(const Instr *__restrict__ prog, uint32_t count, Arena a,
                                                       uint32_t points) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    interpret<K>(prog, count, PlaneArena{a, i});
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

// one lane per point, every output in order
template <bool DF>
__global__ void reduce_resolved_kernel(const double *are, const double *aim, size_t ld, const uint32_t *__restrict__ planes,
                                       const double *__restrict__ weights, const uint32_t *__restrict__ offsets,
                                       uint32_t outputs, double *output, uint32_t points, size_t output_stride,
                                       size_t point_stride) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    for (uint32_t o = 0; o < outputs; o++) {
        size_t out = o * output_stride + i * point_stride;
        double acc = output[out];
        output[out] = reduce_terms<DF>(are, aim, ld, planes, weights, cnst(offsets)[o], cnst(offsets)[o + 1], i, acc);
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

// ---- host: programs --------------------------------------------------------------------------

namespace pacuda {

void contribution_shape(int op, int *c0, int *c1, int *cd) {
    *c0 = op == 4 ? 6 : (op == 5 || op == 6) ? 4 : 2;
    *c1 = 4;
    *cd = (op == 4 || op == 6) ? 4 : op == 5 ? 6 : 2;
}

static std::vector<HostRow> decode_words(const uint64_t *words, uint32_t count, const uint8_t *width) {
    std::vector<HostRow> rows;
    rows.reserve(count);
    for (uint32_t n = 0; n < count; n++) {
        uint64_t w = words[n];
        HostRow r;
        memset(r.f, 0, sizeof r.f);
        if ((w & 15) == 14) {  // symjit binding: kernel 16, inputs 8, output pairs 8, direct 1, data words
            r.op = 14;
            r.f[0] = (uint32_t)(w >> 4) & 0xffff;
            r.f[1] = (uint32_t)(w >> 20) & 0xff;
            r.f[2] = (uint32_t)(w >> 28) & 0xff;
            r.f[3] = (uint32_t)(w >> 36) & 1;
            uint32_t codes = r.f[1] + r.f[2];
            for (uint32_t k = 0; k < codes; k++) {
                uint64_t d = n + 1 + k / 2 < count ? words[n + 1 + k / 2] : 0;
                r.data.push_back((uint32_t)(k % 2 ? d >> 32 : d));
            }
            n += (codes + 1) / 2;
            rows.push_back(std::move(r));
            continue;
        }
        r.op = (int)(w & 15);
        w >>= 4;
        const int P = width[W_P], A = width[W_A], F = width[W_F], X = width[W_X], S = width[W_S], K = width[W_K];
        int fields[8] = {0}, nf = 0;
        auto set = [&](std::initializer_list<int> l) { for (int b : l) fields[nf++] = b; };
        switch (r.op) {
        case 1: set({P, P}); break;
        case 2: case 3: case 4: case 5: case 6: set({1, S, P, P, P, F, F, X}); break;
        case 7: case 8: case 9: set({P, F, X}); break;
        case 10: set({P, K, X}); break;
        case 11: set({P, P, A, X, X, K}); break;
        case 12: set({P, F, X, S, 10}); break;
        case 13: set({1, P, P, P, F, F, X}); break;
        default: break;
        }
        for (int k = 0; k < nf; k++) {
            r.f[k] = (uint32_t)(w & ((1ull << fields[k]) - 1));
            w >>= fields[k];
        }
        rows.push_back(std::move(r));
    }
    return rows;
}

// each maximal run of contribution rows, stably by destination, when no row of the run reads a
// component the run writes and the destinations do not overlap (each destination keeps its order)
static void sort_runs(std::vector<HostRow> &rows) {
    size_t n = 0;
    while (n < rows.size()) {
        if (rows[n].op < 2 || rows[n].op > 6) { n++; continue; }
        size_t end = n;
        while (end < rows.size() && rows[end].op >= 2 && rows[end].op <= 6) end++;
        std::unordered_map<uint32_t, uint32_t> owner;  // written component -> its destination base
        bool safe = true;
        for (size_t k = n; k < end && safe; k++) {
            int c0, c1, cd;
            contribution_shape(rows[k].op, &c0, &c1, &cd);
            for (int q = 0; q < cd; q++) {
                auto it = owner.emplace(rows[k].f[4] + q, rows[k].f[4]);
                if (it.first->second != rows[k].f[4]) safe = false;
            }
        }
        for (size_t k = n; k < end && safe; k++) {
            int c0, c1, cd;
            contribution_shape(rows[k].op, &c0, &c1, &cd);
            for (int q = 0; q < c0 && safe; q++) safe = !owner.count(rows[k].f[2] + q);
            for (int q = 0; q < c1 && safe; q++) safe = !owner.count(rows[k].f[3] + q);
        }
        if (safe)
            std::stable_sort(rows.begin() + n, rows.begin() + end,
                             [](const HostRow &x, const HostRow &y) { return x.f[4] < y.f[4]; });
        n = end;
    }
}

HostProgram build_program(const uint64_t *words, uint32_t count, const uint8_t *width, bool sort) {
    HostProgram p;
    std::vector<HostRow> rows = decode_words(words, count, width);
    if (sort) sort_runs(rows);
    for (size_t n = 0; n < rows.size(); n++) {
        const HostRow &r = rows[n];
        if (r.op == 13 || r.op == 14) {
            p.error = "a symjit row, but this runner has no symjit kernels";
            return p;
        }
        if (r.op < 1 || r.op > 12) {
            p.error = "a word with an unknown op";
            return p;
        }
        Instr o;
        memset(&o, 0, sizeof o);
        int kind, flags = 0;
        if (r.op >= 2 && r.op <= 6) {
            bool first = n == 0 || rows[n - 1].op < 2 || rows[n - 1].op > 6 || rows[n - 1].f[4] != r.f[4];
            bool last = n + 1 == rows.size() || rows[n + 1].op < 2 || rows[n + 1].op > 6 || rows[n + 1].f[4] != r.f[4];
            kind = K_CONTRIB;
            flags = (r.f[0] ? F_INIT : 0) | (first ? F_FIRST : 0) | (last ? F_LAST : 0);
            for (int k = 1; k < 8; k++) o.w[k] = r.f[k];
        } else {
            kind = r.op == 1 ? K_CLEAR : r.op <= 9 ? K_PROP : r.op == 10 ? K_IDENTITY : r.op == 11 ? K_CLOSURE : K_SOURCE;
            for (int k = 0; k < 7; k++) o.w[k + 1] = r.f[k];
        }
        o.w[0] = (uint32_t)kind | (uint32_t)r.op << 8 | (uint32_t)flags << 16;
        p.ops.push_back(o);
    }
    // the extent of every table the program indexes
    auto use = [&](int table, uint64_t first, uint64_t count) {
        if (count) p.need[table] = std::max<uint64_t>(p.need[table], first + count);
    };
    for (const Instr &o : p.ops) {
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
        if (kind == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(op, &c0, &c1, &cd);
            use(N_CURRENT, o.w[2], c0);
            use(N_CURRENT, o.w[3], c1);
            use(N_CURRENT, o.w[4], cd);
            use(N_SCALE, o.w[1], 1);
            use(N_FACTOR, o.w[7], 1);
            if (op == 6) { use(N_FORM, o.w[5], 1); use(N_FORM, o.w[6], 1); }
        } else if (kind == K_CLEAR) {
            use(N_CURRENT, o.w[1], o.w[2]);
        } else if (kind == K_PROP) {
            use(N_CURRENT, o.w[1], op == 9 ? 4 : 2);
            use(N_FORM, o.w[2], 1);
            use(N_FACTOR, o.w[3], 1);
        } else if (kind == K_IDENTITY) {
            use(N_CURRENT, o.w[1], o.w[2]);
            use(N_FACTOR, o.w[3], 1);
        } else if (kind == K_CLOSURE) {
            use(N_CURRENT, o.w[1], o.w[6]);
            use(N_CURRENT, o.w[2], o.w[6]);
            use(N_AMPLITUDE, o.w[3], 1);
            use(N_FACTOR, o.w[4], 1);
            use(N_FACTOR, o.w[5], o.w[6]);
        } else if (kind == K_SOURCE) {
            uint32_t family = o.w[5] & 7;
            use(N_CURRENT, o.w[1], family == 0 ? 1 : family == 1 ? 2 : 4);
            use(N_FORM, o.w[2], 1);
            use(N_FACTOR, o.w[3], 1);
            use(N_SCALE, o.w[4], 1);
        }
    }
    while (p.sources < p.ops.size() && (p.ops[p.sources].w[0] & 0xff) == K_SOURCE) p.sources++;
    for (size_t n = p.sources; n < p.ops.size(); n++)
        if ((p.ops[n].w[0] & 0xff) == K_SOURCE) {
            p.error = "a source row follows other rows";
            return p;
        }
    return p;
}

bool hip_failed(hipError_t e, const char *step) {
    if (e == hipSuccess) return false;
    fprintf(stderr, "pacman recurrence runner: %s: %s\n", step, hipGetErrorString(e));
    return true;
}

hipStream_t runner_stream() {
    static hipStream_t s = nullptr;
    if (!s && hip_failed(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream")) s = nullptr;
    return s;
}

}  // namespace pacuda

// ---- host: the v1 runner state ----------------------------------------------------------------

namespace {

struct Program {
    Instr *ops = nullptr;
    uint32_t count = 0, sources = 0;
    uint64_t need[N_COUNT] = {0, 0, 0, 0, 0};
};

enum { KERNEL_DF64 = 0, KERNEL_F64 = 1, KERNEL_MIXED = 2, KERNEL_MIXED64 = 3 };
struct Options {
    int kernel = KERNEL_DF64;
    bool df64 = true, df64_forms = true;
    int block = 128;
    bool sort = true;
};

template <class T> struct Buf {
    T *p = nullptr;
    size_t capacity = 0;
    // at least n elements; a grown buffer loses its contents
    bool reserve(size_t n, const char *what) {
        if (n <= capacity) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        capacity = 0;
        if (hip_failed(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)), what)) return false;
        capacity = n;
        return true;
    }
};

struct State {
    Buf<double> cre, cim, are, aim, mom, small, external, coefficients, rweights, routput, rrandom, rscale, rtotal;
    Buf<float> fcre, fcim;
    Buf<uint32_t> form_terms, slots, rplanes, roffsets, rhelicity, rchosen;
    bool kept = false, kept_df64 = false;
    size_t kept_stride = 0, kept_points = 0, kept_planes = 0;
    std::unordered_map<uint64_t, Program> programs;
    Program scratch;
    Options options;
    bool runs_started = false;
    uint64_t calls = 0, points = 0, copies = 0, bytes = 0, reductions = 0;
    double seconds = 0;
    bool stats = false;
};
State st;

double now() {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

void print_stats() {
    if (!st.stats || st.calls == 0) return;
    fprintf(stderr,
            "pacman recurrence runner: %llu calls, %llu points, %llu reductions, %llu copies (%.1f MB), %.3f s (%.2f us/point)\n",
            (unsigned long long)st.calls, (unsigned long long)st.points, (unsigned long long)st.reductions,
            (unsigned long long)st.copies, st.bytes / 1e6, st.seconds, st.points ? st.seconds * 1e6 / st.points : 0.0);
}

void initialize() {
    static bool done = false;
    if (!done) {
        done = true;
        atexit(print_stats);
    }
}

unsigned stride_grid(size_t count) { return (unsigned)std::min<size_t>(4096, std::max<size_t>(1, (count + 255) / 256)); }

// the first `points` values of `n` planes of `stride` doubles between host and device
bool copy_planes(void *dst, const void *src, size_t n, size_t stride, size_t points, hipMemcpyKind kind) {
    size_t pitch = stride * sizeof(double);
    if (hip_failed(hipMemcpy2DAsync(dst, pitch, src, pitch, points * sizeof(double), n, kind, runner_stream()), "copy"))
        return false;
    st.copies++;
    st.bytes += n * points * sizeof(double);
    return true;
}

bool copy_runs(double *dre, double *dim, double *hre, double *him, const uint32_t *runs, uint32_t count, uint32_t planes,
               size_t stride, size_t points, bool to_device) {
    for (uint32_t r = 0; r < count; r++) {
        uint32_t first = runs[2 * r], n = runs[2 * r + 1];
        if (first > planes || n > planes - first) return false;
        if (n == 0) continue;
        size_t off = (size_t)first * stride;
        bool ok = to_device ? copy_planes(dre + off, hre + off, n, stride, points, hipMemcpyHostToDevice) &&
                                  copy_planes(dim + off, him + off, n, stride, points, hipMemcpyHostToDevice)
                            : copy_planes(hre + off, dre + off, n, stride, points, hipMemcpyDeviceToHost) &&
                                  copy_planes(him + off, dim + off, n, stride, points, hipMemcpyDeviceToHost);
        if (!ok) return false;
    }
    return true;
}

bool convert_planes(bool to_df, double *p, size_t planes, size_t stride, size_t points) {
    size_t count = planes * points;
    if (count == 0) return true;
    if (to_df) to_df64_planes_kernel<<<stride_grid(count), 256, 0, runner_stream()>>>(p, planes, stride, points);
    else to_f64_planes_kernel<<<stride_grid(count), 256, 0, runner_stream()>>>(p, planes, stride, points);
    return !hip_failed(hipGetLastError(), "conversion");
}

bool convert_runs(bool to_df, double *re, double *im, const uint32_t *runs, uint32_t count, size_t stride, size_t points) {
    for (uint32_t r = 0; r < count; r++) {
        size_t off = (size_t)runs[2 * r] * stride, n = runs[2 * r + 1];
        if (!convert_planes(to_df, re + off, n, stride, points) || !convert_planes(to_df, im + off, n, stride, points))
            return false;
    }
    return true;
}

bool convert_current_runs(bool to_f32, const uint32_t *runs, uint32_t count, size_t stride, size_t points) {
    for (uint32_t r = 0; r < count; r++) {
        size_t off = (size_t)runs[2 * r] * stride, n = runs[2 * r + 1], total = n * points;
        if (total == 0) continue;
        for (int part = 0; part < 2; part++) {
            double *d = (part ? st.cim.p : st.cre.p) + off;
            float *f = (part ? st.fcim.p : st.fcre.p) + off;
            if (to_f32) f64_to_f32_planes_kernel<<<stride_grid(total), 256, 0, runner_stream()>>>(d, f, n, stride, points);
            else f32_to_f64_planes_kernel<<<stride_grid(total), 256, 0, runner_stream()>>>(f, d, n, stride, points);
        }
        if (hip_failed(hipGetLastError(), "conversion")) return false;
    }
    return true;
}

const Program *prepare(const pacman_recurrence_call *call) {
    if (call->words_key) {
        auto it = st.programs.find(call->words_key);
        if (it != st.programs.end()) return &it->second;
    }
    HostProgram hp = build_program(call->words, call->word_count, call->width, st.options.sort);
    if (hp.error) {
        fprintf(stderr, "pacman recurrence runner: %s\n", hp.error);
        return nullptr;
    }
    Program p;
    p.count = (uint32_t)hp.ops.size();
    p.sources = hp.sources;
    memcpy(p.need, hp.need, sizeof p.need);
    // a synchronous copy: the host instructions go out of scope on return
    if (hip_failed(hipMalloc((void **)&p.ops, std::max<size_t>(1, hp.ops.size()) * sizeof(Instr)), "program") ||
        (!hp.ops.empty() &&
         hip_failed(hipMemcpy(p.ops, hp.ops.data(), hp.ops.size() * sizeof(Instr), hipMemcpyHostToDevice), "program")))
        return nullptr;
    if (call->words_key) return &(st.programs[call->words_key] = p);
    if (st.scratch.ops) {
        (void)hipStreamSynchronize(runner_stream());
        (void)hipFree(st.scratch.ops);
    }
    st.scratch = p;
    return &st.scratch;
}

bool launch_program(const Program &p, const Arena &a, uint32_t points) {
    const unsigned block = (unsigned)st.options.block, grid = (points + block - 1) / block;
    const Instr *ops = p.ops + p.sources;
    const uint32_t count = p.count - p.sources;
    hipStream_t s = runner_stream();
    switch (st.options.kernel) {
    case KERNEL_F64: program_kernel<KF64><<<grid, block, 0, s>>>(ops, count, a, points); break;
    case KERNEL_MIXED: program_kernel<KMixed><<<grid, block, 0, s>>>(ops, count, a, points); break;
    case KERNEL_MIXED64: program_kernel<KMixed64><<<grid, block, 0, s>>>(ops, count, a, points); break;
    default: program_kernel<KDF64><<<grid, block, 0, s>>>(ops, count, a, points); break;
    }
    return !hip_failed(hipGetLastError(), "launch");
}

bool launch_sources(const Program &p, const Arena &a, uint32_t points) {
    if (p.sources == 0) return true;
    const unsigned block = 256, grid = (points + block - 1) / block;
    hipStream_t s = runner_stream();
    switch (st.options.kernel) {
    case KERNEL_F64: sources_kernel<double><<<grid, block, 0, s>>>(p.ops, p.sources, a, points); break;
    case KERNEL_MIXED:
    case KERNEL_MIXED64: sources_kernel<float><<<grid, block, 0, s>>>(p.ops, p.sources, a, points); break;
    default: sources_kernel<D><<<grid, block, 0, s>>>(p.ops, p.sources, a, points); break;
    }
    return !hip_failed(hipGetLastError(), "sources");
}

int run_locked(const pacman_recurrence_call *in) {
    st.runs_started = true;
    initialize();
    const size_t v1_size = offsetof(pacman_recurrence_call, zero_amplitude);
    const size_t v2_size = offsetof(pacman_recurrence_call, external_momenta);
    if (!in || in->abi_version != PACMAN_RECURRENCE_ABI ||
        (in->struct_size != sizeof(pacman_recurrence_call) && in->struct_size != v1_size && in->struct_size != v2_size))
        return PACMAN_RECURRENCE_INVALID;
    pacman_recurrence_call full;
    memset(&full, 0, sizeof full);
    memcpy(&full, in, in->struct_size);
    const pacman_recurrence_call *call = &full;
    if (call->points == 0 || call->points > call->stride || !call->words || (call->zero_amplitude_runs && !call->zero_amplitude))
        return PACMAN_RECURRENCE_INVALID;
    st.kept = false;
    const bool keep = call->flags & PACMAN_RECURRENCE_KEEP_AMPLITUDES;
    const double start = st.stats ? now() : 0;
    const size_t stride = call->stride, points = call->points;
    const size_t current_doubles = (size_t)call->current_planes * stride;
    const size_t amplitude_doubles = (size_t)call->amplitude_planes * stride;
    hipStream_t s = runner_stream();
    if (!s) return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (!st.cre.reserve(current_doubles, "current arena") || !st.cim.reserve(current_doubles, "current arena") ||
        !st.are.reserve(amplitude_doubles, "amplitudes") || !st.aim.reserve(amplitude_doubles, "amplitudes"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    const Options opt = st.options;
    const bool mixed = opt.kernel == KERNEL_MIXED || opt.kernel == KERNEL_MIXED64;
    if (mixed && (!st.fcre.reserve(current_doubles, "current arena (f32)") ||
                  !st.fcim.reserve(current_doubles, "current arena (f32)")))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    const size_t momentum_doubles = (size_t)call->momentum_planes * stride;
    const size_t small = 2 * ((size_t)call->scale_count + call->factor_count);
    if (!st.mom.reserve(std::max<size_t>(1, momentum_doubles), "momenta") ||
        !st.small.reserve(std::max<size_t>(1, small), "factors"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    double *dsre = st.small.p, *dsim = dsre + call->scale_count;
    double *dfre = dsim + call->scale_count, *dfim = dfre + call->factor_count;
    const bool forms_on_device = call->flags & PACMAN_RECURRENCE_FORMS_ON_DEVICE;
    if (forms_on_device) {
        const uint32_t form_count = call->momentum_planes / 4;
        const size_t external_doubles = (size_t)call->external_count * 4 * points;
        if (!call->external_momenta || !call->form_terms || (call->term_count && (!call->term_slots || !call->term_coefficients)))
            return PACMAN_RECURRENCE_INVALID;
        for (uint32_t f = 0; f < form_count; f++)
            if (call->form_terms[2 * f] > call->term_count || call->form_terms[2 * f + 1] > call->term_count - call->form_terms[2 * f])
                return PACMAN_RECURRENCE_INVALID;
        for (uint32_t t = 0; t < call->term_count; t++)
            if (call->term_slots[t] >= call->external_count) return PACMAN_RECURRENCE_INVALID;
        if (!st.external.reserve(std::max<size_t>(1, external_doubles), "external momenta") ||
            !st.form_terms.reserve(std::max<size_t>(1, 2 * (size_t)form_count), "form terms") ||
            !st.coefficients.reserve(std::max<size_t>(1, call->term_count), "form terms") ||
            !st.slots.reserve(std::max<size_t>(1, call->term_count), "form terms"))
            return PACMAN_RECURRENCE_DEVICE_ERROR;
        if ((external_doubles && hip_failed(hipMemcpyAsync(st.external.p, call->external_momenta, external_doubles * 8,
                                                           hipMemcpyHostToDevice, s), "external momenta")) ||
            (form_count && hip_failed(hipMemcpyAsync(st.form_terms.p, call->form_terms, 2 * (size_t)form_count * 4,
                                                     hipMemcpyHostToDevice, s), "form terms")) ||
            (call->term_count &&
             (hip_failed(hipMemcpyAsync(st.slots.p, call->term_slots, call->term_count * 4, hipMemcpyHostToDevice, s), "form terms") ||
              hip_failed(hipMemcpyAsync(st.coefficients.p, call->term_coefficients, call->term_count * 8,
                                        hipMemcpyHostToDevice, s), "form terms"))))
            return PACMAN_RECURRENCE_DEVICE_ERROR;
        st.copies += 1;
        st.bytes += external_doubles * 8;
        if (form_count) {
            fill_forms_kernel<<<(unsigned)((points + 255) / 256), 256, 0, s>>>(st.external.p, (uint32_t)points, st.form_terms.p,
                                                                               form_count, st.slots.p, st.coefficients.p,
                                                                               st.mom.p, stride);
            if (hip_failed(hipGetLastError(), "forms")) return PACMAN_RECURRENCE_DEVICE_ERROR;
        }
    }
    if ((momentum_doubles && !forms_on_device &&
         !copy_planes(st.mom.p, call->momenta, call->momentum_planes, stride, points, hipMemcpyHostToDevice)) ||
        (call->scale_count &&
         (hip_failed(hipMemcpyAsync(dsre, call->scale_re, call->scale_count * 8, hipMemcpyHostToDevice, s), "scales") ||
          hip_failed(hipMemcpyAsync(dsim, call->scale_im, call->scale_count * 8, hipMemcpyHostToDevice, s), "scales"))) ||
        (call->factor_count &&
         (hip_failed(hipMemcpyAsync(dfre, call->factor_re, call->factor_count * 8, hipMemcpyHostToDevice, s), "factors") ||
          hip_failed(hipMemcpyAsync(dfim, call->factor_im, call->factor_count * 8, hipMemcpyHostToDevice, s), "factors"))))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (!copy_runs(st.cre.p, st.cim.p, call->current_re, call->current_im, call->upload_current, call->upload_current_runs,
                   call->current_planes, stride, points, true) ||
        !copy_runs(st.are.p, st.aim.p, call->amplitude_re, call->amplitude_im, call->upload_amplitude,
                   call->upload_amplitude_runs, call->amplitude_planes, stride, points, true))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    for (uint32_t r = 0; r < call->zero_amplitude_runs; r++) {
        uint32_t first = call->zero_amplitude[2 * r], n = call->zero_amplitude[2 * r + 1];
        if (first > call->amplitude_planes || n > call->amplitude_planes - first) return PACMAN_RECURRENCE_INVALID;
        if (n == 0) continue;
        size_t pitch = stride * 8, width = points * 8;
        if (hip_failed(hipMemset2DAsync(st.are.p + (size_t)first * stride, pitch, 0, width, n, s), "zero") ||
            hip_failed(hipMemset2DAsync(st.aim.p + (size_t)first * stride, pitch, 0, width, n, s), "zero"))
            return PACMAN_RECURRENCE_DEVICE_ERROR;
    }
    Arena a;
    a.cre = st.cre.p; a.cim = st.cim.p; a.are = st.are.p; a.aim = st.aim.p;
    a.fcre = st.fcre.p; a.fcim = st.fcim.p; a.mom = st.mom.p;
    a.sre = dsre; a.sim = dsim; a.fre = dfre; a.fim = dfim; a.ld = stride;
    const Program *prepared = prepare(call);
    if (!prepared) return PACMAN_RECURRENCE_DEVICE_ERROR;
    // a program indexing past this call's tables would read or write outside the arena
    if (!program_fits(prepared->need, call->current_planes, call->amplitude_planes, call->momentum_planes / 4,
                      call->scale_count, call->factor_count)) {
        fprintf(stderr, "pacman recurrence runner: the program indexes past the call's planes or tables\n");
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    }
    if (!launch_sources(*prepared, a, (uint32_t)points)) return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (mixed && !convert_current_runs(true, call->upload_current, call->upload_current_runs, stride, points))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (opt.df64) {
        bool ok = (mixed || convert_runs(true, st.cre.p, st.cim.p, call->upload_current, call->upload_current_runs, stride, points)) &&
                  convert_runs(true, st.are.p, st.aim.p, call->upload_amplitude, call->upload_amplitude_runs, stride, points) &&
                  (!opt.df64_forms || convert_planes(true, st.mom.p, call->momentum_planes, stride, points));
        if (ok && small) {
            to_df64_kernel<<<stride_grid(small), 256, 0, s>>>(st.small.p, small);
            ok = !hip_failed(hipGetLastError(), "conversion");
        }
        if (!ok) return PACMAN_RECURRENCE_DEVICE_ERROR;
    }
    if (!launch_program(*prepared, a, (uint32_t)points)) return PACMAN_RECURRENCE_DEVICE_ERROR;
    const uint32_t amplitude_runs = keep ? 0 : call->download_amplitude_runs;
    if (opt.df64 &&
        (!convert_runs(false, st.are.p, st.aim.p, call->download_amplitude, amplitude_runs, stride, points) ||
         (mixed ? !convert_current_runs(false, call->download_current, call->download_current_runs, stride, points)
                : !convert_runs(false, st.cre.p, st.cim.p, call->download_current, call->download_current_runs, stride, points))))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (!copy_runs(st.are.p, st.aim.p, call->amplitude_re, call->amplitude_im, call->download_amplitude, amplitude_runs,
                   call->amplitude_planes, stride, points, false) ||
        !copy_runs(st.cre.p, st.cim.p, call->current_re, call->current_im, call->download_current,
                   call->download_current_runs, call->current_planes, stride, points, false))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (opt.df64 &&
        (!convert_runs(true, st.are.p, st.aim.p, call->download_amplitude, amplitude_runs, stride, points) ||
         (!mixed && !convert_runs(true, st.cre.p, st.cim.p, call->download_current, call->download_current_runs, stride, points))))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (hip_failed(hipStreamSynchronize(s), "kernel")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (keep) {
        st.kept = true;
        st.kept_df64 = opt.df64;
        st.kept_stride = stride;
        st.kept_points = points;
        st.kept_planes = call->amplitude_planes;
    }
    st.calls++;
    st.points += points;
    if (st.stats) st.seconds += now() - start;
    return PACMAN_RECURRENCE_OK;
}

// validation shared by reduce_resolved and reduce_sample
int check_outputs(const uint32_t *offsets, uint32_t outputs, const uint32_t *planes, const double *weights, uint32_t points) {
    if (!st.kept || points != st.kept_points) return PACMAN_RECURRENCE_INVALID;
    uint32_t count = offsets[outputs];
    if (offsets[0] != 0 || (count && (!planes || !weights))) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t o = 0; o < outputs; o++)
        if (offsets[o] > offsets[o + 1]) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t k = 0; k < count; k++)
        if (planes[k] >= st.kept_planes) return PACMAN_RECURRENCE_INVALID;
    return PACMAN_RECURRENCE_OK;
}

bool upload_reduction(const uint32_t *planes, const double *weights, uint32_t count, const uint32_t *offsets, uint32_t outputs) {
    hipStream_t s = runner_stream();
    if (!st.rplanes.reserve(std::max<size_t>(count, 1), "reduction") || !st.rweights.reserve(std::max<size_t>(count, 1), "reduction") ||
        (offsets && !st.roffsets.reserve(outputs + 1, "reduction")))
        return false;
    return !((count && (hip_failed(hipMemcpyAsync(st.rplanes.p, planes, count * 4, hipMemcpyHostToDevice, s), "reduction") ||
                        hip_failed(hipMemcpyAsync(st.rweights.p, weights, count * 8, hipMemcpyHostToDevice, s), "reduction"))) ||
             (offsets && hip_failed(hipMemcpyAsync(st.roffsets.p, offsets, (outputs + 1) * 4, hipMemcpyHostToDevice, s), "reduction")));
}

int reduce_resolved_locked(const pacman_recurrence_reduce_resolved *r) {
    initialize();
    if (!r || r->abi_version != PACMAN_RECURRENCE_ABI || r->struct_size != sizeof(pacman_recurrence_reduce_resolved) ||
        !r->offsets || !r->output || r->outputs == 0 || r->outputs > 65535)
        return PACMAN_RECURRENCE_INVALID;
    int c = check_outputs(r->offsets, r->outputs, r->planes, r->weights, r->points);
    if (c) return c;
    if (r->points == 0) return PACMAN_RECURRENCE_OK;
    const size_t span = (size_t)(r->outputs - 1) * r->output_stride + (size_t)(r->points - 1) * r->point_stride + 1;
    const bool zero = r->flags & PACMAN_RECURRENCE_OUTPUT_ZERO;
    const double start = st.stats ? now() : 0;
    hipStream_t s = runner_stream();
    const uint32_t count = r->offsets[r->outputs];
    // OUTPUT_ZERO clears the whole span on the device (the values between the outputs come back
    // as zeros too); otherwise the span goes up and comes back
    if (!upload_reduction(r->planes, r->weights, count, r->offsets, r->outputs) || !st.routput.reserve(span, "reduction") ||
        (zero ? hip_failed(hipMemsetAsync(st.routput.p, 0, span * 8, s), "reduction")
              : hip_failed(hipMemcpyAsync(st.routput.p, r->output, span * 8, hipMemcpyHostToDevice, s), "reduction")))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    const unsigned grid = (r->points + 255) / 256;
    if (st.kept_df64)
        reduce_resolved_kernel<true><<<grid, 256, 0, s>>>(st.are.p, st.aim.p, st.kept_stride, st.rplanes.p, st.rweights.p,
                                                          st.roffsets.p, r->outputs, st.routput.p, r->points,
                                                          r->output_stride, r->point_stride);
    else
        reduce_resolved_kernel<false><<<grid, 256, 0, s>>>(st.are.p, st.aim.p, st.kept_stride, st.rplanes.p, st.rweights.p,
                                                           st.roffsets.p, r->outputs, st.routput.p, r->points,
                                                           r->output_stride, r->point_stride);
    if (hip_failed(hipGetLastError(), "reduction") ||
        hip_failed(hipMemcpyAsync(r->output, st.routput.p, span * 8, hipMemcpyDeviceToHost, s), "reduction") ||
        hip_failed(hipStreamSynchronize(s), "reduction"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    st.copies += zero ? 1 : 2;
    st.bytes += (zero ? 1 : 2) * span * 8;
    st.reductions++;
    if (st.stats) st.seconds += now() - start;
    return PACMAN_RECURRENCE_OK;
}

int reduce_sample_locked(const pacman_recurrence_reduce_sample *r) {
    initialize();
    if (!r || r->abi_version != PACMAN_RECURRENCE_ABI || r->struct_size != sizeof(pacman_recurrence_reduce_sample) ||
        !r->offsets || !r->helicity || !r->random || !r->total || !r->chosen || r->outputs == 0)
        return PACMAN_RECURRENCE_INVALID;
    int c = check_outputs(r->offsets, r->outputs, r->planes, r->weights, r->points);
    if (c) return c;
    if (r->points == 0) return PACMAN_RECURRENCE_OK;
    const double start = st.stats ? now() : 0;
    hipStream_t s = runner_stream();
    const size_t n = r->points;
    const uint32_t count = r->offsets[r->outputs];
    if (!upload_reduction(r->planes, r->weights, count, r->offsets, r->outputs) ||
        !st.rhelicity.reserve(r->outputs, "sampling") || !st.rrandom.reserve(n, "sampling") ||
        !st.rscale.reserve(n, "sampling") || !st.rtotal.reserve(n, "sampling") || !st.rchosen.reserve(n, "sampling"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (hip_failed(hipMemcpyAsync(st.rhelicity.p, r->helicity, r->outputs * 4, hipMemcpyHostToDevice, s), "sampling") ||
        hip_failed(hipMemcpyAsync(st.rrandom.p, r->random, n * 8, hipMemcpyHostToDevice, s), "sampling") ||
        (r->scale && hip_failed(hipMemcpyAsync(st.rscale.p, r->scale, n * 8, hipMemcpyHostToDevice, s), "sampling")))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    const bool factor = r->numerator != r->denominator;
    const unsigned grid = (unsigned)((n + 255) / 256);
    const double *scale = r->scale ? st.rscale.p : nullptr;
    if (st.kept_df64)
        reduce_sample_kernel<true><<<grid, 256, 0, s>>>(st.are.p, st.aim.p, st.kept_stride, st.rplanes.p, st.rweights.p,
                                                        st.roffsets.p, st.rhelicity.p, r->outputs, r->points, st.rrandom.p,
                                                        scale, factor, r->numerator, r->denominator, st.rtotal.p, st.rchosen.p);
    else
        reduce_sample_kernel<false><<<grid, 256, 0, s>>>(st.are.p, st.aim.p, st.kept_stride, st.rplanes.p, st.rweights.p,
                                                         st.roffsets.p, st.rhelicity.p, r->outputs, r->points, st.rrandom.p,
                                                         scale, factor, r->numerator, r->denominator, st.rtotal.p, st.rchosen.p);
    if (hip_failed(hipGetLastError(), "sampling") ||
        hip_failed(hipMemcpyAsync(r->total, st.rtotal.p, n * 8, hipMemcpyDeviceToHost, s), "sampling") ||
        hip_failed(hipMemcpyAsync(r->chosen, st.rchosen.p, n * 4, hipMemcpyDeviceToHost, s), "sampling") ||
        hip_failed(hipStreamSynchronize(s), "sampling"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    st.copies += r->scale ? 4 : 3;
    st.bytes += n * (8 * (r->scale ? 3 : 2) + 4);
    st.reductions++;
    if (st.stats) st.seconds += now() - start;
    return PACMAN_RECURRENCE_OK;
}

int reduce_locked(const pacman_recurrence_reduce *r) {
    initialize();
    if (!r || r->abi_version != PACMAN_RECURRENCE_ABI || r->struct_size != sizeof(pacman_recurrence_reduce) ||
        (r->count && (!r->planes || !r->weights)) || !r->output)
        return PACMAN_RECURRENCE_INVALID;
    if (!st.kept || r->points != st.kept_points) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t k = 0; k < r->count; k++)
        if (r->planes[k] >= st.kept_planes) return PACMAN_RECURRENCE_INVALID;
    const double start = st.stats ? now() : 0;
    hipStream_t s = runner_stream();
    const size_t bytes = (size_t)r->points * 8;
    if (!upload_reduction(r->planes, r->weights, r->count, nullptr, 0) ||
        !st.routput.reserve(std::max<uint32_t>(r->points, 1), "reduction") ||
        (bytes && hip_failed(hipMemcpyAsync(st.routput.p, r->output, bytes, hipMemcpyHostToDevice, s), "reduction")))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (r->points) {
        const unsigned grid = (r->points + 255) / 256;
        if (st.kept_df64)
            reduce_kernel<true><<<grid, 256, 0, s>>>(st.are.p, st.aim.p, st.kept_stride, st.rplanes.p, st.rweights.p,
                                                     r->count, st.routput.p, r->points);
        else
            reduce_kernel<false><<<grid, 256, 0, s>>>(st.are.p, st.aim.p, st.kept_stride, st.rplanes.p, st.rweights.p,
                                                      r->count, st.routput.p, r->points);
        if (hip_failed(hipGetLastError(), "reduction") ||
            hip_failed(hipMemcpyAsync(r->output, st.routput.p, bytes, hipMemcpyDeviceToHost, s), "reduction"))
            return PACMAN_RECURRENCE_DEVICE_ERROR;
    }
    if (hip_failed(hipStreamSynchronize(s), "reduction")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    st.copies += 2;
    st.bytes += 2 * bytes;
    st.reductions++;
    if (st.stats) st.seconds += now() - start;
    return PACMAN_RECURRENCE_OK;
}

// One lock serialises the entry points; a KEEP run keeps it until the same thread's reduction
// (runner/recurrence_runner.cu:1725-1748).
std::mutex runner_lock;
thread_local bool holding_runner_lock = false;
void acquire_runner_lock() {
    if (!holding_runner_lock) {
        runner_lock.lock();
        holding_runner_lock = true;
    }
}
void release_runner_lock() {
    if (holding_runner_lock) {
        holding_runner_lock = false;
        runner_lock.unlock();
    }
}

}  // namespace

extern "C" int pacman_recurrence_v1_run(const pacman_recurrence_call *in) {
    acquire_runner_lock();
    int status = run_locked(in);
    if (!(status == PACMAN_RECURRENCE_OK && st.kept)) release_runner_lock();
    return status;
}

extern "C" int pacman_recurrence_v1_configure(const pacman_recurrence_options *o) {
    if (!o || o->abi_version != PACMAN_RECURRENCE_ABI || o->struct_size < sizeof(pacman_recurrence_options))
        return PACMAN_RECURRENCE_INVALID;
    if (o->kernel > PACMAN_RECURRENCE_KERNEL_MIXED64 || (o->block && (o->block < 32 || o->block > 1024)))
        return PACMAN_RECURRENCE_INVALID;
    acquire_runner_lock();
    int status = PACMAN_RECURRENCE_OK;
    Options n;
    n.kernel = (int)o->kernel;
    n.df64 = n.kernel != KERNEL_F64;
    n.df64_forms = n.kernel == KERNEL_DF64 || n.kernel == KERNEL_MIXED;
    n.block = o->block ? (int)o->block : 128;
    n.sort = !o->no_sort;
    const bool same = n.kernel == st.options.kernel && n.block == st.options.block && n.sort == st.options.sort;
    if (st.runs_started && !same) {
        fprintf(stderr, "pacman recurrence runner: another configuration after the first run\n");
        status = PACMAN_RECURRENCE_INVALID;
    } else {
        st.options = n;
        st.stats = st.stats || o->stats != 0;
    }
    release_runner_lock();
    return status;
}

extern "C" uint32_t pacman_recurrence_v1_capabilities(void) {
    return PACMAN_RECURRENCE_CAP_SOURCES | PACMAN_RECURRENCE_CAP_FORMS | PACUDA_CAP_FUSED;
}

// This build carries no generated symjit kernels (op 13/14 rows are refused), so no SHA-256
// names one.
extern "C" int pacman_recurrence_v1_symjit_kernel(const char *sha256_hex) {
    (void)sha256_hex;
    return -1;
}

extern "C" int pacman_recurrence_v1_test_sources(const uint32_t *codes, const double *masses, const double *momenta,
                                                 uint32_t count, double *out_re, double *out_im) {
    acquire_runner_lock();
    uint32_t *dc = nullptr;
    double *dm = nullptr, *dp = nullptr, *dre = nullptr, *dim = nullptr;
    hipStream_t s = runner_stream();
    size_t n = std::max<uint32_t>(count, 1);
    bool ok = s && !hip_failed(hipMalloc((void **)&dc, n * 4), "test") && !hip_failed(hipMalloc((void **)&dm, n * 8), "test") &&
              !hip_failed(hipMalloc((void **)&dp, 4 * n * 8), "test") && !hip_failed(hipMalloc((void **)&dre, 4 * n * 8), "test") &&
              !hip_failed(hipMalloc((void **)&dim, 4 * n * 8), "test");
    if (ok && count) {
        ok = !hip_failed(hipMemcpyAsync(dc, codes, count * 4, hipMemcpyHostToDevice, s), "test") &&
             !hip_failed(hipMemcpyAsync(dm, masses, count * 8, hipMemcpyHostToDevice, s), "test") &&
             !hip_failed(hipMemcpyAsync(dp, momenta, 4 * (size_t)count * 8, hipMemcpyHostToDevice, s), "test");
        if (ok) {
            test_sources_kernel<<<(count + 127) / 128, 128, 0, s>>>(dc, dm, dp, count, dre, dim);
            ok = !hip_failed(hipGetLastError(), "test") &&
                 !hip_failed(hipMemcpyAsync(out_re, dre, 4 * (size_t)count * 8, hipMemcpyDeviceToHost, s), "test") &&
                 !hip_failed(hipMemcpyAsync(out_im, dim, 4 * (size_t)count * 8, hipMemcpyDeviceToHost, s), "test") &&
                 !hip_failed(hipStreamSynchronize(s), "test");
        }
    }
    (void)hipFree(dc); (void)hipFree(dm); (void)hipFree(dp); (void)hipFree(dre); (void)hipFree(dim);
    release_runner_lock();
    return ok ? PACMAN_RECURRENCE_OK : PACMAN_RECURRENCE_DEVICE_ERROR;
}

extern "C" void *pacman_recurrence_v1_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

extern "C" void pacman_recurrence_v1_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

extern "C" int pacman_recurrence_v1_reduce_sample(const pacman_recurrence_reduce_sample *r) {
    acquire_runner_lock();
    int status = reduce_sample_locked(r);
    release_runner_lock();
    return status;
}

extern "C" int pacman_recurrence_v1_reduce_resolved(const pacman_recurrence_reduce_resolved *r) {
    acquire_runner_lock();
    int status = reduce_resolved_locked(r);
    release_runner_lock();
    return status;
}

extern "C" int pacman_recurrence_v1_reduce(const pacman_recurrence_reduce *r) {
    acquire_runner_lock();
    int status = reduce_locked(r);
    release_runner_lock();
    return status;
}
