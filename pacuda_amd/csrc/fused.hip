// SPDX-License-Identifier: 0BSD
//
// The fused evaluation (include/pacuda_amd.h): whole plan runs of one or many plans in one
// persistent launch.
//
// Layout in HBM: the external momenta stay where the caller put them (SoA planes of points);
// each resident wavefront owns a scratch slot holding its arena for 64 points, plane-major with
// one vector per lane and plane: current planes (a complex value, 16 bytes per lane; 8 in the
// mixed kernels), amplitude planes (16 bytes), momentum forms (four doubles, 32 bytes). A lane's
// access is one load at a uniform plane base plus its lane offset. A slot is reused for every
// 64-point unit the wavefront takes, so the arena's footprint is (resident wavefronts x slot
// size), not (points x plan size); HBM carries the momenta in and |M|^2 out, plus whatever of
// the slots the L2 and the MALL do not hold.
//
// Work queue: the items' 64-point units are numbered item by item; a wavefront takes the next
// unit with one atomic add of lane 0, so plans of uneven length and items of uneven size
// balance across the compute units. The grid is sized to the resident wavefronts only; every
// wavefront leaves when the counter passes the last unit.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <array>
#include <unordered_map>
#include <vector>

#include "pacman_recurrence.h"
#include "pacuda_amd.h"
#include "plan_compiler.hpp"
#include "coop.hpp"
#include "program.hpp"
#include "runtime.hpp"

using namespace pacuda;

namespace {

constexpr int LDS_VARIANTS = 5;
// the cooperative shapes (wavefronts per unit, units per compute unit): one 64-point unit per
// workgroup, LDS_BYTES / units per CU of LDS for it (its top COOP_CONTROL bytes hold the unit
// number), so (LDS_BYTES / units - COOP_CONTROL) / 1 KB planes of 64 f64 values
constexpr int COOP_SHAPES = 4;
constexpr uint32_t COOP_WAVES[COOP_SHAPES] = {16, 8, 8, 4};
constexpr uint32_t COOP_UNITS[COOP_SHAPES] = {1, 1, 2, 4};
constexpr uint32_t COOP_CONTROL = 256;
// the cooperative kernels built: f64 in every shape, the other policies in shapes 0 and 2 (per
// class of current, 0 for 16-byte currents and 1 for the f32 ones, the programs they need)
constexpr bool coop_built(int cls, int shape) { return cls == 0 || shape == 0 || shape == 2; }
// the default launch's shape for the kernels whose cooperative launch is the faster one (f64,
// mixed, mixed64, compensated64: 8 wavefronts per unit, two units per CU; the df64-arithmetic
// kernels are VALU-bound and keep one wavefront per unit)
constexpr int COOP_DEFAULT = 2;
static_assert(COOP_WAVES[COOP_DEFAULT] == 8 && COOP_UNITS[COOP_DEFAULT] == 2, "the default cooperative shape");
// ... and, when the default shape's programs would move more than this many bytes per point
// through the slot (averaged over the launch's points), shape 0: one unit per CU with 16
// wavefronts and twice the LDS per unit (measured: g g > g g g g in f64, 90 KB per point at
// shape 2, runs 7% faster in shape 0; its f32 currents, 37 KB, and g g > g g g, 22 KB, run
// faster in shape 2)
constexpr uint64_t COOP_WIDE_BYTES = 64 * 1024;

struct PlanDev {
    const Instr *prog_lds[LDS_VARIANTS];  // with forwarding and LDS-resident currents (LDS_CAPS[v] planes)
    const Instr *prog_reg[LDS_VARIANTS];  // the same with a register tier of REG_PLANES planes after them
    const Instr *prog_rlds[LDS_VARIANTS], *prog_rreg[LDS_VARIANTS];  // the f64 kernel's (sources made again)
    uint32_t count_rlds[LDS_VARIANTS], count_rreg[LDS_VARIANTS];     // their instruction counts
    uint32_t reg_forms[FORM_REGS];        // the form each register slot holds (UINT32_MAX: none), every program
    uint32_t zero_fill;                   // 1: some amplitude plane has no F_ZERO closure, so all start zeroed
    const Instr *prog_fwd;                // with forwarding only (launches that give the wavefronts no LDS)
    const Instr *prog_plain;              // neither (kernels without forwarding)
    uint32_t count, count_plain, sources;
    uint32_t lds_planes[LDS_VARIANTS];    // LDS planes each variant uses per wavefront (0: no variant)
    const double *c64;  // scales re, im, factors re, im (f64)
    const double *cdf;  // the same in the df64 layout
    uint32_t scale_count, factor_count;
    uint32_t current_planes, amplitude_planes, form_count, outputs;
    const uint32_t *form_terms, *term_slots;
    const double *coefficients;
    const uint32_t *offsets, *planes;
    const double *weights;
    const uint32_t *bindings;             // the program's symjit bindings
    double numerator, denominator;
    uint32_t factor, pad;
    // the cooperative programs per class of current (0: 16 bytes, f64 and df64; 1: 8, the f32
    // ones) and shape (coop_prog null: none for the plan)
    const Instr *coop_prog[2][COOP_SHAPES];     // the wavefronts' streams, one after another
    const uint32_t *coop_start[2][COOP_SHAPES];  // their offsets (wavefronts + 1)
    const uint32_t *coop_split[2][COOP_SHAPES];  // per wavefront: where its first phase ends
    const uint32_t *coop_red[2][COOP_SHAPES];    // per term of the reduction (`planes`): its amplitude's field
    const uint32_t *coop_form[2][COOP_SHAPES];   // per form: FORM_LDS | LDS plane, or the slot form
    uint32_t coop_slot[2][COOP_SHAPES];          // slot current planes
};

struct ItemDev {
    const PlanDev *plan;
    const double *external;
    double *resolved, *total;
    uint32_t points, ext_stride;
    // a v1 run on these kernels (amp_re not null): instead of the reduction, the amplitude planes
    // the reduction names (`planes`, one per output) go to these plane arrays, amp_stride values
    // apart, in the kernel's amplitude layout
    double *amp_re, *amp_im;
    size_t amp_stride;
};

constexpr int LANES = 64;
// LDS-resident current planes per wavefront of the program variants: a launch takes the largest
// whose LDS fits the wavefronts it asks per CU (f64, 256-thread blocks: 40 at 4 wavefronts, 20 at
// 8, 13 at 12, 10 at 16; 64-thread blocks: 26 at 6)
constexpr uint32_t LDS_CAPS[LDS_VARIANTS] = {10, 13, 20, 26, 40};
// (a dual field names its on-chip slot in six bits: LDS and register-tier slots below 64)
static_assert(LDS_CAPS[LDS_VARIANTS - 1] + REG_PLANES <= 64, "on-chip slots of dual fields");
constexpr uint32_t LDS_BYTES = 160 * 1024;  // LDS per CU (gfx950), the most one workgroup may take

// an amplitude's raw value as a double: the df64 kernels' pairs are summed in f64
template <bool DF> PCD double amplitude_value(double raw) {
    if constexpr (DF) {
        const flt2 v = __builtin_bit_cast(flt2, raw);
        return (double)v.x + (double)v.y;
    } else {
        return raw;
    }
}

__global__ void reset_counter(uint32_t *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

// The kernels' register tiers: TIER 2 keeps REG_PLANES current planes and four forms in
// registers, built for two wavefronts per SIMD in the f64, df64 and compensated policies (up to
// 256 VGPRs) and three in the mixed ones (168 VGPRs); TIER 1 keeps only the FORM_REGS forms, for
// three wavefronts (f64, df64, compensated) or four (mixed); TIER 0 keeps neither, for four (128
// VGPRs). (The compensated policy's df64 vertices and accumulators need the wide budgets.)
template <class K> constexpr bool f32_currents() { return __is_same(typename K::real, float); }
template <class K> constexpr bool narrow_registers() { return f32_currents<K>() && !compensated<K>(); }
template <class K, int TIER> constexpr int waves_per_simd() {
    return TIER == 2 ? (narrow_registers<K>() ? 3 : 2) : TIER == 1 ? (narrow_registers<K>() ? 4 : 3) : 4;
}
template <class K, int TIER> constexpr int register_forms() {
    return TIER == 2 ? 4 : TIER == 1 ? (int)FORM_REGS : 0;
}

template <class K, int TIER, bool SYMJIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_simd<K, TIER>())))
fused_kernel(const ItemDev *__restrict__ items, const uint32_t *__restrict__ unit_start, uint32_t item_count,
             uint32_t *counter, double *scratch, size_t slot_doubles, uint32_t lds_per_wave, uint32_t lds_variant,
             uint32_t lds_cap) {
    typedef typename K::real R;
    // df64 pairs for the program's constants (all but the f64-arithmetic policies) and for the
    // amplitude planes
    constexpr bool DF = !__is_same(typename K::acc, double);
    constexpr bool DF_AMP = !__is_same(typename K::amp, double);
    constexpr bool DF_FORMS = __is_same(typename K::den, D);
    const uint32_t lane = threadIdx.x & (LANES - 1);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / LANES) + threadIdx.x / LANES);
    double *slot = scratch + (size_t)wave * slot_doubles;
    const uint32_t units = cnst(unit_start)[item_count];
    for (;;) {
        uint32_t taken = 0;
        if (lane == 0) taken = atomicAdd(counter, 1u);
        const uint32_t unit = __builtin_amdgcn_readfirstlane(taken);
        if (unit >= units) break;
        // the item of this unit (uniform binary search)
        uint32_t lo = 0, hi = item_count;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) / 2;
            if (cnst(unit_start)[mid] <= unit) lo = mid;
            else hi = mid;
        }
        const ItemDev it = cnst(items)[lo];
        const PlanDev pl = *cnst(it.plan);
        const uint32_t i = (unit - cnst(unit_start)[lo]) * LANES + lane;
        const bool active = i < it.points;
        const size_t ie = active ? i : it.points - 1;  // inactive lanes repeat the last point
        const uint32_t C = pl.current_planes, A = pl.amplitude_planes, F = pl.form_count;
        typedef typename Packed<R>::T V;
        // forms in registers where the kernel has room for them (up to three wavefronts per SIMD)
        constexpr bool REG = TIER == 2;
        constexpr int NF = register_forms<K, TIER>();
        SlotArena<R, REG, NF> a;
        RegFile<R> regs;
        FormRegs<NF ? NF : 1> fregs;
        a.regs = &regs;
        a.fregs = &fregs;
        a.lds_cap = lds_cap;
        a.slot = __builtin_amdgcn_make_buffer_rsrc(slot, (short)0, (int)(slot_doubles * sizeof(double)), BUFFER_FORMAT);
        a.amp_base = C * 1024u;
        a.form_base = a.amp_base + A * 1024u;
        a.idx = lane;
        a.bind_ = pl.bindings;
        a.point_ = i;
        a.points_ = it.points;
        a.lds = (threadIdx.x / LANES) * lds_per_wave * (uint32_t)(sizeof(V) / 2) + lane * (uint32_t)sizeof(V);
        // the momentum forms in f64 (runner/recurrence_runner.cu:745-764), four per lane
        for (uint32_t f = 0; f < F; f++) {
            const uint32_t first = cnst(pl.form_terms)[2 * f], n = cnst(pl.form_terms)[2 * f + 1];
            double v[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                v[k] = 0.0;
                for (uint32_t t = 0; t < n; t++) {
                    double term = cnst(pl.coefficients)[first + t] *
                                  it.external[((size_t)cnst(pl.term_slots)[first + t] * 4 + k) * it.ext_stride + ie];
                    v[k] = t == 0 ? 0.0 + term : v[k] + term;
                }
            }
            a.put_form(f, v);
        }
        if (pl.zero_fill || !K::forwarding)
            for (uint32_t q = 0; q < A; q++) a.template put_amp<double>(q, cx(0.0, 0.0));
        // sources in f64 from the f64 constants, then the program in the kernel's layout
        const uint32_t S = pl.scale_count, X = pl.factor_count;
        a.sre_ = pl.c64; a.sim_ = pl.c64 + S; a.fre_ = pl.c64 + 2 * S; a.fim_ = pl.c64 + 2 * S + X;
        // (selects, not an index: a runtime index into the local copy of the plan would put it in scratch)
        constexpr bool REMAKE = __is_same(R, double) && __is_same(typename K::den, double);
        const Instr *prog;
        uint32_t count = pl.count;
        if constexpr (REMAKE && REG) {
            prog = !lds_per_wave       ? pl.prog_fwd
                   : lds_variant == 0 ? pl.prog_rreg[0]
                   : lds_variant == 1 ? pl.prog_rreg[1]
                   : lds_variant == 2 ? pl.prog_rreg[2]
                   : lds_variant == 3 ? pl.prog_rreg[3]
                                      : pl.prog_rreg[4];
            count = !lds_per_wave       ? pl.count
                    : lds_variant == 0 ? pl.count_rreg[0]
                    : lds_variant == 1 ? pl.count_rreg[1]
                    : lds_variant == 2 ? pl.count_rreg[2]
                    : lds_variant == 3 ? pl.count_rreg[3]
                                       : pl.count_rreg[4];
        } else if constexpr (REMAKE) {
            prog = !lds_per_wave       ? pl.prog_fwd
                   : lds_variant == 0 ? pl.prog_rlds[0]
                   : lds_variant == 1 ? pl.prog_rlds[1]
                   : lds_variant == 2 ? pl.prog_rlds[2]
                   : lds_variant == 3 ? pl.prog_rlds[3]
                                      : pl.prog_rlds[4];
            count = !lds_per_wave       ? pl.count
                    : lds_variant == 0 ? pl.count_rlds[0]
                    : lds_variant == 1 ? pl.count_rlds[1]
                    : lds_variant == 2 ? pl.count_rlds[2]
                    : lds_variant == 3 ? pl.count_rlds[3]
                                       : pl.count_rlds[4];
        } else if constexpr (REG) {
            prog = !lds_per_wave       ? pl.prog_fwd
                   : lds_variant == 0 ? pl.prog_reg[0]
                   : lds_variant == 1 ? pl.prog_reg[1]
                   : lds_variant == 2 ? pl.prog_reg[2]
                   : lds_variant == 3 ? pl.prog_reg[3]
                                      : pl.prog_reg[4];
        } else {
            prog = !lds_per_wave       ? pl.prog_fwd
                   : lds_variant == 0 ? pl.prog_lds[0]
                   : lds_variant == 1 ? pl.prog_lds[1]
                   : lds_variant == 2 ? pl.prog_lds[2]
                   : lds_variant == 3 ? pl.prog_lds[3]
                                      : pl.prog_lds[4];
        }
        run_sources_point<R>(prog, pl.sources, a);
        if constexpr (DF_FORMS) {
            for (uint32_t f = 0; f < F; f++) {
                double raw[4];
                a.form(f, raw);
#pragma unroll
                for (int k = 0; k < 4; k++) raw[k] = to_raw(narrow<D>(raw[k]));
                a.put_form(f, raw);
            }
        }
        // the program's constants: df64 pairs for every kernel but f64 (as the v1 path converts them)
        const double *cons = DF ? pl.cdf : pl.c64;
        a.sre_ = cons; a.sim_ = cons + S; a.fre_ = cons + 2 * S; a.fim_ = cons + 2 * S + X;
        if constexpr (NF > 0) {  // the register-resident forms, in the kernel's form layout
#pragma unroll
            for (uint32_t q = 0; q < (uint32_t)NF; q++) {
                if (pl.reg_forms[q] < F) {
                    double raw[4];
                    a.form(pl.reg_forms[q], raw);
                    fregs.set(q, raw);
                }
            }
        }
        if constexpr (K::forwarding)
            interpret<K, SYMJIT>(prog + pl.sources, count - pl.sources, a);
        else
            interpret<K, SYMJIT>(pl.prog_plain + pl.sources, pl.count_plain - pl.sources, a);
        if (it.amp_re) {  // a v1 run: the amplitude planes, for the v1 reductions
            if (active)
                for (uint32_t k = 0; k < pl.outputs; k++) {
                    const uint32_t q = cnst(pl.planes)[k];
                    const Cx<double> amp = a.template amp<double>(q);
                    it.amp_re[(size_t)q * it.amp_stride + i] = amp.re;
                    it.amp_im[(size_t)q * it.amp_stride + i] = amp.im;
                }
            continue;
        }
        // the reduction (runner/recurrence_runner.cu:1511-1527) and the helicity sum
        double t = 0.0;
        for (uint32_t o = 0; o < pl.outputs; o++) {
            double acc = 0.0;
            for (uint32_t k = cnst(pl.offsets)[o]; k < cnst(pl.offsets)[o + 1]; k++) {
                const Cx<double> amp = a.template amp<double>(cnst(pl.planes)[k]);
                double re = amplitude_value<DF_AMP>(amp.re), im = amplitude_value<DF_AMP>(amp.im);
                acc = acc + cnst(pl.weights)[k] * __builtin_fma(re, re, im * im);
            }
            if (it.resolved && active) it.resolved[(size_t)i * pl.outputs + o] = acc;
            double v = pl.factor ? (acc * pl.numerator) / pl.denominator : acc;
            t = t + v;
        }
        if (it.total && active) it.total[i] = t;
    }
}

// The cooperative kernel: each workgroup takes one 64-point unit at a time off the same work
// queue, and its W wavefronts share it. Per unit: the momentum forms (form f by wavefront f mod W,
// into the unit's LDS region or slot), a barrier, each wavefront's stream of the plan's
// cooperative program (whose barrier rows separate its phases; currents, forms and amplitudes in
// LDS or in the workgroup's slot), a barrier, then wavefront 0 reduces the amplitudes as
// fused_kernel does while the others (and then wavefront 0) compute the next unit's forms: the
// cooperative programs never give an amplitude a form's LDS room, and when the next unit belongs
// to another plan a barrier ends the reduction first. A fetch row in one wavefront's stream takes
// the next unit off the queue while the current one runs. Every precision policy: the first
// phase (the sources) runs with the f64 constants, as fused_kernel runs its sources, and the
// policies whose denominators are df64 then rewrite the forms in df64 between two barriers.
template <class K, int SHAPE, bool SYMJIT, int W = COOP_WAVES[SHAPE], int UNITS = COOP_UNITS[SHAPE]>
__global__ void __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(W * UNITS / 4)))
coop_kernel(const ItemDev *__restrict__ items, const uint32_t *__restrict__ unit_start, uint32_t item_count,
            uint32_t *counter, double *scratch, size_t slot_doubles, uint32_t lds_bytes) {
    typedef typename K::real R;
    typedef typename Packed<R>::T V;
    constexpr int CLASS = sizeof(V) == 16 ? 0 : 1;  // the programs of 16-byte or 8-byte currents
    constexpr bool DF = !__is_same(typename K::acc, double);    // df64 constants after the sources
    constexpr bool DF_AMP = !__is_same(typename K::amp, double);
    constexpr bool DF_FORMS = __is_same(typename K::den, D);
    extern __shared__ __attribute__((aligned(16))) char pacuda_lds[];
    const uint32_t lane = threadIdx.x & (LANES - 1);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / LANES);
    volatile uint32_t *control = reinterpret_cast<volatile uint32_t *>(pacuda_lds + lds_bytes - COOP_CONTROL);
    const uint32_t units = cnst(unit_start)[item_count];
    SlotArena<R, false, 0, true> a;
    a.regs = nullptr;
    a.fregs = nullptr;
    a.lds_cap = 0;
    a.slot = __builtin_amdgcn_make_buffer_rsrc(scratch + (size_t)blockIdx.x * slot_doubles, (short)0,
                                               (int)(slot_doubles * sizeof(double)), BUFFER_FORMAT);
    a.idx = lane;
    a.bind_ = nullptr;
    a.lds = lane * (uint32_t)sizeof(V);
    a.counter = counter;
    a.next_unit = control + 1;
    // the item of a unit (uniform binary search) and the lane's point in it
    auto locate = [&](uint32_t u, ItemDev &it, uint32_t &i) {
        uint32_t lo = 0, hi = item_count;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) / 2;
            if (cnst(unit_start)[mid] <= u) lo = mid;
            else hi = mid;
        }
        it = cnst(items)[lo];
        i = (u - cnst(unit_start)[lo]) * LANES + lane;
    };
    // the arena's regions for a unit of `plan`, then the unit's momentum forms in f64
    // (runner/recurrence_runner.cu:745-764), form f by wavefront f mod W
    auto forms = [&](const ItemDev &it, uint32_t i) {
        const PlanDev *plan = it.plan;
        a.amp_base = cnst(plan)->coop_slot[CLASS][SHAPE] * 1024u;
        a.form_base = a.amp_base + cnst(plan)->amplitude_planes * 1024u;
        const size_t ie = i < it.points ? i : it.points - 1;  // inactive lanes repeat the last point
        const uint32_t F = cnst(plan)->form_count;
        const uint32_t *form_terms = cnst(plan)->form_terms, *term_slots = cnst(plan)->term_slots;
        const uint32_t *form_field = cnst(plan)->coop_form[CLASS][SHAPE];
        const double *coefficients = cnst(plan)->coefficients;
        for (uint32_t f = wave; f < F; f += W) {
            const uint32_t first = cnst(form_terms)[2 * f], n = cnst(form_terms)[2 * f + 1];
            double v[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                v[k] = 0.0;
                for (uint32_t t = 0; t < n; t++) {
                    double term = cnst(coefficients)[first + t] *
                                  it.external[((size_t)cnst(term_slots)[first + t] * 4 + k) * it.ext_stride + ie];
                    v[k] = t == 0 ? 0.0 + term : v[k] + term;
                }
            }
            a.put_form(cnst(form_field)[f], v);
        }
    };
    if (threadIdx.x == 0) control[0] = atomicAdd(counter, 1u);
    __syncthreads();
    uint32_t unit = __builtin_amdgcn_readfirstlane(control[0]);
    ItemDev it;
    uint32_t i = 0;
    if (unit < units) {
        locate(unit, it, i);
        forms(it, i);
    }
    __syncthreads();
    while (unit < units) {
        const PlanDev *plan = it.plan;
        a.point_ = i;
        a.points_ = it.points;
        a.bind_ = cnst(plan)->bindings;
        const double *c64 = cnst(plan)->c64, *cons = DF ? cnst(plan)->cdf : c64;
        const uint32_t S = cnst(plan)->scale_count, X = cnst(plan)->factor_count;
        a.sre_ = c64; a.sim_ = c64 + S; a.fre_ = c64 + 2 * S; a.fim_ = c64 + 2 * S + X;
        const Instr *prog = cnst(plan)->coop_prog[CLASS][SHAPE];
        const uint32_t first = cnst(cnst(plan)->coop_start[CLASS][SHAPE])[wave];
        const uint32_t end = cnst(cnst(plan)->coop_start[CLASS][SHAPE])[wave + 1];
        const uint32_t mid = cnst(cnst(plan)->coop_split[CLASS][SHAPE])[wave];  // (a barrier row)
        // (one wavefront's stream takes the next unit off the queue: a fetch row in the phase that
        // hides the atomic's round trip best)
        interpret<K, SYMJIT, true>(prog + first, mid - first, a);  // the sources, in f64
        a.sre_ = cons; a.sim_ = cons + S; a.fre_ = cons + 2 * S; a.fim_ = cons + 2 * S + X;
        if constexpr (DF_FORMS) {  // the forms in df64 once no source reads them
            __syncthreads();  // (the barrier row at `mid`)
            const uint32_t F = cnst(plan)->form_count;
            const uint32_t *form_field = cnst(plan)->coop_form[CLASS][SHAPE];
            for (uint32_t f = wave; f < F; f += W) {
                double raw[4];
                a.form(cnst(form_field)[f], raw);
#pragma unroll
                for (int c = 0; c < 4; c++) raw[c] = to_raw(narrow<D>(raw[c]));
                a.put_form(cnst(form_field)[f], raw);
            }
            __syncthreads();
            interpret<K, SYMJIT, true>(prog + mid + 1, end - mid - 1, a);
        } else {
            interpret<K, SYMJIT, true>(prog + mid, end - mid, a);
        }
        __syncthreads();
        // (wavefront 0's reduction is the unit's critical path: it issues first)
        if (wave == 0) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(0);
        const uint32_t next = __builtin_amdgcn_readfirstlane(control[1]);
        ItemDev nit = it;
        uint32_t ni = 0;
        if (next < units) locate(next, nit, ni);
        if (wave == 0 && it.amp_re) {  // a v1 run: the amplitude planes, for the v1 reductions
            const uint32_t outputs = cnst(plan)->outputs;
            const uint32_t *planes = cnst(plan)->planes, *fields = cnst(plan)->coop_red[CLASS][SHAPE];
            if (i < it.points)
                for (uint32_t k = 0; k < outputs; k++) {
                    const Cx<double> amp = a.template amp<double>(cnst(fields)[k]);
                    const size_t at = (size_t)cnst(planes)[k] * it.amp_stride + i;
                    it.amp_re[at] = amp.re;
                    it.amp_im[at] = amp.im;
                }
        } else if (wave == 0) {
            // the reduction (runner/recurrence_runner.cu:1511-1527) and the helicity sum
            const uint32_t outputs = cnst(plan)->outputs;
            const uint32_t *offsets = cnst(plan)->offsets, *fields = cnst(plan)->coop_red[CLASS][SHAPE];
            const double *weights = cnst(plan)->weights;
            const bool factor = cnst(plan)->factor;
            const double numerator = cnst(plan)->numerator, denominator = cnst(plan)->denominator;
            const bool active = i < it.points;
            double t = 0.0;
            for (uint32_t o = 0; o < outputs; o++) {
                double acc = 0.0;
                for (uint32_t k = cnst(offsets)[o]; k < cnst(offsets)[o + 1]; k++) {
                    const Cx<double> amp = a.template amp<double>(cnst(fields)[k]);
                    const double re = amplitude_value<DF_AMP>(amp.re), im = amplitude_value<DF_AMP>(amp.im);
                    acc = acc + cnst(weights)[k] * __builtin_fma(re, re, im * im);
                }
                if (it.resolved && active) it.resolved[(size_t)i * outputs + o] = acc;
                double v = factor ? (acc * numerator) / denominator : acc;
                t = t + v;
            }
            if (it.total && active) it.total[i] = t;
        }
        if (next < units) {
            if (nit.plan != plan) __syncthreads();  // (another plan's rooms: after the reduction)
            forms(nit, ni);
        }
        __syncthreads();
        unit = next;
        it = nit;
        i = ni;
    }
}

}  // namespace

struct pacuda_plan {
    PlanDev host;         // with device pointers
    PlanDev *device = nullptr;
    std::vector<void *> allocations;
    uint64_t slot_doubles = 0;
    uint64_t coop_slot_doubles[2][COOP_SHAPES] = {};  // a workgroup's slot per class and shape (0: no program)
    uint64_t coop_slot_bytes[2][COOP_SHAPES] = {};    // the bytes a point moves through it
    bool symjit = false;  // the program holds symjit rows
};

namespace {

template <class T> T *upload(pacuda_plan *p, const T *src, size_t n, bool *ok) {
    if (!*ok) return nullptr;
    T *d = nullptr;
    if (hip_failed(hipMalloc((void **)&d, std::max<size_t>(n, 1) * sizeof(T)), "plan") ||
        (n && hip_failed(hipMemcpy(d, src, n * sizeof(T), hipMemcpyHostToDevice), "plan"))) {
        *ok = false;
        return nullptr;
    }
    p->allocations.push_back(d);
    return d;
}

// every current plane the program reads has been written by it before
bool reads_after_writes(const std::vector<Instr> &ops, uint32_t planes, const std::vector<uint32_t> *bindings = nullptr) {
    std::vector<char> w(planes, 0);
    auto written = [&](uint32_t base, uint32_t n) {
        for (uint32_t k = 0; k < n; k++)
            if ((uint64_t)base + k >= planes || !w[base + k]) return false;
        return true;
    };
    auto mark = [&](uint32_t base, uint32_t n) {
        for (uint32_t k = 0; k < n; k++)
            if ((uint64_t)base + k < planes) w[base + k] = 1;
    };
    for (const Instr &o : ops) {
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff, flags = o.w[0] >> 16;
        if (kind == K_SOURCE) {
            uint32_t family = o.w[5] & 7;
            mark(o.w[1], family == 0 ? 1 : family == 1 ? 2 : 4);
        } else if (kind == K_CLEAR) {
            mark(o.w[1], o.w[2]);
        } else if (kind == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(op, &c0, &c1, &cd);
            if (!written(o.w[2], c0) || !written(o.w[3], c1)) return false;
            if ((flags & F_FIRST) && !(flags & F_INIT) && !written(o.w[4], cd)) return false;
            if (flags & F_LAST) mark(o.w[4], cd);
        } else if (kind == K_PROP) {
            if (!written(o.w[1], op == 9 ? 4 : 2)) return false;
        } else if (kind == K_IDENTITY) {
            if (!written(o.w[1], o.w[2])) return false;
        } else if (kind == K_CLOSURE) {
            if (!written(o.w[1], o.w[6]) || !written(o.w[2], o.w[6])) return false;
        } else if (kind == K_SYMJIT) {
            // the components its binding's codes read, then its outputs (read first when the row
            // does not initialise its destination)
            if (!bindings) return false;
            const std::vector<uint32_t> &b = *bindings;
            const size_t at = o.w[6];
            if (at + 4 > b.size()) return false;
            const uint32_t inputs = b[at + 1], pairs = b[at + 2];
            if (at + 4 + inputs + pairs > b.size()) return false;
            for (uint32_t k = 0; k < inputs; k++) {
                const uint32_t code = b[at + 4 + k], which = (code >> 3) & 1, index = code >> 5;
                if ((code & 7) == SB_PARENT && !written((which ? o.w[2] : o.w[1]) + index, 1)) return false;
                if ((code & 7) == SB_DESTINATION && !written(o.w[3] + index, 1)) return false;
            }
            for (uint32_t j = 0; j < pairs; j++) {
                const uint32_t plane = o.w[3] + b[at + 4 + inputs + j];
                if (!(flags & F_INIT) && !written(plane, 1)) return false;
                mark(plane, 1);
            }
        }
    }
    return true;
}

// Arena compaction: currents are objects (a plane range written as a unit: a source, a
// contribution destination, a clear); an object lives from the row that first writes it to the
// last row that touches it. Objects get physical plane ranges first-fit from a free list as the
// program runs, so the slot holds only what is live at once. Values are unchanged (only plane
// indices are rewritten). Every access must name a whole object; otherwise the program keeps
// its planes. Returns the compacted current plane count (0: not compacted).
uint32_t compact_planes(std::vector<Instr> &ops, uint32_t planes) {
    struct Obj { uint32_t base, len, first, last, phys; };
    std::vector<int> owner(planes, -1);  // plane -> object (by its base)
    std::vector<Obj> objs;
    auto object_at = [&](uint32_t base, uint32_t len, uint32_t row, bool write) -> int {
        if (base + len > planes || len == 0) return -2;
        int o = owner[base];
        if (o >= 0) {
            if (objs[o].base != base || objs[o].len != len) return -2;
            objs[o].last = row;
            return o;
        }
        if (!write) return -2;
        for (uint32_t k = 0; k < len; k++)
            if (owner[base + k] >= 0) return -2;
        objs.push_back({base, len, row, row, 0});
        for (uint32_t k = 0; k < len; k++) owner[base + k] = (int)objs.size() - 1;
        return (int)objs.size() - 1;
    };
    // accesses per row: (operand slot index in w[], length, write?)
    struct Acc { int word; uint32_t len; bool write; };
    std::vector<std::vector<Acc>> access(ops.size());
    for (size_t n = 0; n < ops.size(); n++) {
        const Instr &o = ops[n];
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff, flags = o.w[0] >> 16;
        auto &a = access[n];
        if (kind == K_SOURCE) {
            uint32_t family = o.w[5] & 7;
            a.push_back({1, family == 0 ? 1u : family == 1 ? 2u : 4u, true});
        } else if (kind == K_CLEAR) {
            a.push_back({1, o.w[2], true});
        } else if (kind == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(op, &c0, &c1, &cd);
            a.push_back({2, (uint32_t)c0, false});
            a.push_back({3, (uint32_t)c1, false});
            a.push_back({4, (uint32_t)cd, (flags & F_FIRST) && (flags & F_INIT)});
        } else if (kind == K_PROP) {
            a.push_back({1, op == 9 ? 4u : 2u, false});
        } else if (kind == K_IDENTITY) {
            a.push_back({1, o.w[2], false});
        } else if (kind == K_CLOSURE) {
            a.push_back({1, o.w[6], false});
            a.push_back({2, o.w[6], false});
        } else {
            return 0;
        }
        for (const Acc &x : a)
            if (x.len && object_at(o.w[x.word], x.len, (uint32_t)n, x.write) < 0) return 0;
    }
    // first-fit allocation in row order
    std::vector<std::pair<uint32_t, uint32_t>> free_list;  // (start, length), sorted by start
    uint32_t top = 0;
    std::vector<std::vector<int>> born(ops.size()), dies(ops.size());
    for (size_t k = 0; k < objs.size(); k++) {
        born[objs[k].first].push_back((int)k);
        dies[objs[k].last].push_back((int)k);
    }
    auto release = [&](uint32_t start, uint32_t len) {
        free_list.push_back({start, len});
        std::sort(free_list.begin(), free_list.end());
        std::vector<std::pair<uint32_t, uint32_t>> merged;
        for (auto &f : free_list) {
            if (!merged.empty() && merged.back().first + merged.back().second == f.first) merged.back().second += f.second;
            else merged.push_back(f);
        }
        free_list.swap(merged);
    };
    for (size_t n = 0; n < ops.size(); n++) {
        if (n > 0)
            for (int k : dies[n - 1]) release(objs[k].phys, objs[k].len);
        for (int k : born[n]) {
            uint32_t len = objs[k].len;
            bool placed = false;
            for (auto &f : free_list)
                if (f.second >= len) {
                    objs[k].phys = f.first;
                    f.first += len;
                    f.second -= len;
                    placed = true;
                    break;
                }
            if (!placed) {
                objs[k].phys = top;
                top += len;
            }
            free_list.erase(std::remove_if(free_list.begin(), free_list.end(),
                                           [](const std::pair<uint32_t, uint32_t> &f) { return f.second == 0; }),
                            free_list.end());
        }
    }
    // replay the placement: a read must find its own object in every physical plane, and a new
    // object may only take planes whose previous holder is dead
    std::vector<int> held(top, -1);
    for (size_t n = 0; n < ops.size(); n++) {
        for (int pass = 0; pass < 2; pass++)
            for (const Acc &x : access[n]) {
                if (!x.len) continue;
                const int k = owner[ops[n].w[x.word]];
                const bool birth = objs[k].first == n;
                if (birth != (pass == 1)) continue;
                for (uint32_t j = 0; j < x.len; j++) {
                    int &h = held[objs[k].phys + j];
                    if (birth) {
                        if (h >= 0 && h != k && objs[h].last >= n) return 0;
                        h = k;
                    } else if (h != k) {
                        return 0;
                    }
                }
            }
    }
    for (size_t n = 0; n < ops.size(); n++)
        for (const Acc &x : access[n])
            if (x.len) ops[n].w[x.word] = objs[owner[ops[n].w[x.word]]].phys;
    return top;
}

// The current-plane ranges row o touches: operand word, length, and whether the row creates the
// range (a source, a clear, the first row of an initialising run). -1 for a kind not known here.
struct Access { int word; uint32_t len; bool write; };
int row_accesses(const Instr &o, Access *a) {
    const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff, flags = o.w[0] >> 16;
    if (kind == K_SOURCE) {
        const uint32_t family = o.w[5] & 7;
        a[0] = {1, family == 0 ? 1u : family == 1 ? 2u : 4u, true};
        return 1;
    }
    if (kind == K_CLEAR) { a[0] = {1, o.w[2], true}; return 1; }
    if (kind == K_CONTRIB) {
        int c0, c1, cd;
        contribution_shape(op, &c0, &c1, &cd);
        a[0] = {2, (uint32_t)c0, false};
        a[1] = {3, (uint32_t)c1, false};
        a[2] = {4, (uint32_t)cd, (flags & F_FIRST) && (flags & F_INIT)};
        return 3;
    }
    if (kind == K_PROP) { a[0] = {1, op == 9 ? 4u : 2u, false}; return 1; }
    if (kind == K_IDENTITY) { a[0] = {1, o.w[2], false}; return 1; }
    if (kind == K_CLOSURE) { a[0] = {1, o.w[6], false}; a[1] = {2, o.w[6], false}; return 2; }
    return -1;
}

// Row blocks and their dependences. The rows form blocks: each leading source, each contribution
// run (FIRST to LAST on one destination: its partial sums live in the accumulators, so it stays
// whole), each other row. A block depends on the earlier blocks that write a plane it reads or
// writes, or read a plane it writes, and closures into one amplitude keep their order; every
// order that honours these dependences runs the same operations on the same operands, and adds
// the same terms to each destination in the same order, so every value keeps its bits. The
// leading sources stay in front, in order (the kernels run them as a phase of their own).
struct BlockGraph {
    struct Block {
        uint32_t first, count;
        std::vector<uint32_t> reads, writes;  // plane indices
        int amplitude = -1;
    };
    std::vector<Block> blocks;
    std::vector<std::vector<uint32_t>> succ, pred;
    uint32_t lead = 0;  // the leading source blocks
    bool ok = false;
};

BlockGraph block_graph(const std::vector<Instr> &ops, uint32_t sources, uint32_t planes) {
    BlockGraph g;
    auto &blocks = g.blocks;
    for (size_t n = 0; n < ops.size();) {
        const int kind = ops[n].w[0] & 0xff;
        BlockGraph::Block b;
        b.first = (uint32_t)n;
        size_t m = n;
        if (kind == K_CONTRIB) {
            if (!((ops[n].w[0] >> 16) & F_FIRST)) return g;
            while (m < ops.size() && (ops[m].w[0] & 0xff) == K_CONTRIB && ops[m].w[4] == ops[n].w[4] &&
                   !((ops[m].w[0] >> 16) & F_LAST))
                m++;
            if (m == ops.size() || (ops[m].w[0] & 0xff) != K_CONTRIB || ops[m].w[4] != ops[n].w[4]) return g;
        }
        b.count = (uint32_t)(m - n + 1);
        for (size_t r = n; r <= m; r++) {
            const Instr &o = ops[r];
            const int k = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
            auto add = [&](std::vector<uint32_t> &v, uint32_t base, uint32_t len) {
                for (uint32_t j = 0; j < len; j++) v.push_back(base + j);
            };
            if (k == K_SOURCE || k == K_CLEAR) {
                Access a[3];
                row_accesses(o, a);
                add(b.writes, o.w[a[0].word], a[0].len);
            } else if (k == K_CONTRIB) {
                int c0, c1, cd;
                contribution_shape(op, &c0, &c1, &cd);
                add(b.reads, o.w[2], c0);
                add(b.reads, o.w[3], c1);
                add(b.reads, o.w[4], cd);  // (read when the run adds to the old value)
                add(b.writes, o.w[4], cd);
            } else if (k == K_PROP || k == K_IDENTITY) {
                const uint32_t len = k == K_PROP ? (op == 9 ? 4u : 2u) : o.w[2];
                add(b.reads, o.w[1], len);
                add(b.writes, o.w[1], len);
            } else if (k == K_CLOSURE) {
                add(b.reads, o.w[1], o.w[6]);
                add(b.reads, o.w[2], o.w[6]);
                b.amplitude = (int)o.w[3];
            } else {
                return g;
            }
        }
        for (auto *v : {&b.reads, &b.writes}) {
            std::sort(v->begin(), v->end());
            v->erase(std::unique(v->begin(), v->end()), v->end());
            if (!v->empty() && v->back() >= planes) return g;
        }
        blocks.push_back(std::move(b));
        n = m + 1;
    }
    const uint32_t B = (uint32_t)blocks.size();
    g.succ.assign(B, {});
    g.pred.assign(B, {});
    std::vector<int> writer(planes, -1);
    std::vector<std::vector<uint32_t>> readers(planes);
    std::unordered_map<int, uint32_t> last_amplitude;
    auto edge = [&](uint32_t from, uint32_t to) {
        if (from == to) return;
        g.succ[from].push_back(to);
        g.pred[to].push_back(from);
    };
    for (uint32_t b = 0; b < B; b++) {
        const BlockGraph::Block &x = blocks[b];
        for (uint32_t p : x.reads)
            if (writer[p] >= 0) edge((uint32_t)writer[p], b);
        for (uint32_t p : x.writes) {
            if (writer[p] >= 0) edge((uint32_t)writer[p], b);
            for (uint32_t r : readers[p]) edge(r, b);
        }
        if (x.amplitude >= 0) {
            auto it = last_amplitude.find(x.amplitude);
            if (it != last_amplitude.end()) edge(it->second, b);
            last_amplitude[x.amplitude] = b;
        }
        for (uint32_t p : x.reads) readers[p].push_back(b);
        for (uint32_t p : x.writes) {
            writer[p] = (int)b;
            readers[p].clear();
        }
    }
    while (g.lead < B && blocks[g.lead].first < sources) {
        if (!g.pred[g.lead].empty()) return g;
        g.lead++;
    }
    g.ok = true;
    return g;
}

// Row scheduling: a list scheduler builds one dependence-honouring order: among the ready blocks
// it takes the one with the highest score, ties to the earliest. A plane counts as recent while
// fewer than `window` plane touches have passed since its last one (a stand-in for the on-chip
// tiers); the score adds `w_recent` per recent plane a block reads and `w_last` per plane it
// reads for the last time. Returns an empty program when a row is not understood.
std::vector<Instr> schedule_rows(const std::vector<Instr> &ops, uint32_t sources, uint32_t planes, uint32_t window,
                                 int w_recent, int w_last) {
    const BlockGraph g = block_graph(ops, sources, planes);
    if (!g.ok) return {};
    const auto &blocks = g.blocks;
    const uint32_t B = (uint32_t)blocks.size();
    std::vector<uint32_t> indeg(B);
    for (uint32_t b = 0; b < B; b++) indeg[b] = (uint32_t)g.pred[b].size();
    const auto &succ = g.succ;
    std::vector<uint32_t> pending_reads(planes, 0);  // unscheduled blocks reading each plane
    for (const auto &x : blocks)
        for (uint32_t p : x.reads) pending_reads[p]++;
    std::vector<uint64_t> touched(planes, 0);  // clock of the last touch + 1 (0: never)
    uint64_t clock = 0;
    std::vector<uint32_t> ready;
    std::vector<Instr> out;
    out.reserve(ops.size());
    auto emit = [&](uint32_t b) {
        const auto &x = blocks[b];
        for (uint32_t r = 0; r < x.count; r++) out.push_back(ops[x.first + r]);
        for (uint32_t p : x.reads) pending_reads[p]--;
        for (uint32_t p : x.reads) touched[p] = ++clock;
        for (uint32_t p : x.writes) touched[p] = ++clock;
        for (uint32_t s : succ[b])
            if (--indeg[s] == 0) ready.push_back(s);
    };
    const uint32_t lead = g.lead;
    for (uint32_t b = lead; b < B; b++)
        if (indeg[b] == 0) ready.push_back(b);
    for (uint32_t b = 0; b < lead; b++) emit(b);
    ready.erase(std::remove_if(ready.begin(), ready.end(), [&](uint32_t r) { return r < lead; }), ready.end());
    while (!ready.empty()) {
        size_t best = 0;
        int64_t best_score = INT64_MIN;
        for (size_t k = 0; k < ready.size(); k++) {
            const auto &x = blocks[ready[k]];
            int64_t s = 0;
            for (uint32_t p : x.reads) {
                if (touched[p] && clock - (touched[p] - 1) < window) s += w_recent;
                if (pending_reads[p] == 1) s += w_last;
            }
            s = s * (int64_t)B * 2 - ready[k];  // ties: the earliest block
            if (s > best_score) {
                best_score = s;
                best = k;
            }
        }
        const uint32_t b = ready[best];
        ready[best] = ready.back();
        ready.pop_back();
        emit(b);
    }
    if (out.size() != ops.size()) return {};
    return out;
}

// Demand-driven scheduling: the closures (the program's sinks) are taken in the order `sink_key`
// gives, and each is preceded by its not yet scheduled predecessor blocks, depth first, in the
// order `child_key` gives. Values then live from their producer to consumers that follow soon
// after, so the peak of live currents -- the slot's size -- and the reuse distances drop.
// Dependence-honouring, so every value keeps its bits. sink_key 0: the closures in program
// order, 1: reversed, 2: amplitude planes in Gray-code order, 3: in bit-reversed order;
// child_key 0: the largest subtree first (Sethi-Ullman), 1: the smallest first, 2: the
// earliest block first, 3: the latest first.
std::vector<Instr> schedule_demand(const std::vector<Instr> &ops, uint32_t sources, uint32_t planes, int sink_key,
                                   int child_key = 0) {
    const BlockGraph g = block_graph(ops, sources, planes);
    if (!g.ok) return {};
    const auto &blocks = g.blocks;
    const uint32_t B = (uint32_t)blocks.size();
    std::vector<uint8_t> done(B, 0);
    std::vector<Instr> out;
    out.reserve(ops.size());
    // unscheduled ancestors of b (the blocks an emission of b would pull in)
    std::vector<uint32_t> mark(B, 0);
    uint32_t epoch = 0;
    auto ancestors = [&](uint32_t b) {
        epoch++;
        uint32_t n = 0;
        std::vector<uint32_t> stack{b};
        mark[b] = epoch;
        while (!stack.empty()) {
            const uint32_t x = stack.back();
            stack.pop_back();
            for (uint32_t p : g.pred[x])
                if (!done[p] && mark[p] != epoch) {
                    mark[p] = epoch;
                    n++;
                    stack.push_back(p);
                }
        }
        return n;
    };
    // iterative depth-first emission of b after its unscheduled predecessors
    auto visit = [&](uint32_t root) {
        if (done[root]) return;
        std::vector<std::pair<uint32_t, std::vector<uint32_t>>> stack;
        auto children = [&](uint32_t b) {
            std::vector<std::pair<int64_t, uint32_t>> c;
            for (uint32_t p : g.pred[b])
                if (!done[p]) {
                    const int64_t a = child_key < 2 ? (int64_t)ancestors(p) : 0;
                    c.push_back({child_key == 0 ? a : child_key == 1 ? -a : child_key == 2 ? -(int64_t)p : (int64_t)p, p});
                }
            std::sort(c.begin(), c.end(), [](const auto &x, const auto &y) {
                return x.first != y.first ? x.first < y.first : x.second > y.second;
            });
            std::vector<uint32_t> v;  // popped from the back: the largest subtree first
            for (const auto &x : c) v.push_back(x.second);
            return v;
        };
        stack.push_back({root, children(root)});
        while (!stack.empty()) {
            auto &top = stack.back();
            while (!top.second.empty() && done[top.second.back()]) top.second.pop_back();
            if (top.second.empty()) {
                const uint32_t b = top.first;
                stack.pop_back();
                if (done[b]) continue;
                done[b] = 1;
                const auto &x = blocks[b];
                for (uint32_t r = 0; r < x.count; r++) out.push_back(ops[x.first + r]);
                continue;
            }
            const uint32_t next = top.second.back();
            top.second.pop_back();
            stack.push_back({next, children(next)});
        }
    };
    for (uint32_t b = 0; b < g.lead; b++) visit(b);
    std::vector<uint32_t> sinks;
    for (uint32_t b = g.lead; b < B; b++)
        if (g.succ[b].empty()) sinks.push_back(b);
    auto gray_rank = [](uint32_t a) {  // the position of a in the Gray-code sequence
        uint32_t r = 0;
        for (; a; a >>= 1) r ^= a;
        return r;
    };
    auto bit_reverse = [](uint32_t a) {
        uint32_t r = 0;
        for (int k = 0; k < 16; k++) r |= ((a >> k) & 1u) << (15 - k);
        return r;
    };
    auto key = [&](uint32_t b) -> uint64_t {
        const int amp = blocks[b].amplitude;
        const uint64_t a = amp < 0 ? 0xffffu : (uint32_t)amp;
        switch (sink_key) {
        case 1: return ((uint64_t)(B - b));
        case 2: return ((uint64_t)gray_rank((uint32_t)a) << 32) | b;
        case 3: return ((uint64_t)bit_reverse((uint32_t)a) << 32) | b;
        default: return b;
        }
    };
    std::stable_sort(sinks.begin(), sinks.end(), [&](uint32_t x, uint32_t y) { return key(x) < key(y); });
    for (uint32_t b : sinks) visit(b);
    for (uint32_t b = g.lead; b < B; b++) visit(b);
    if (out.size() != ops.size()) return {};
    return out;
}

// Structural zeros. A real value of the program is known to be zero for every input when it is
// a product with a known zero or a sum of known zeros: a massless vector source's time component
// (ext_gluon writes 0 there) and the imaginary part of its last component, a constant 0, and
// what the row formulas make of those. The pass runs the row formulas themselves (the same
// templates the kernels run) over this two-point lattice, in program order, and records per
// operand which components are zero in both parts; the fused kernels then neither load nor store
// such a component in the slot (a load gives +0). Replacing a zero of either sign by +0 leaves
// every nonzero value of the program unchanged (x + 0 = x, x * 0 = 0 and no current is a
// divisor), so the outputs keep their bits whenever the reference's are finite.
struct ZeroVal {
    bool zero;  // known to be zero (of either sign) for every input
    ZeroVal() = default;
    constexpr ZeroVal(double v) : zero(v == 0.0) {}
    static ZeroVal known(bool z) {
        ZeroVal r;
        r.zero = z;
        return r;
    }
};
inline ZeroVal operator+(ZeroVal a, ZeroVal b) { return ZeroVal::known(a.zero && b.zero); }
inline ZeroVal operator-(ZeroVal a, ZeroVal b) { return ZeroVal::known(a.zero && b.zero); }
inline ZeroVal operator-(ZeroVal a) { return a; }
inline ZeroVal operator*(ZeroVal a, ZeroVal b) { return ZeroVal::known(a.zero || b.zero); }

// mask fields (the fused programs only): contribution: w[0] bits 22-27 parent 0, bits 28-31
// parent 1, w[1] bits 16-21 the destination's old value, w[7] bits 16-21 the stored value
// (the scale slot and factor keep bits 0-15); propagator and identity: w[4] = load mask | store
// mask << 8 (propagator) or << 16 (identity); closure: w[7] = parent 0 | parent 1 << 8; source:
// w[6] = store mask. Bit k: component k is zero.
void zero_masks(std::vector<Instr> &ops, const pacuda_plan_desc *d) {
    typedef Cx<ZeroVal> ZC;
    const ZeroVal N = ZeroVal::known(false);
    const ZC unknown = cx(N, N);
    std::vector<ZC> st(d->current_planes, unknown);
    auto zc = [](double re, double im) { return cx(ZeroVal(re), ZeroVal(im)); };
    auto factor = [&](uint32_t k) { return zc(d->factor_re[k], d->factor_im[k]); };
    auto mask = [](const ZC *v, uint32_t n) {
        uint32_t m = 0;
        for (uint32_t k = 0; k < n; k++)
            if (v[k].re.zero && v[k].im.zero) m |= 1u << k;
        return m;
    };
    ZC acc[6];
    for (Instr &o : ops) {
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff, flags = o.w[0] >> 16;
        if (kind == K_SOURCE) {
            const uint32_t family = o.w[5] & 7, c = family == 0 ? 1 : family == 1 ? 2 : 4;
            ZC v[4] = {unknown, unknown, unknown, unknown};
            if (family == 0) v[0] = zc(1.0, 0.0);
            if (family >= 3 && d->scale_re[o.w[4]] == 0.0) {  // ext_gluon
                v[0] = zc(0.0, 0.0);
                v[3] = cx(N, ZeroVal(0.0));
            }
            const ZC f = factor(o.w[3]);
            for (uint32_t k = 0; k < c; k++) st[o.w[1] + k] = identity_value(f.re, f.im, v[k]);
            o.w[6] = mask(&st[o.w[1]], c);
        } else if (kind == K_CLEAR) {
            for (uint32_t k = 0; k < o.w[2]; k++) st[o.w[1] + k] = zc(0.0, 0.0);
        } else if (kind == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(op, &c0, &c1, &cd);
            ZC l[6], r[4], v[6];
            ZeroVal q0[4] = {N, N, N, N}, q1[4] = {N, N, N, N};
            for (int k = 0; k < c0; k++) l[k] = st[o.w[2] + k];
            for (int k = 0; k < c1; k++) r[k] = st[o.w[3] + k];
            switch (op) {
            case 2: contribution_values<2>(l, r, q0, q1, v); break;
            case 3: contribution_values<3>(l, r, q0, q1, v); break;
            case 4: contribution_values<4>(l, r, q0, q1, v); break;
            case 5: contribution_values<5>(l, r, q0, q1, v); break;
            default: contribution_values<6>(l, r, q0, q1, v); break;
            }
            const ZC scale = mul(zc(d->scale_re[o.w[1]], d->scale_im[o.w[1]]), factor(o.w[7]));
            uint32_t old = 0, stored = 0;
            if ((flags & F_FIRST) && !(flags & F_INIT)) old = mask(&st[o.w[4]], cd);
            for (int k = 0; k < cd; k++) {
                const ZC c = mul(v[k], scale);
                if (!(flags & F_FIRST)) acc[k] = add(acc[k], c);
                else if (flags & F_INIT) acc[k] = c;
                else acc[k] = add(st[o.w[4] + k], c);
            }
            if (flags & F_LAST) {
                for (int k = 0; k < cd; k++) st[o.w[4] + k] = acc[k];
                stored = mask(acc, cd);
            }
            o.w[0] |= mask(l, c0) << 22 | mask(r, c1) << 28;
            o.w[1] |= old << 16;
            o.w[7] |= stored << 16;
        } else if (kind == K_PROP) {
            const uint32_t m = op == 9 ? 4 : 2;
            ZC c[4], v[4];
            ZeroVal p[4] = {N, N, N, N};
            for (uint32_t k = 0; k < m; k++) c[k] = st[o.w[1] + k];
            if (op == 9) propagator_values<9>(c, p, N, factor(o.w[3]), v);
            else if (op == 7) propagator_values<7>(c, p, N, factor(o.w[3]), v);
            else propagator_values<8>(c, p, N, factor(o.w[3]), v);
            o.w[4] = mask(c, m) | mask(v, m) << 8;
            for (uint32_t k = 0; k < m; k++) st[o.w[1] + k] = v[k];
        } else if (kind == K_IDENTITY) {
            const uint32_t n = o.w[2];
            const ZC f = factor(o.w[3]);
            uint32_t in = 0, out = 0;
            for (uint32_t k = 0; k < n; k++) {
                ZC &u = st[o.w[1] + k];
                if (u.re.zero && u.im.zero) in |= (k < 16 ? 1u : 0u) << (k & 15);
                u = identity_value(f.re, f.im, u);
                if (u.re.zero && u.im.zero) out |= (k < 16 ? 1u : 0u) << (k & 15);
            }
            o.w[4] = n <= 16 ? in | out << 16 : 0;
        } else if (kind == K_CLOSURE) {
            const uint32_t n = o.w[6];
            o.w[7] = n <= 8 ? mask(&st[o.w[1]], n) | mask(&st[o.w[2]], n) << 8 : 0;
        }
    }
}

// Forwarding: an in-place row (a propagator, or an identity of at most 6 components) whose
// planes were last touched by the final row of the contribution run that produces them (or by
// an in-place row on the same planes) moves right after that row; the producer keeps its result
// in the accumulators (F_HOLD) and the in-place row takes it from there (F_FROMACC), so the
// store and the load of those values disappear. Moving is safe: no row in between touches the
// planes. Bit-identical: the same operations on the same values.
uint32_t forward_in_place(std::vector<Instr> &ops) {
    uint32_t moved = 0;
    for (size_t x = 0; x < ops.size(); x++) {
        const Instr X = ops[x];
        const int kind = X.w[0] & 0xff, op = (X.w[0] >> 8) & 0xff;
        if (kind != K_PROP && kind != K_IDENTITY) continue;
        const uint32_t base = X.w[1], m = kind == K_PROP ? (op == 9 ? 4u : 2u) : X.w[2];
        if (m == 0 || m > 6) continue;
        size_t y = x;
        bool found = false, unknown = false;
        while (y > 0 && !found) {
            y--;
            Access a[3];
            const int na = row_accesses(ops[y], a);
            if (na < 0) { unknown = true; break; }
            for (int k = 0; k < na; k++) {
                const uint32_t b = ops[y].w[a[k].word], l = a[k].len;
                if (b < base + m && base < b + l) found = true;
            }
        }
        if (!found || unknown) continue;
        Instr &Y = ops[y];
        const int yk = Y.w[0] & 0xff, yop = (Y.w[0] >> 8) & 0xff, yflags = Y.w[0] >> 16;
        bool ok = false;
        if (yk == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(yop, &c0, &c1, &cd);
            ok = (yflags & F_LAST) && Y.w[4] == base && (uint32_t)cd == m;
        } else if (yk == K_PROP || yk == K_IDENTITY) {
            const uint32_t ym = yk == K_PROP ? (yop == 9 ? 4u : 2u) : Y.w[2];
            ok = Y.w[1] == base && ym == m;
        }
        if (!ok) continue;
        Y.w[0] |= (uint32_t)F_HOLD << 16;
        Instr Z = X;
        Z.w[0] |= (uint32_t)F_FROMACC << 16;
        ops.erase(ops.begin() + x);
        ops.insert(ops.begin() + y + 1, Z);
        moved++;
    }
    return moved;
}

// LDS residency: often-touched currents move from the slot to the wavefront's LDS region.
// Objects are rebuilt from the final program (a creating write starts one at its base; every
// other access must name a whole live object), ranked by accesses per plane per square root of
// their lifetime in rows, and placed first-fit in [0, capacity) planes over their lifetimes; a replay checks the
// placement (a new object only takes slots no live object holds, every access finds its own) and
// the program is rewritten to planes LDS_PLANE + base + slot. Returns the planes used (0:
// unchanged). With `only` set, only objects of `only` planes are placed, at slots that are
// multiples of it (the register tier: whole four-plane currents); objects already on chip
// (planes >= LDS_PLANE) are left where they are.
uint32_t assign_lds(std::vector<Instr> &ops, uint32_t capacity, uint32_t base = 0, uint32_t only = 0) {
    struct Obj { uint32_t base, len, first, last, touches, slot; bool lds; };
    std::vector<Obj> objs;
    std::vector<int> live;  // plane -> object living there
    std::vector<std::array<int, 3>> obj_of(ops.size());
    for (size_t n = 0; n < ops.size(); n++) {
        Access acc[3];
        const int na = row_accesses(ops[n], acc);
        if (na < 0) return 0;
        obj_of[n] = {-1, -1, -1};
        for (int k = 0; k < na; k++) {
            const uint32_t at = ops[n].w[acc[k].word], len = acc[k].len;
            if (len == 0) continue;
            if (at >= LDS_PLANE && at - LDS_PLANE + len <= LDS_PLANE) continue;  // on chip already
            if (at >= LDS_PLANE || at + len > LDS_PLANE) return 0;
            if (live.size() < at + len) live.resize(at + len, -1);
            int o = live[at];
            if (acc[k].write) {
                objs.push_back({at, len, (uint32_t)n, (uint32_t)n, 0, 0, false});
                o = (int)objs.size() - 1;
                for (uint32_t j = 0; j < len; j++) live[at + j] = o;
            } else if (o < 0 || objs[o].base != at || objs[o].len != len) {
                return 0;
            }
            objs[o].last = (uint32_t)n;
            objs[o].touches += len;
            obj_of[n][k] = o;
        }
    }
    std::vector<uint32_t> order(objs.size());
    for (uint32_t k = 0; k < order.size(); k++) order[k] = k;
    // accesses per plane, per square root of the lifetime in rows: the plain ratio (per row of
    // lifetime) crowds the tiers with short-lived objects and leaves the often-read long-lived
    // ones in the slot; measured on the gluon plans the square root leaves 5-7% fewer slot reads
    auto density = [&](uint32_t k) {
        const Obj &o = objs[k];
        return (double)o.touches / ((double)o.len * sqrt((double)(o.last - o.first + 1)));
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return density(x) > density(y); });
    std::vector<uint8_t> busy((size_t)ops.size() * capacity, 0);  // [row][slot]
    uint32_t used = 0;
    for (uint32_t k : order) {
        Obj &o = objs[k];
        if (o.len > capacity || (only && o.len != only)) continue;
        for (uint32_t s0 = 0; s0 + o.len <= capacity && !o.lds; s0 += only ? only : 1) {
            bool fits = true;
            for (uint32_t r = o.first; r <= o.last && fits; r++)
                for (uint32_t j = 0; j < o.len; j++)
                    if (busy[(size_t)r * capacity + s0 + j]) { fits = false; break; }
            if (!fits) continue;
            for (uint32_t r = o.first; r <= o.last; r++)
                for (uint32_t j = 0; j < o.len; j++) busy[(size_t)r * capacity + s0 + j] = 1;
            o.slot = s0;
            o.lds = true;
            used = std::max(used, s0 + o.len);
        }
    }
    if (!used) return 0;
    // replay the LDS placement
    std::vector<int> held(capacity, -1);
    for (size_t n = 0; n < ops.size(); n++) {
        Access acc[3];
        const int na = row_accesses(ops[n], acc);
        for (int pass = 0; pass < 2; pass++)
            for (int k = 0; k < na; k++) {
                const int oi = obj_of[n][k];
                if (oi < 0 || !objs[oi].lds) continue;
                const Obj &o = objs[oi];
                const bool birth = acc[k].write;
                if (birth != (pass == 1)) continue;
                for (uint32_t j = 0; j < o.len; j++) {
                    int &h = held[o.slot + j];
                    if (birth) {
                        if (h >= 0 && h != oi && objs[h].last >= n) return 0;
                        h = oi;
                    } else if (h != oi) {
                        return 0;
                    }
                }
            }
    }
    for (size_t n = 0; n < ops.size(); n++) {
        Access acc[3];
        const int na = row_accesses(ops[n], acc);
        for (int k = 0; k < na; k++) {
            const int oi = obj_of[n][k];
            if (oi >= 0 && objs[oi].lds) ops[n].w[acc[k].word] = LDS_PLANE + base + objs[oi].slot;
        }
    }
    return used;
}

// On-chip allocation with fills and write-through. The program is straight-line code, so the
// on-chip slots are allocated as registers are in a basic block: the current objects (as in
// assign_lds) move on chip for parts of their lives. An access to an object in the slot takes
// on-chip room when there is some — free slots, or slots whose holders are next read later than
// this object is (the furthest-next-read rule) — and the object stays there until it is evicted
// or its value dies. A read that brings it on chip loads from the slot and keeps the copy
// (DUAL_FILL); writes while on chip go on chip only, and when an object is evicted before a later
// read, its last write also stores to the slot (DUAL_THROUGH, set afterwards: nothing reads the
// slot copy in between). An access that finds no room goes to the slot. LDS slots are
// [0, lds_cap) (any object, contiguous), the register tier [lds_cap, lds_cap + reg) (four-plane
// objects at multiples of four). With `recompute`, a read that would fill an on-chip copy of a
// current a leading vector-source row makes (and no row writes again) is preceded instead by a
// copy of that source row, which makes the value again into the same room (the same operations
// on the same f64 inputs: the same bits), so the slot load disappears; reads that find no room
// still load the slot's copy, and evicting such a current never needs a store. A
// replay of the rewritten program checks that every read finds the value last written (in the
// slot only its unmasked components). With `keep_copies`, a current made again that is created
// on chip also stores to the slot (else a read that finds no room after the current has left
// finds no slot copy, and the replay fails). Returns false with the program unchanged when an
// access does not name a whole object or the replay fails.
bool allocate_on_chip(std::vector<Instr> &ops, uint32_t lds_cap, uint32_t reg, uint32_t *lds_used,
                      uint32_t *reg_used, bool recompute = false, bool keep_copies = false) {
    struct Acc {
        uint32_t row;
        int word;
        uint32_t len;
        bool rd, wr, create;
        uint32_t rmask, wmask;
        int obj;
        uint32_t field;
    };
    std::vector<Acc> acc;
    for (size_t n = 0; n < ops.size(); n++) {
        const Instr &o = ops[n];
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff, flags = o.w[0] >> 16;
        const uint32_t r = (uint32_t)n;
        if (kind == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(op, &c0, &c1, &cd);
            acc.push_back({r, 2, (uint32_t)c0, true, false, false, (o.w[0] >> 22) & 0x3f, 0, -1, 0});
            acc.push_back({r, 3, (uint32_t)c1, true, false, false, o.w[0] >> 28, 0, -1, 0});
            const bool rd = (flags & F_FIRST) && !(flags & F_INIT), wr = (flags & F_LAST) && !(flags & F_HOLD),
                       cr = (flags & F_FIRST) && (flags & F_INIT);
            if (rd || wr || cr)
                acc.push_back({r, 4, (uint32_t)cd, rd, wr, cr, (o.w[1] >> 16) & 0x3f, (o.w[7] >> 16) & 0x3f, -1, 0});
        } else if (kind == K_PROP) {
            acc.push_back({r, 1, op == 9 ? 4u : 2u, !(flags & F_FROMACC), !(flags & F_HOLD), false, o.w[4] & 0xff,
                           (o.w[4] >> 8) & 0xff, -1, 0});
        } else if (kind == K_IDENTITY) {
            acc.push_back({r, 1, o.w[2], !(flags & F_FROMACC), !(flags & F_HOLD), false, o.w[4] & 0xffff, o.w[4] >> 16,
                           -1, 0});
        } else if (kind == K_CLOSURE) {
            acc.push_back({r, 1, o.w[6], true, false, false, o.w[7] & 0xff, 0, -1, 0});
            acc.push_back({r, 2, o.w[6], true, false, false, (o.w[7] >> 8) & 0xff, 0, -1, 0});
        } else if (kind == K_CLEAR) {
            acc.push_back({r, 1, o.w[2], false, true, true, 0, 0, -1, 0});
        } else if (kind == K_SOURCE) {
            Access a[3];
            row_accesses(o, a);
            acc.push_back({r, 1, a[0].len, false, true, true, 0, o.w[6] & 15, -1, 0});
        } else {
            return false;
        }
    }
    // objects: a creating access starts one; every other access names a whole live object
    struct Obj { uint32_t base, len; };
    std::vector<Obj> objs;
    std::vector<int> live;
    uint32_t slot_top = 0;
    for (Acc &x : acc) {
        if (x.len == 0) continue;
        const uint32_t at = ops[x.row].w[x.word];
        if (at >= LDS_PLANE || at + x.len > LDS_PLANE || x.len > 63) return false;
        if (live.size() < at + x.len) live.resize(at + x.len, -1);
        slot_top = std::max(slot_top, at + x.len);
        if (x.create) {
            objs.push_back({at, x.len});
            for (uint32_t j = 0; j < x.len; j++) live[at + j] = (int)objs.size() - 1;
        }
        const int o = live[at];
        if (o < 0 || objs[o].base != at || objs[o].len != x.len) return false;
        x.obj = o;
        x.field = at;
    }
    // the currents made again instead of stored: made by a source row, never written again
    std::vector<char> again(objs.size(), 0);
    std::vector<uint32_t> maker(objs.size(), 0);  // the access that makes each object
    if (recompute) {
        for (size_t i = 0; i < acc.size(); i++) {
            const Acc &x = acc[i];
            if (x.obj < 0) continue;
            if (x.create) {  // (vector sources: a spinor's two components are cheaper to load)
                again[x.obj] = (ops[x.row].w[0] & 0xff) == K_SOURCE && (ops[x.row].w[5] & 7) >= 3;
                maker[x.obj] = (uint32_t)i;
            } else if (x.wr) {
                again[x.obj] = 0;
            }
        }
    }
    // the next read of the value each access leaves (none when the object is written over or dies first)
    const uint32_t NONE = UINT32_MAX;
    std::vector<uint32_t> next_read(acc.size(), NONE), ahead(objs.size(), NONE);
    for (size_t i = acc.size(); i-- > 0;) {
        const Acc &x = acc[i];
        if (x.obj < 0 || !(x.rd || x.wr)) continue;
        next_read[i] = ahead[x.obj];
        if (x.rd) ahead[x.obj] = (uint32_t)i;
        else ahead[x.obj] = NONE;  // a write over the value
    }
    const uint32_t cap = lds_cap + reg;
    std::vector<int> occ(cap, -1), pos(objs.size(), -1);
    std::vector<uint32_t> next_of(objs.size(), NONE), last_write(objs.size(), 0);
    std::vector<char> dirty(objs.size(), 0);
    uint32_t used_lds = 0, used_reg = 0;
    auto leave = [&](int o) {
        for (uint32_t j = 0; j < objs[o].len; j++) occ[pos[o] + j] = -1;
        pos[o] = -1;
    };
    // (the objects a row names are not displaced by its own accesses: a row may read its operands
    // one component at a time, interleaved)
    // and the room of a value that dies in a row is free from the next row on
    std::vector<uint32_t> named_in(objs.size(), UINT32_MAX);
    std::vector<int> dying;
    struct Remake { uint32_t before; int obj; uint32_t slot; };  // a source row made again before access `before`
    std::vector<Remake> remakes;
    for (size_t i = 0; i < acc.size(); i++) {
        Acc &x = acc[i];
        if (i == 0 || acc[i - 1].row != x.row) {
            for (int o : dying) leave(o);
            dying.clear();
            for (size_t j = i; j < acc.size() && acc[j].row == x.row; j++)
                if (acc[j].obj >= 0) named_in[acc[j].obj] = x.row;
        }
        if (x.obj < 0 || !(x.rd || x.wr)) continue;
        const int o = x.obj;
        const uint32_t t = next_read[i], n = x.len, p = objs[o].base;
        dying.erase(std::remove(dying.begin(), dying.end(), o), dying.end());
        const bool remake = again[o] && x.rd && pos[o] < 0;  // (made again into on-chip room)
        if (pos[o] >= 0) {
            x.field = LDS_PLANE + (uint32_t)pos[o];
            if (x.wr && !again[o]) {
                dirty[o] = 1;
                last_write[o] = (uint32_t)i;
            }
        } else if (t != NONE || remake) {
            // the room with the latest next read among the holders it displaces (free room first)
            int best = -1;
            uint64_t best_key = 0;
            auto consider = [&](uint32_t s) {
                uint32_t soonest = NONE;
                int dirt = 0;
                for (uint32_t j = 0; j < n; j++) {
                    const int v = occ[s + j];
                    if (v < 0) continue;
                    if (next_of[v] <= t || named_in[v] == x.row) return;
                    soonest = std::min(soonest, next_of[v]);
                    dirt += dirty[v];
                }
                const uint64_t key = ((uint64_t)soonest << 8) | (uint64_t)(255 - std::min(dirt, 255));
                if (best < 0 || key > best_key) {
                    best = (int)s;
                    best_key = key;
                }
            };
            if (reg && n == 4 && (!x.rd || remake))  // (a read's copy goes to LDS)
                for (uint32_t s = lds_cap; s + 4 <= cap; s += 4) consider(s);
            for (uint32_t s = 0; s + n <= lds_cap; s++) consider(s);

            if (best >= 0) {
                for (uint32_t j = 0; j < n; j++) {
                    const int v = occ[best + j];
                    if (v < 0) continue;
                    if (dirty[v]) {  // its last write also stores to the slot
                        Acc &w = acc[last_write[v]];
                        w.field = (w.field & DUAL) ? (w.field | DUAL_THROUGH)
                                                   : dual_field(DUAL_THROUGH, objs[v].base, w.field - LDS_PLANE);
                        dirty[v] = 0;
                    }
                    leave(v);
                }
                pos[o] = best;
                for (uint32_t j = 0; j < n; j++) occ[best + j] = o;
                if (remake) {  // (else the read loads the slot's copy)
                    remakes.push_back({(uint32_t)i, o, (uint32_t)best});
                    x.field = LDS_PLANE + (uint32_t)best;
                } else if (x.create && again[o] && keep_copies) {
                    // (a current made again also keeps its slot copy, for the reads that find no
                    // room after it has left)
                    x.field = dual_field(DUAL_THROUGH, p, (uint32_t)best);
                } else {
                    x.field = x.rd ? dual_field(DUAL_FILL, p, (uint32_t)best) : LDS_PLANE + (uint32_t)best;
                }
                dirty[o] = x.wr && !again[o];
                if (x.wr) last_write[o] = (uint32_t)i;
                if ((uint32_t)best >= lds_cap) used_reg = std::max(used_reg, (uint32_t)best + n - lds_cap);
                else used_lds = std::max(used_lds, (uint32_t)best + n);
            }
        }
        next_of[o] = t;
        if (t == NONE && pos[o] >= 0) {  // the value dies here
            dirty[o] = 0;
            dying.push_back(o);
        }
    }
    if (!used_lds && !used_reg) return false;
    // replay: each location holds (object, version) of its last write
    std::vector<std::pair<int, uint32_t>> slot(slot_top, {-1, 0}), chip(cap, {-1, 0});
    std::vector<uint32_t> version(objs.size(), 0);
    // (within a row, no on-chip copy a read keeps may overlap another access's on-chip planes: the
    // kernels may interleave a row's operand reads component by component)
    std::vector<std::array<uint32_t, 3>> row_chip;  // (first, end, is a copy) of the row's accesses
    size_t next_remake = 0;
    for (size_t i = 0; i < acc.size(); i++) {
        const Acc &x = acc[i];
        if (i == 0 || acc[i - 1].row != x.row) row_chip.clear();
        for (; next_remake < remakes.size() && remakes[next_remake].before == i; next_remake++) {
            const Remake &m = remakes[next_remake];  // the value itself again (no new version)
            for (uint32_t j = 0; j < objs[m.obj].len; j++) chip[m.slot + j] = {m.obj, version[m.obj]};
        }
        if (x.obj < 0 || !(x.rd || x.wr)) continue;
        if (x.rd && x.field >= LDS_PLANE) {
            const bool copy = (x.field & DUAL_FILL) != 0;
            const uint32_t s0 = (x.field & DUAL) ? dual_chip(x.field) : x.field - LDS_PLANE;
            for (const auto &c : row_chip)
                if ((copy || c[2]) && s0 < c[1] && c[0] < s0 + x.len) return false;
            row_chip.push_back({s0, s0 + x.len, copy ? 1u : 0u});
        }
        const std::pair<int, uint32_t> now{x.obj, version[x.obj]};
        const uint32_t f = x.field;
        const bool dual = f >= LDS_PLANE && (f & DUAL), onchip = f >= LDS_PLANE && !(f & DUAL);
        const uint32_t q = dual ? (f & (LDS_PLANE - 1)) : f, s = dual ? dual_chip(f) : f - LDS_PLANE;
        if ((dual || onchip) && (s + x.len > cap || (s >= lds_cap && (x.len != 4 || (s - lds_cap) % 4))))
            return false;
        if (x.rd && dual && (f & DUAL_FILL) && s + x.len > lds_cap) return false;
        if (x.rd) {
            if (onchip || (dual && !(f & DUAL_FILL))) {
                for (uint32_t j = 0; j < x.len; j++)
                    if (chip[s + j] != now) return false;
            } else {
                for (uint32_t j = 0; j < x.len; j++)
                    if (!((x.rmask >> j) & 1) && slot[q + j] != now) return false;
                if (dual)
                    for (uint32_t j = 0; j < x.len; j++) chip[s + j] = now;
            }
        }
        if (x.wr) {
            const std::pair<int, uint32_t> next{x.obj, ++version[x.obj]};
            if (onchip || dual)
                for (uint32_t j = 0; j < x.len; j++) chip[s + j] = next;
            if (!onchip && (!dual || (f & DUAL_THROUGH)))
                for (uint32_t j = 0; j < x.len; j++)
                    if (!((x.wmask >> j) & 1)) slot[q + j] = next;
        }
    }
    for (const Acc &x : acc)
        if (x.obj >= 0) ops[x.row].w[x.word] = x.field;
    if (!remakes.empty()) {  // the source rows made again, each before the row that reads it
        std::vector<Instr> out;
        out.reserve(ops.size() + remakes.size());
        size_t k = 0;
        for (uint32_t r = 0; r < ops.size(); r++) {
            for (; k < remakes.size() && acc[remakes[k].before].row == r; k++) {
                Instr m = ops[acc[maker[remakes[k].obj]].row];
                m.w[1] = LDS_PLANE + remakes[k].slot;
                m.w[6] = 0;
                out.push_back(m);
            }
            out.push_back(ops[r]);
        }
        ops.swap(out);
    }
    *lds_used = used_lds;
    *reg_used = used_reg;
    return true;
}

// The slot's share of a fused program's work: the current components it loads from and stores
// to the slot per point (on-chip planes and known-zero components take none).
uint64_t slot_accesses(const std::vector<Instr> &ops) {
    uint64_t n = 0;
    auto count = [&](uint32_t f, uint32_t len, uint32_t zero, bool load = true) {
        if (f >= LDS_PLANE && !(f & (load ? DUAL_FILL : DUAL_THROUGH))) return;
        for (uint32_t j = 0; j < len; j++) n += !((zero >> j) & 1);
    };
    for (const Instr &o : ops) {
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff, flags = o.w[0] >> 16;
        if (kind == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(op, &c0, &c1, &cd);
            count(o.w[2], c0, (o.w[0] >> 22) & 0x3f);
            count(o.w[3], c1, (o.w[0] >> 28) & 0xf);
            if ((flags & F_FIRST) && !(flags & F_INIT)) count(o.w[4], cd, (o.w[1] >> 16) & 0x3f);
            if ((flags & F_LAST) && !(flags & F_HOLD)) count(o.w[4], cd, (o.w[7] >> 16) & 0x3f, false);
        } else if (kind == K_PROP) {
            const uint32_t m = op == 9 ? 4 : 2;
            if (!(flags & F_FROMACC)) count(o.w[1], m, o.w[4] & 0xff);
            if (!(flags & F_HOLD)) count(o.w[1], m, (o.w[4] >> 8) & 0xff, false);
        } else if (kind == K_IDENTITY) {
            if (!(flags & F_FROMACC)) count(o.w[1], o.w[2], o.w[4] & 0xffff);
            if (!(flags & F_HOLD)) count(o.w[1], o.w[2], o.w[4] >> 16, false);
        } else if (kind == K_CLOSURE) {
            count(o.w[1], o.w[6], o.w[7] & 0xff);
            count(o.w[2], o.w[6], (o.w[7] >> 8) & 0xff);
        } else if (kind == K_CLEAR) {
            count(o.w[1], o.w[2], 0, false);
        } else if (kind == K_SOURCE) {
            Access a[3];
            row_accesses(o, a);
            count(o.w[1], a[0].len, o.w[6], false);
        }
    }
    return n;
}

struct FusedState {
    std::mutex lock;
    double *scratch = nullptr;
    size_t scratch_doubles = 0;
    uint32_t *counter = nullptr;
    ItemDev *items = nullptr;
    uint32_t *unit_start = nullptr;
    size_t item_capacity = 0;
    hipEvent_t start = nullptr, stop = nullptr;
    // recorded after each launch on the stream it went to: the next call, on whatever stream,
    // waits for it before it rewrites the item table, the counter or the slots
    hipEvent_t last = nullptr;
    bool last_pending = false;
};
FusedState fs;

int compute_units() {
    static int cus = 0;
    if (!cus) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || v <= 0) v = 256;
        cus = v;
    }
    return cus;
}

}  // namespace

namespace {

// The host half of plan creation: validation, the program, and the fused path's passes.
struct Compiled {
    uint32_t sources = 0, planes = 0;
    std::vector<Instr> plain, fwd, lds[LDS_VARIANTS], reg[LDS_VARIANTS];
    uint32_t lds_used[LDS_VARIANTS] = {}, reg_used[LDS_VARIANTS] = {};
    // the same for the f64 kernel, whose programs make sources again instead of storing them
    std::vector<Instr> rlds[LDS_VARIANTS], rreg[LDS_VARIANTS];
    uint32_t rreg_used[LDS_VARIANTS] = {};
    uint32_t reg_forms[FORM_REGS];
    bool zero_fill = true;
    std::vector<uint32_t> bindings;
    bool symjit = false;
    CoopProgram coop[2][COOP_SHAPES];  // the cooperative programs per class of current (ok false: none)
};

// F_ZERO on the first closure of every amplitude plane (closures are the only rows that touch
// amplitudes); returns whether some plane in [0, amplitude_planes) has none, i.e. whether the
// kernel must still zero the amplitude planes before the program.
bool mark_first_closures(std::vector<Instr> &ops, uint32_t amplitude_planes) {
    std::vector<uint8_t> seen(amplitude_planes, 0);
    uint32_t first = 0;
    for (Instr &o : ops) {
        if ((o.w[0] & 0xff) != K_CLOSURE || o.w[3] >= amplitude_planes) continue;
        if (!seen[o.w[3]]) {
            seen[o.w[3]] = 1;
            o.w[0] |= (uint32_t)F_ZERO << 16;
            first++;
        }
    }
    return first < amplitude_planes;
}

// The register-resident forms: the FORM_REGS forms the program's rows (the leading sources aside,
// which run before the registers are filled) read most often go to register slots 0.. in that order;
// their form fields in the rows are rewritten to f | (slot + 1) << FORM_SLOT_SHIFT. The same
// choice serves every program of a plan (they differ only in where currents live).
void choose_form_regs(const std::vector<Instr> &ops, uint32_t form_count, uint32_t *reg_forms) {
    std::vector<uint64_t> uses(form_count, 0);
    for (const Instr &o : ops) {
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
        if (kind == K_CONTRIB && op == 6) {
            if (o.w[5] < form_count) uses[o.w[5]]++;
            if (o.w[6] < form_count) uses[o.w[6]]++;
        } else if (kind == K_PROP && o.w[2] < form_count) {
            uses[o.w[2]]++;
        }
    }
    std::vector<uint32_t> order(form_count);
    for (uint32_t f = 0; f < form_count; f++) order[f] = f;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return uses[a] > uses[b]; });
    for (uint32_t s = 0; s < FORM_REGS; s++)
        reg_forms[s] = s < form_count && uses[order[s]] ? order[s] : UINT32_MAX;
}

void mark_form_regs(std::vector<Instr> &ops, const uint32_t *reg_forms, uint32_t sources) {
    auto mark = [&](uint32_t &w) {
        for (uint32_t s = 0; s < FORM_REGS; s++)
            if (reg_forms[s] == w) {
                w |= (s + 1) << FORM_SLOT_SHIFT;
                return;
            }
    };
    for (size_t n = 0; n < ops.size(); n++) {
        Instr &o = ops[n];
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
        if (kind == K_CONTRIB && op == 6) {
            mark(o.w[5]);
            mark(o.w[6]);
        } else if (kind == K_PROP || (kind == K_SOURCE && n >= sources)) {  // (a source row made again)
            mark(o.w[2]);
        }
    }
}

// The cooperative programs of `prog` (over `planes` current planes): per class of current and
// shape, the candidate (phases as early or as late as possible, three LDS rankings) that leaves
// the fewest current accesses in the slot
void build_coop(Compiled &c, const std::vector<Instr> &prog, uint32_t planes, const pacuda_plan_desc *d) {
    const uint32_t count = d->offsets[d->outputs];
    std::vector<uint8_t> used(d->amplitude_planes, 0);
    for (uint32_t k = 0; k < count; k++) used[d->planes[k]] = 1;
    for (int cls = 0; cls < 2; cls++)
        for (int s = 0; s < COOP_SHAPES; s++) {
            if (!coop_built(cls, s)) continue;
            const uint32_t plane_bytes = cls ? 512 : 1024;
            for (int late = 0; late < 2; late++)
                for (int rank = 0; rank < 3; rank++) {
                    CoopProgram q = coop_program(prog, planes, d->amplitude_planes, d->form_count, used, COOP_WAVES[s],
                                                 (LDS_BYTES / COOP_UNITS[s] - COOP_CONTROL) / plane_bytes, plane_bytes,
                                                 c.bindings, late, rank);
                    CoopProgram &best = c.coop[cls][s];
                    if (q.ok && (!best.ok || q.slot_accesses < best.slot_accesses)) best = std::move(q);
                }
        }
}

int compile_plan(const pacuda_plan_desc *d, Compiled &c) {
    if (!d || d->abi_version != PACMAN_RECURRENCE_ABI || d->struct_size != sizeof(pacuda_plan_desc) || !d->words ||
        !d->offsets || (d->form_count && !d->form_terms) || (d->term_count && (!d->term_slots || !d->term_coefficients)) ||
        (d->scale_count && (!d->scale_re || !d->scale_im)) || (d->factor_count && (!d->factor_re || !d->factor_im)))
        return PACMAN_RECURRENCE_INVALID;
    // (plane and form indices share their fields with the on-chip markers of the fused passes)
    if (d->current_planes >= LDS_PLANE || d->form_count >= (1u << FORM_SLOT_SHIFT)) return PACMAN_RECURRENCE_INVALID;
    const uint32_t count = d->offsets[d->outputs];
    if (d->offsets[0] != 0 || (count && (!d->planes || !d->weights))) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t o = 0; o < d->outputs; o++)
        if (d->offsets[o] > d->offsets[o + 1]) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t k = 0; k < count; k++)
        if (d->planes[k] >= d->amplitude_planes) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t f = 0; f < d->form_count; f++)
        if (d->form_terms[2 * f] > d->term_count || d->form_terms[2 * f + 1] > d->term_count - d->form_terms[2 * f])
            return PACMAN_RECURRENCE_INVALID;
    for (uint32_t t = 0; t < d->term_count; t++)
        if (d->term_slots[t] >= d->external_count) return PACMAN_RECURRENCE_INVALID;
    CompiledProgram hp = compile_program(d->words, d->word_count, d->width, !d->no_sort,
                                         {PACUDA_KERNEL_COUNT, PACUDA_MAX_IN, PACUDA_MAX_OUT});
    if (hp.error) {
        fprintf(stderr, "pacman recurrence runner: %s\n", hp.error);
        return PACMAN_RECURRENCE_INVALID;
    }
    if (!program_fits(hp.need, d->current_planes, d->amplitude_planes, d->form_count, d->scale_count, d->factor_count)) {
        fprintf(stderr, "pacman recurrence runner: the program indexes past the plan's planes or tables\n");
        return PACMAN_RECURRENCE_INVALID;
    }
    c.bindings = hp.bindings;
    c.symjit = hp.symjit;
    if (hp.symjit) {
        // Symjit rows read and write single components their bindings name, which the passes
        // below (whole-object compaction, forwarding, on-chip placement, zero masks) do not model:
        // such plans run their program as built, every current plane in the slot.
        c.planes = d->current_planes;
        c.sources = hp.sources;
        c.plain = c.fwd = hp.ops;
        // (every variant is the program as built, which places nothing on chip: the plan joins
        // launches of any LDS variant without taking LDS)
        for (int v = 0; v < LDS_VARIANTS; v++) {
            c.lds[v] = c.reg[v] = c.rlds[v] = c.rreg[v] = hp.ops;
            c.lds_used[v] = 1;
        }
        for (uint32_t &f : c.reg_forms) f = UINT32_MAX;
        c.zero_fill = true;
        // the cooperative programs model a symjit row's accesses from its binding (whole parents,
        // the destination created or read and written whole); the first closure of each
        // amplitude starts from zero there
        std::vector<Instr> coop = hp.ops;
        mark_first_closures(coop, d->amplitude_planes);
        build_coop(c, coop, d->current_planes, d);
        return PACMAN_RECURRENCE_OK;
    }
    if (!reads_after_writes(hp.ops, d->current_planes)) {
        fprintf(stderr, "pacman recurrence runner: the program reads a current plane before writing it\n");
        return PACMAN_RECURRENCE_INVALID;
    }
    // (the mask fields share the scale slot's and factor's words)
    if (d->scale_count > 0x10000 || d->factor_count > 0x10000) {
        fprintf(stderr, "pacman recurrence runner: the fused path takes at most 65536 scales and factors\n");
        return PACMAN_RECURRENCE_INVALID;
    }
    zero_masks(hp.ops, d);
    const std::vector<Instr> built = hp.ops;
    // the slot keeps only the live currents
    c.planes = d->current_planes;
    if (const uint32_t compacted = compact_planes(hp.ops, d->current_planes)) c.planes = compacted;
    c.sources = hp.sources;
    c.plain = hp.ops;
    forward_in_place(hp.ops);
    c.zero_fill = mark_first_closures(hp.ops, d->amplitude_planes);
    c.fwd = hp.ops;
    build_coop(c, c.fwd, c.planes, d);
    // the on-chip programs: the passes over the rows as built and over each candidate order, the
    // allocator's placement (the static one where it gives up), keeping per variant and tier the
    // program that leaves the fewest slot accesses
    std::vector<std::vector<Instr>> orders{built};
    static const int weights[][2] = {{1, 0}, {1, 1}, {0, 1}, {2, 1}, {1, 2}, {4, 1}};
    for (uint32_t window : {24u, 72u})
        for (const auto &w : weights) {
            std::vector<Instr> o = schedule_rows(built, hp.sources, d->current_planes, window, w[0], w[1]);
            if (!o.empty()) orders.push_back(std::move(o));
        }
    // (the closures in bit-reversed amplitude order were the best sink order on the gluon plans,
    // ahead of program, reversed and Gray-code orders, under each predecessor order)
    for (int child = 0; child < 4; child++) {
        std::vector<Instr> o = schedule_demand(built, hp.sources, d->current_planes, 3, child);
        if (!o.empty()) orders.push_back(std::move(o));
    }
    uint64_t best[2][2][LDS_VARIANTS];  // [sources made again][tier][variant]
    bool have[2][2][LDS_VARIANTS] = {};
    uint32_t lds_of[2][2][LDS_VARIANTS] = {};
    for (const std::vector<Instr> &order : orders) {
        std::vector<Instr> ops = order;
        uint32_t planes = d->current_planes;
        if (const uint32_t compacted = compact_planes(ops, d->current_planes)) planes = compacted;
        forward_in_place(ops);
        mark_first_closures(ops, d->amplitude_planes);
        for (int again = 0; again < 2; again++)
            for (int v = 0; v < LDS_VARIANTS; v++)
                for (int tier = 0; tier < 2; tier++) {  // LDS only; LDS and the register tier
                    const uint32_t reg = tier ? REG_PLANES : 0;
                    std::vector<Instr> prog = ops;
                    uint32_t lds_used = 0, reg_used = 0;
                    // (programs that make sources again: first without slot copies of the made
                    // currents, then, where the replay needs them, with)
                    if (!allocate_on_chip(prog, LDS_CAPS[v], reg, &lds_used, &reg_used, again) &&
                        !(again && allocate_on_chip(prog, LDS_CAPS[v], reg, &lds_used, &reg_used, true, true))) {
                        if (have[again][tier][v]) continue;
                        prog = ops;
                        lds_used = assign_lds(prog, LDS_CAPS[v]);
                        if (reg) reg_used = assign_lds(prog, reg, LDS_CAPS[v], 4);
                    }
                    // (a source row made again is charged as two slot accesses: its f64 arithmetic)
                    uint64_t cost = slot_accesses(prog);
                    for (size_t r = c.sources; r < prog.size(); r++) cost += 2 * ((prog[r].w[0] & 0xff) == K_SOURCE);
                    if (have[again][tier][v] && cost >= best[again][tier][v]) continue;
                    have[again][tier][v] = true;
                    best[again][tier][v] = cost;
                    lds_of[again][tier][v] = lds_used;
                    (again ? (tier ? c.rreg : c.rlds) : (tier ? c.reg : c.lds))[v] = std::move(prog);
                    if (tier) (again ? c.rreg_used : c.reg_used)[v] = reg_used;
                    c.planes = std::max(c.planes, planes);
                }
    }
    // the f64 kernel runs the other kernels' program where making sources again does not pay
    for (int v = 0; v < LDS_VARIANTS; v++)
        for (int tier = 0; tier < 2; tier++)
            if (best[1][tier][v] >= best[0][tier][v]) {
                (tier ? c.rreg : c.rlds)[v] = (tier ? c.reg : c.lds)[v];
                if (tier) c.rreg_used[v] = c.reg_used[v];
                lds_of[1][tier][v] = lds_of[0][tier][v];
            }
    // (the LDS region of a wavefront holds what any program of a variant places there)
    for (int v = 0; v < LDS_VARIANTS; v++)
        c.lds_used[v] = std::max(std::max(lds_of[0][0][v], lds_of[0][1][v]), std::max(lds_of[1][0][v], lds_of[1][1][v]));
    choose_form_regs(c.fwd, d->form_count, c.reg_forms);
    for (int v = 0; v < LDS_VARIANTS; v++)
        for (auto *prog : {&c.lds[v], &c.reg[v], &c.rlds[v], &c.rreg[v]}) mark_form_regs(*prog, c.reg_forms, c.sources);
    return PACMAN_RECURRENCE_OK;
}

}  // namespace

extern "C" int pacuda_plan_program(const pacuda_plan_desc *d, uint32_t variant, uint32_t *out, uint32_t capacity,
                                   uint32_t *count, uint32_t *sources, uint32_t *planes, uint32_t *lds_planes) {
    if (!count || variant >= 2 + 4 * LDS_VARIANTS) return PACMAN_RECURRENCE_INVALID;
    Compiled c;
    const int st = compile_plan(d, c);
    if (st) return st;
    // 2 + v: LDS variant v, + LDS_VARIANTS: with the register tier, + 2 LDS_VARIANTS: the f64
    // kernel's programs (sources made again)
    const uint32_t set = variant < 2 ? 0 : (variant - 2) / LDS_VARIANTS, v = variant < 2 ? 0 : (variant - 2) % LDS_VARIANTS;
    const bool reg = set == 1 || set == 3, again = set >= 2;
    const std::vector<Instr> &ops = variant == 0 ? c.plain
                                    : variant == 1 ? c.fwd
                                    : again ? (reg ? c.rreg[v] : c.rlds[v])
                                            : (reg ? c.reg[v] : c.lds[v]);
    *count = (uint32_t)ops.size();
    if (sources) *sources = c.sources;
    if (planes) *planes = c.planes;
    // on-chip slots the program names: LDS slots, then (register variants) the register tier
    const uint32_t reg_used = again ? c.rreg_used[v] : c.reg_used[v];
    if (lds_planes) *lds_planes = variant < 2 ? 0 : reg && reg_used ? LDS_CAPS[v] + reg_used : c.lds_used[v];
    if (out) {
        if (capacity < ops.size()) return PACMAN_RECURRENCE_INVALID;
        memcpy(out, ops.data(), ops.size() * sizeof(Instr));
    }
    return PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_plan_coop_program(const pacuda_plan_desc *d, uint32_t shape, uint32_t serial, uint32_t *out,
                                        uint32_t capacity, uint32_t *count, uint32_t *info, uint32_t *starts) {
    if (!count || shape >= 2u * COOP_SHAPES) return PACMAN_RECURRENCE_INVALID;
    Compiled c;
    const int st = compile_plan(d, c);
    if (st) return st;
    const CoopProgram &q = c.coop[shape / COOP_SHAPES][shape % COOP_SHAPES];
    if (!q.ok) return PACMAN_RECURRENCE_INVALID;
    const std::vector<Instr> &ops = serial ? q.serial : q.ops;
    *count = (uint32_t)ops.size();
    if (info) {
        const uint32_t v[6] = {q.waves, q.phases, q.lds_planes, q.slot_planes, (uint32_t)q.slot_accesses,
                               (uint32_t)q.lds_accesses};
        memcpy(info, v, sizeof v);
    }
    if (starts) memcpy(starts, q.start.data(), q.start.size() * sizeof(uint32_t));
    if (out) {
        if (capacity < ops.size()) return PACMAN_RECURRENCE_INVALID;
        memcpy(out, ops.data(), ops.size() * sizeof(Instr));
    }
    return PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_plan_bindings(const pacuda_plan_desc *d, uint32_t *out, uint32_t capacity, uint32_t *count) {
    if (!count) return PACMAN_RECURRENCE_INVALID;
    Compiled c;
    const int st = compile_plan(d, c);
    if (st) return st;
    *count = (uint32_t)c.bindings.size();
    if (out) {
        if (capacity < c.bindings.size()) return PACMAN_RECURRENCE_INVALID;
        memcpy(out, c.bindings.data(), c.bindings.size() * sizeof(uint32_t));
    }
    return PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_plan_create(const pacuda_plan_desc *d, pacuda_plan **out) {
    if (!out) return PACMAN_RECURRENCE_INVALID;
    *out = nullptr;
    Compiled cp;
    if (const int st = compile_plan(d, cp)) return st;
    const uint32_t count = d->offsets[d->outputs];
    const uint32_t planes = cp.planes;
    pacuda_plan *p = new pacuda_plan();
    const uint32_t S = d->scale_count, X = d->factor_count;
    std::vector<double> c64(2 * (size_t)(S + X)), cdf;
    for (uint32_t k = 0; k < S; k++) { c64[k] = d->scale_re[k]; c64[S + k] = d->scale_im[k]; }
    for (uint32_t k = 0; k < X; k++) { c64[2 * S + k] = d->factor_re[k]; c64[2 * S + X + k] = d->factor_im[k]; }
    cdf = c64;
    for (double &v : cdf) {  // df64 (hi, lo) = ((float)x, (float)(x - hi)) in the double's 8 bytes
        float hl[2];
        hl[0] = (float)v;
        hl[1] = (float)(v - (double)hl[0]);
        memcpy(&v, hl, 8);
    }
    bool ok = true;
    PlanDev &h = p->host;
    memset(&h, 0, sizeof h);
    for (int v = 0; v < LDS_VARIANTS; v++) {
        h.prog_lds[v] = upload(p, cp.lds[v].data(), cp.lds[v].size(), &ok);
        h.prog_reg[v] = upload(p, cp.reg[v].data(), cp.reg[v].size(), &ok);
        h.prog_rlds[v] = upload(p, cp.rlds[v].data(), cp.rlds[v].size(), &ok);
        h.prog_rreg[v] = upload(p, cp.rreg[v].data(), cp.rreg[v].size(), &ok);
        h.count_rlds[v] = (uint32_t)cp.rlds[v].size();
        h.count_rreg[v] = (uint32_t)cp.rreg[v].size();
        h.lds_planes[v] = cp.lds_used[v];
    }
    h.count = (uint32_t)cp.fwd.size();
    h.prog_plain = upload(p, cp.plain.data(), cp.plain.size(), &ok);
    h.count_plain = (uint32_t)cp.plain.size();
    h.prog_fwd = upload(p, cp.fwd.data(), cp.fwd.size(), &ok);
    h.sources = cp.sources;
    memcpy(h.reg_forms, cp.reg_forms, sizeof h.reg_forms);
    h.zero_fill = cp.zero_fill ? 1 : 0;
    h.c64 = upload(p, c64.data(), c64.size(), &ok);
    h.cdf = upload(p, cdf.data(), cdf.size(), &ok);
    h.scale_count = S;
    h.factor_count = X;
    h.current_planes = planes;
    h.amplitude_planes = d->amplitude_planes;
    h.form_count = d->form_count;
    h.outputs = d->outputs;
    h.form_terms = upload(p, d->form_terms, 2 * (size_t)d->form_count, &ok);
    h.term_slots = upload(p, d->term_slots, d->term_count, &ok);
    h.coefficients = upload(p, d->term_coefficients, d->term_count, &ok);
    h.offsets = upload(p, d->offsets, (size_t)d->outputs + 1, &ok);
    h.planes = upload(p, d->planes, count, &ok);
    h.weights = upload(p, d->weights, count, &ok);
    h.bindings = upload(p, cp.bindings.data(), cp.bindings.size(), &ok);
    p->symjit = cp.symjit;
    h.numerator = d->numerator;
    h.denominator = d->denominator;
    h.factor = d->numerator != d->denominator;
    for (int cls = 0; cls < 2; cls++)
        for (int s = 0; s < COOP_SHAPES; s++) {
            const CoopProgram &q = cp.coop[cls][s];
            if (!q.ok) continue;
            h.coop_prog[cls][s] = upload(p, q.ops.data(), q.ops.size(), &ok);
            h.coop_start[cls][s] = upload(p, q.start.data(), q.start.size(), &ok);
            h.coop_split[cls][s] = upload(p, q.split.data(), q.split.size(), &ok);
            h.coop_form[cls][s] = upload(p, q.form_field.data(), q.form_field.size(), &ok);
            std::vector<uint32_t> red(count);
            for (uint32_t k = 0; k < count; k++) red[k] = q.amp_field[d->planes[k]];
            h.coop_red[cls][s] = upload(p, red.data(), red.size(), &ok);
            h.coop_slot[cls][s] = q.slot_planes;
            // (the slot's current planes at 1 KB whatever the class, as fused_kernel lays them out)
            p->coop_slot_doubles[cls][s] =
                (uint64_t)LANES * (2ull * q.slot_planes + 2ull * d->amplitude_planes + 4ull * d->form_count);
            p->coop_slot_bytes[cls][s] = q.slot_accesses * (cls ? 8u : 16u);
        }
    p->device = upload(p, &h, 1, &ok);
    p->slot_doubles = (uint64_t)LANES * (2ull * planes + 2ull * d->amplitude_planes + 4ull * d->form_count);
    if (!ok) {
        pacuda_plan_destroy(p);
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    }
    *out = p;
    return PACMAN_RECURRENCE_OK;
}

extern "C" void pacuda_plan_destroy(pacuda_plan *p) {
    if (!p) return;
    (void)hipDeviceSynchronize();
    for (void *a : p->allocations) (void)hipFree(a);
    delete p;
}

extern "C" int pacuda_plan_info(const pacuda_plan *p, uint32_t *instructions, uint64_t *slot_doubles) {
    if (!p) return PACMAN_RECURRENCE_INVALID;
    if (instructions) *instructions = p->host.count;
    if (slot_doubles) *slot_doubles = p->slot_doubles;
    return PACMAN_RECURRENCE_OK;
}

namespace {

// A launch's schedule: the item table in queue order, the units' first indices, the grid.
// Items go into the queue longest first (instructions x units, stable), so the units of long
// plans are handed out before the short ones and the launch's tail is short.
struct Schedule {
    std::vector<ItemDev> items;
    std::vector<uint32_t> starts;
    uint64_t slot = 0, units = 0, waves = 0;
    uint32_t grid = 0, block = 0, per_cu = 0;
    uint32_t lds_planes[LDS_VARIANTS] = {};  // per variant, the most any item's plan uses
    bool lds_all[LDS_VARIANTS] = {true, true, true, true, true};  // every plan has the variant
    bool symjit = false;  // some item's plan holds symjit rows
    int coop = -1;        // the cooperative shape, or -1
    int coop_class = 0;   // its programs' class of current (0: 16 bytes, 1: the f32 kernels' 8)
    uint64_t slots = 0;   // scratch slots: one per wavefront, or per workgroup in a cooperative launch
    uint32_t slots_per_block() const { return coop >= 0 ? 1u : block / LANES; }
};

bool coop_default(uint32_t kernel) {
    return kernel == PACMAN_RECURRENCE_KERNEL_F64 || kernel == PACMAN_RECURRENCE_KERNEL_MIXED ||
           kernel == PACMAN_RECURRENCE_KERNEL_MIXED64 || kernel == PACUDA_KERNEL_COMPENSATED64;
}

// the kernels whose currents are f32
bool f32_kernel(uint32_t kernel) {
    return kernel == PACMAN_RECURRENCE_KERNEL_MIXED || kernel == PACMAN_RECURRENCE_KERNEL_MIXED64 ||
           kernel == PACUDA_KERNEL_COMPENSATED || kernel == PACUDA_KERNEL_COMPENSATED64;
}

// The LDS variant of a launch, its LDS elements per wavefront and bytes per workgroup. A variant
// qualifies when every plan has it and the workgroups a CU must hold for the requested
// wavefronts fit in its LDS; the largest one is taken, else none (per_wave 0).
void lds_shape(uint32_t kernel, const Schedule &sc, uint32_t *per_wave, size_t *bytes, uint32_t *variant) {
    *per_wave = 0;
    *bytes = 0;
    *variant = 0;
    // current bytes per plane and lane: f64 and df64 (a float pair) 8, the mixed kernels' f32 4
    const size_t elem = f32_kernel(kernel) ? 4 : 8;
    const uint32_t waves_per_block = sc.block / LANES;
    const uint32_t blocks_per_cu = (sc.per_cu + waves_per_block - 1) / waves_per_block;
    for (int v = LDS_VARIANTS - 1; v >= 0; v--) {
        if (!sc.lds_all[v] || !sc.lds_planes[v]) continue;
        const size_t elems = 2ull * sc.lds_planes[v] * LANES;
        const size_t b = elem * elems * waves_per_block;
        if (b * blocks_per_cu > LDS_BYTES) continue;
        *per_wave = (uint32_t)elems;
        *bytes = b;
        *variant = (uint32_t)v;
        return;
    }
}

int schedule(const pacuda_item *items, uint32_t count, const pacuda_launch *l, Schedule &sc) {
    if (!l || l->abi_version != PACMAN_RECURRENCE_ABI || l->struct_size != sizeof(pacuda_launch) ||
        l->kernel > PACUDA_KERNEL_COMPENSATED64 || (!l->unit_waves && ((l->block % LANES) != 0 || l->block > 256)) ||
        l->waves_per_cu > 64 || (count && !items))
        return PACMAN_RECURRENCE_INVALID;
    if (l->unit_waves) {  // a cooperative launch: its shape
        const uint32_t per_cu = l->waves_per_cu ? l->waves_per_cu : 16;
        for (int k = 0; k < COOP_SHAPES; k++)
            if (COOP_WAVES[k] == l->unit_waves && COOP_WAVES[k] * COOP_UNITS[k] == per_cu) sc.coop = k;
        sc.coop_class = f32_kernel(l->kernel) ? 1 : 0;
        if (sc.coop < 0 || (l->kernel != PACMAN_RECURRENCE_KERNEL_F64 && !coop_built(1, sc.coop))) {
            fprintf(stderr, "pacuda: no cooperative launch of kernel %u with %u wavefronts per unit and %u per CU\n",
                    l->kernel, l->unit_waves, per_cu);
            return PACMAN_RECURRENCE_INVALID;
        }
    } else if (!l->waves_per_cu && !l->block && coop_default(l->kernel)) {
        // the default launch: cooperative, 8 wavefronts per unit and two units per CU (the
        // fastest shape measured for these kernels), when every plan has its program; 16
        // wavefronts and one unit per CU when the points would move more than COOP_WIDE_BYTES
        // each through the slot and every plan has that program too
        const int cls = f32_kernel(l->kernel) ? 1 : 0;
        bool all = count > 0, wide_ok = true;
        uint64_t points = 0, bytes = 0;
        for (uint32_t k = 0; k < count && all; k++) {
            all = items[k].plan && items[k].plan->coop_slot_doubles[cls][COOP_DEFAULT];
            if (!all) break;
            wide_ok = wide_ok && items[k].plan->coop_slot_doubles[cls][0];
            points += items[k].points;
            bytes += (uint64_t)items[k].points * items[k].plan->coop_slot_bytes[cls][COOP_DEFAULT];
        }
        if (all) {
            sc.coop = wide_ok && bytes > COOP_WIDE_BYTES * points ? 0 : COOP_DEFAULT;
            sc.coop_class = cls;
        }
    }
    std::vector<uint32_t> order(count);
    std::vector<uint64_t> cost(count);
    for (uint32_t k = 0; k < count; k++) {
        const pacuda_item &it = items[k];
        if (!it.plan || (it.points && (!it.external || it.ext_stride < it.points))) return PACMAN_RECURRENCE_INVALID;
        order[k] = k;
        cost[k] = (uint64_t)it.plan->host.count * ((it.points + LANES - 1) / LANES);
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    sc.items.resize(count);
    sc.starts.assign(count + 1, 0);
    for (uint32_t j = 0; j < count; j++) {
        const pacuda_item &it = items[order[j]];
        ItemDev &d = sc.items[j];
        d.plan = it.plan->device;
        d.external = it.external;
        d.resolved = it.resolved;
        d.total = it.total;
        d.points = it.points;
        d.ext_stride = it.ext_stride;
        sc.starts[j] = (uint32_t)sc.units;
        sc.units += (it.points + LANES - 1) / LANES;
        if (sc.coop >= 0 && !it.plan->coop_slot_doubles[sc.coop_class][sc.coop]) {
            fprintf(stderr, "pacuda: a plan of the launch has no cooperative program (symjit rows)\n");
            return PACMAN_RECURRENCE_INVALID;
        }
        sc.slot = std::max<uint64_t>(sc.slot, sc.coop >= 0 ? it.plan->coop_slot_doubles[sc.coop_class][sc.coop]
                                                           : it.plan->slot_doubles);
        sc.symjit = sc.symjit || it.plan->symjit;
        for (int v = 0; v < LDS_VARIANTS; v++) {
            sc.lds_planes[v] = std::max(sc.lds_planes[v], it.plan->host.lds_planes[v]);
            sc.lds_all[v] = sc.lds_all[v] && it.plan->host.lds_planes[v] > 0;
        }
    }
    if (sc.coop >= 0 && sc.symjit && sc.coop != 0 && sc.coop != 2) {
        fprintf(stderr, "pacuda: plans with symjit rows run cooperatively with 16 wavefronts per CU only\n");
        return PACMAN_RECURRENCE_INVALID;
    }
    // (a slot is addressed by 32-bit buffer offsets)
    if (sc.units >= 0xffffffffull || sc.slot * sizeof(double) > 0x7fffffffull) return PACMAN_RECURRENCE_INVALID;
    sc.starts[count] = (uint32_t)sc.units;
    if (sc.coop >= 0) {  // one slot per workgroup, COOP_UNITS workgroups per CU
        sc.block = COOP_WAVES[sc.coop] * LANES;
        sc.per_cu = COOP_WAVES[sc.coop] * COOP_UNITS[sc.coop];
        sc.grid = (uint32_t)std::min<uint64_t>((uint64_t)compute_units() * COOP_UNITS[sc.coop], std::max<uint64_t>(sc.units, 1));
        sc.waves = (uint64_t)sc.grid * COOP_WAVES[sc.coop];
        sc.slots = sc.grid;
        return PACMAN_RECURRENCE_OK;
    }
    sc.block = l->block ? l->block : 256;
    // default occupancy per kernel (measured on g g > g g g and g g > g g g g)
    const uint32_t per_cu = l->waves_per_cu                                  ? l->waves_per_cu
                            : l->kernel == PACMAN_RECURRENCE_KERNEL_F64     ? 12
                            : l->kernel == PACMAN_RECURRENCE_KERNEL_DF64    ? 16
                                                                            : 16;
    sc.per_cu = per_cu;
    const uint32_t waves_per_block = sc.block / LANES;
    uint64_t waves = std::min<uint64_t>((uint64_t)compute_units() * per_cu, std::max<uint64_t>(sc.units, 1));
    sc.grid = (uint32_t)((waves + waves_per_block - 1) / waves_per_block);
    sc.waves = (uint64_t)sc.grid * waves_per_block;
    sc.slots = sc.waves;
    return PACMAN_RECURRENCE_OK;
}

template <class K, int TIER, bool SJ> bool allow_lds() {
    static const bool ok = !hip_failed(hipFuncSetAttribute(reinterpret_cast<const void *>(fused_kernel<K, TIER, SJ>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, LDS_BYTES),
                                       "fused kernel LDS");
    return ok;
}

// The register tier of a launch of `kernel` at `per_cu` wavefronts per CU: the largest whose
// wavefronts fit the CU's registers
template <class K> int tier_of(uint32_t per_cu) {
    return per_cu <= 4u * (uint32_t)waves_per_simd<K, 2>()   ? 2
           : per_cu <= 4u * (uint32_t)waves_per_simd<K, 1>() ? 1
                                                             : 0;
}
int tier_of(uint32_t kernel, uint32_t per_cu) {
    if (kernel == PACUDA_KERNEL_COMPENSATED) return tier_of<KComp>(per_cu);
    if (kernel == PACUDA_KERNEL_COMPENSATED64) return tier_of<KComp64>(per_cu);
    return f32_kernel(kernel) ? tier_of<KMixed>(per_cu) : tier_of<KF64>(per_cu);
}

template <class K> bool allow_lds(int tier, bool sj) {
    if (sj) return tier == 2 ? allow_lds<K, 2, true>() : tier == 1 ? allow_lds<K, 1, true>() : allow_lds<K, 0, true>();
    return tier == 2 ? allow_lds<K, 2, false>() : tier == 1 ? allow_lds<K, 1, false>() : allow_lds<K, 0, false>();
}

template <class K, int TIER>
void launch_tier(const Schedule &sc, hipStream_t s, size_t bytes, const ItemDev *items, const uint32_t *starts,
                 uint32_t count, uint32_t *counter, double *scratch, uint32_t per_wave, uint32_t variant) {
    if (sc.symjit)
        fused_kernel<K, TIER, true><<<sc.grid, sc.block, bytes, s>>>(items, starts, count, counter, scratch, sc.slot,
                                                                     per_wave, variant, LDS_CAPS[variant]);
    else
        fused_kernel<K, TIER, false><<<sc.grid, sc.block, bytes, s>>>(items, starts, count, counter, scratch, sc.slot,
                                                                      per_wave, variant, LDS_CAPS[variant]);
}

template <class K>
bool launch(const Schedule &sc, hipStream_t s, int tier, size_t bytes, const ItemDev *items, const uint32_t *starts,
            uint32_t count, uint32_t *counter, double *scratch, uint32_t per_wave, uint32_t variant) {
    if (bytes && !allow_lds<K>(tier, sc.symjit)) return false;
    if (tier == 1) launch_tier<K, 1>(sc, s, bytes, items, starts, count, counter, scratch, per_wave, variant);
    else if (tier == 2) launch_tier<K, 2>(sc, s, bytes, items, starts, count, counter, scratch, per_wave, variant);
    else launch_tier<K, 0>(sc, s, bytes, items, starts, count, counter, scratch, per_wave, variant);
    return true;
}

template <class K, int SHAPE, bool SJ> bool allow_coop_lds() {
    static const bool ok = !hip_failed(hipFuncSetAttribute(reinterpret_cast<const void *>(coop_kernel<K, SHAPE, SJ>),
                                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                                           LDS_BYTES / COOP_UNITS[SHAPE]),
                                       "cooperative kernel LDS");
    return ok;
}

template <class K, int SHAPE, bool SJ>
bool launch_coop_kernel(const Schedule &sc, hipStream_t s, const ItemDev *items, const uint32_t *starts,
                        uint32_t count, uint32_t *counter, double *scratch, bool run) {
    if (!allow_coop_lds<K, SHAPE, SJ>()) return false;
    if (!run) return true;
    const uint32_t bytes = LDS_BYTES / COOP_UNITS[SHAPE];
    reset_counter<<<1, 64, 0, s>>>(counter);
    coop_kernel<K, SHAPE, SJ><<<sc.grid, COOP_WAVES[SHAPE] * LANES, bytes, s>>>(items, starts, count, counter,
                                                                            scratch, sc.slot, bytes);
    return true;
}

// (the instance with symjit rows when a plan of the launch holds some: shapes 0 and 2)
template <class K, int SHAPE>
bool launch_coop_shape(const Schedule &sc, hipStream_t s, const ItemDev *items, const uint32_t *starts, uint32_t count,
                       uint32_t *counter, double *scratch, bool run) {
    if constexpr (SHAPE == 0 || SHAPE == 2) {
        if (sc.symjit) return launch_coop_kernel<K, SHAPE, true>(sc, s, items, starts, count, counter, scratch, run);
    } else {
        if (sc.symjit) return false;  // (schedule refuses it)
    }
    return launch_coop_kernel<K, SHAPE, false>(sc, s, items, starts, count, counter, scratch, run);
}

// the policies other than f64 run shapes 0 and 2 only (coop_built)
template <class K>
bool launch_coop_policy(const Schedule &sc, hipStream_t s, const ItemDev *items, const uint32_t *starts,
                        uint32_t count, uint32_t *counter, double *scratch, bool run) {
    if (sc.coop == 0) return launch_coop_shape<K, 0>(sc, s, items, starts, count, counter, scratch, run);
    if (sc.coop == 2) return launch_coop_shape<K, 2>(sc, s, items, starts, count, counter, scratch, run);
    return false;
}
static_assert(COOP_SHAPES == 4, "the cooperative kernels instantiated below");

// a cooperative launch of `kernel` (run false: only the LDS attribute, set before any graph capture)
bool launch_coop(uint32_t kernel, const Schedule &sc, hipStream_t s, const ItemDev *items, const uint32_t *starts,
                 uint32_t count, uint32_t *counter, double *scratch, bool run = true) {
    switch (kernel) {
    case PACMAN_RECURRENCE_KERNEL_F64:
        switch (sc.coop) {
        case 0: return launch_coop_shape<KF64, 0>(sc, s, items, starts, count, counter, scratch, run);
        case 1: return launch_coop_shape<KF64, 1>(sc, s, items, starts, count, counter, scratch, run);
        case 2: return launch_coop_shape<KF64, 2>(sc, s, items, starts, count, counter, scratch, run);
        default: return launch_coop_shape<KF64, 3>(sc, s, items, starts, count, counter, scratch, run);
        }
    case PACMAN_RECURRENCE_KERNEL_MIXED:
        return launch_coop_policy<KMixed>(sc, s, items, starts, count, counter, scratch, run);
    case PACMAN_RECURRENCE_KERNEL_MIXED64:
        return launch_coop_policy<KMixed64>(sc, s, items, starts, count, counter, scratch, run);
    case PACUDA_KERNEL_COMPENSATED:
        return launch_coop_policy<KComp>(sc, s, items, starts, count, counter, scratch, run);
    case PACUDA_KERNEL_COMPENSATED64:
        return launch_coop_policy<KComp64>(sc, s, items, starts, count, counter, scratch, run);
    default:
        return launch_coop_policy<KDF64>(sc, s, items, starts, count, counter, scratch, run);
    }
}

// the LDS attribute of the kernel a launch of `sc` needs, set once (before any graph capture)
bool lds_ready(uint32_t kernel, const Schedule &sc) {
    if (sc.coop >= 0) return launch_coop(kernel, sc, nullptr, nullptr, nullptr, 0, nullptr, nullptr, false);
    uint32_t per_wave, variant;
    size_t bytes;
    lds_shape(kernel, sc, &per_wave, &bytes, &variant);
    if (!bytes) return true;
    const int tier = tier_of(kernel, sc.per_cu);
    if (kernel == PACMAN_RECURRENCE_KERNEL_F64) return allow_lds<KF64>(tier, sc.symjit);
    if (kernel == PACMAN_RECURRENCE_KERNEL_MIXED) return allow_lds<KMixed>(tier, sc.symjit);
    if (kernel == PACMAN_RECURRENCE_KERNEL_MIXED64) return allow_lds<KMixed64>(tier, sc.symjit);
    if (kernel == PACUDA_KERNEL_COMPENSATED) return allow_lds<KComp>(tier, sc.symjit);
    if (kernel == PACUDA_KERNEL_COMPENSATED64) return allow_lds<KComp64>(tier, sc.symjit);
    return allow_lds<KDF64>(tier, sc.symjit);
}

bool launch_fused(uint32_t kernel, const Schedule &sc, hipStream_t s, const ItemDev *items, const uint32_t *starts,
                  uint32_t count, uint32_t *counter, double *scratch) {
    if (sc.coop >= 0) return launch_coop(kernel, sc, s, items, starts, count, counter, scratch);
    uint32_t per_wave, variant;
    size_t bytes;
    lds_shape(kernel, sc, &per_wave, &bytes, &variant);
    reset_counter<<<1, 64, 0, s>>>(counter);
    const int tier = tier_of(kernel, sc.per_cu);
    switch (kernel) {
    case PACMAN_RECURRENCE_KERNEL_F64:
        return launch<KF64>(sc, s, tier, bytes, items, starts, count, counter, scratch, per_wave, variant);
    case PACMAN_RECURRENCE_KERNEL_MIXED:
        return launch<KMixed>(sc, s, tier, bytes, items, starts, count, counter, scratch, per_wave, variant);
    case PACMAN_RECURRENCE_KERNEL_MIXED64:
        return launch<KMixed64>(sc, s, tier, bytes, items, starts, count, counter, scratch, per_wave, variant);
    case PACUDA_KERNEL_COMPENSATED:
        return launch<KComp>(sc, s, tier, bytes, items, starts, count, counter, scratch, per_wave, variant);
    case PACUDA_KERNEL_COMPENSATED64:
        return launch<KComp64>(sc, s, tier, bytes, items, starts, count, counter, scratch, per_wave, variant);
    default:
        return launch<KDF64>(sc, s, tier, bytes, items, starts, count, counter, scratch, per_wave, variant);
    }
}

}  // namespace

namespace {

// the launch of a schedule of `count` items on `s` with the library's shared item table, counter
// and slots; kernel_ms as pacuda_launch's
int launch_schedule(const Schedule &sc, uint32_t count, uint32_t kernel, hipStream_t s, float *kernel_ms) {
    if (count == 0 || sc.units == 0) return PACMAN_RECURRENCE_OK;
    std::lock_guard<std::mutex> guard(fs.lock);
    const size_t need = (size_t)sc.slots * sc.slot;
    if (!s) return PACMAN_RECURRENCE_DEVICE_ERROR;
    // the previous launch (on any stream) reads the item table and uses the counter and slots
    // until it ends
    if (fs.last_pending) {
        if (hip_failed(hipEventSynchronize(fs.last), "previous launch")) return PACMAN_RECURRENCE_DEVICE_ERROR;
        fs.last_pending = false;
    }
    if (need > fs.scratch_doubles) {
        if (fs.scratch) {
            (void)hipDeviceSynchronize();
            (void)hipFree(fs.scratch);
        }
        fs.scratch = nullptr;
        fs.scratch_doubles = 0;
        if (hip_failed(hipMalloc((void **)&fs.scratch, need * sizeof(double)), "scratch")) return PACMAN_RECURRENCE_DEVICE_ERROR;
        fs.scratch_doubles = need;
    }
    if (count > fs.item_capacity) {
        (void)hipDeviceSynchronize();
        if (fs.items) (void)hipFree(fs.items);
        if (fs.unit_start) (void)hipFree(fs.unit_start);
        fs.items = nullptr;
        fs.unit_start = nullptr;
        fs.item_capacity = 0;
        if (hip_failed(hipMalloc((void **)&fs.items, count * sizeof(ItemDev)), "items") ||
            hip_failed(hipMalloc((void **)&fs.unit_start, (count + 1) * sizeof(uint32_t)), "items"))
            return PACMAN_RECURRENCE_DEVICE_ERROR;
        fs.item_capacity = count;
    }
    if (!fs.counter && hip_failed(hipMalloc((void **)&fs.counter, 64), "counter")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (!fs.start && (hip_failed(hipEventCreate(&fs.start), "event") || hip_failed(hipEventCreate(&fs.stop), "event") ||
                      hip_failed(hipEventCreateWithFlags(&fs.last, hipEventDisableTiming), "event")))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    // the item table: the previous launch has ended (above), and the host copies go out of scope
    // on return, so copy synchronously; the caller's earlier work on `s` (which may write the
    // momenta this launch reads) stays ordered before the launch by the stream itself
    if (hip_failed(hipMemcpy(fs.items, sc.items.data(), count * sizeof(ItemDev), hipMemcpyHostToDevice), "items") ||
        hip_failed(hipMemcpy(fs.unit_start, sc.starts.data(), (count + 1) * sizeof(uint32_t), hipMemcpyHostToDevice), "items"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (kernel_ms) (void)hipEventRecord(fs.start, s);
    if (!launch_fused(kernel, sc, s, fs.items, fs.unit_start, count, fs.counter, fs.scratch) ||
        hip_failed(hipGetLastError(), "fused launch") || hip_failed(hipEventRecord(fs.last, s), "event"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    fs.last_pending = true;
    if (kernel_ms) {
        if (hip_failed(hipEventRecord(fs.stop, s), "event") || hip_failed(hipEventSynchronize(fs.stop), "kernel"))
            return PACMAN_RECURRENCE_DEVICE_ERROR;
        float ms = 0;
        if (hip_failed(hipEventElapsedTime(&ms, fs.start, fs.stop), "event")) return PACMAN_RECURRENCE_DEVICE_ERROR;
        *kernel_ms = ms;
    }
    return PACMAN_RECURRENCE_OK;
}

}  // namespace

extern "C" int pacuda_evaluate(const pacuda_item *items, uint32_t count, const pacuda_launch *l) {
    Schedule sc;
    int st = schedule(items, count, l, sc);
    if (st) return st;
    return launch_schedule(sc, count, l->kernel, l->stream ? (hipStream_t)l->stream : runner_stream(), l->kernel_ms);
}

// ---- v1 runs on the fused kernels ---------------------------------------------------------------

namespace pacuda {

pacuda_plan *v1_fused_plan(const pacman_recurrence_call &c, bool sort, std::vector<uint32_t> &amplitudes) {
    amplitudes.clear();
    if (c.momentum_planes % 4 || c.current_planes >= LDS_PLANE) return nullptr;
    const CompiledProgram hp =
        compile_program(c.words, c.word_count, c.width, sort, {PACUDA_KERNEL_COUNT, PACUDA_MAX_IN, PACUDA_MAX_OUT});
    const uint32_t forms = c.momentum_planes / 4;
    if (hp.error || !program_fits(hp.need, c.current_planes, c.amplitude_planes, forms, c.scale_count, c.factor_count) ||
        !reads_after_writes(hp.ops, c.current_planes, &hp.bindings))
        return nullptr;
    std::vector<char> written(c.amplitude_planes, 0);
    for (const Instr &o : hp.ops)
        if ((o.w[0] & 0xff) == K_CLOSURE) written[o.w[3]] = 1;  // (in range: program_fits)
    for (uint32_t q = 0; q < c.amplitude_planes; q++)
        if (written[q]) amplitudes.push_back(q);
    if (amplitudes.empty()) return nullptr;
    // the plan's "reduction": one output per written amplitude plane, which a v1 launch stores
    const uint32_t n = (uint32_t)amplitudes.size();
    std::vector<uint32_t> offsets(n + 1);
    for (uint32_t k = 0; k <= n; k++) offsets[k] = k;
    const std::vector<double> weights(n, 1.0);
    pacuda_plan_desc d;
    memset(&d, 0, sizeof d);
    d.abi_version = PACMAN_RECURRENCE_ABI;
    d.struct_size = sizeof d;
    d.words = c.words;
    d.word_count = c.word_count;
    memcpy(d.width, c.width, sizeof d.width);
    d.current_planes = c.current_planes;
    d.amplitude_planes = c.amplitude_planes;
    d.form_count = forms;
    d.scale_re = c.scale_re;
    d.scale_im = c.scale_im;
    d.scale_count = c.scale_count;
    d.factor_re = c.factor_re;
    d.factor_im = c.factor_im;
    d.factor_count = c.factor_count;
    d.external_count = c.external_count;
    d.term_count = c.term_count;
    d.form_terms = c.form_terms;
    d.term_slots = c.term_slots;
    d.term_coefficients = c.term_coefficients;
    d.outputs = n;
    d.offsets = offsets.data();
    d.planes = amplitudes.data();
    d.weights = weights.data();
    d.numerator = d.denominator = 1.0;
    d.no_sort = sort ? 0 : 1;
    pacuda_plan *p = nullptr;
    return pacuda_plan_create(&d, &p) == PACMAN_RECURRENCE_OK ? p : nullptr;
}

int v1_fused_run(const pacuda_plan *plan, uint32_t kernel, const double *external, uint32_t points, uint32_t ext_stride,
                 double *are, double *aim, size_t stride, hipStream_t s) {
    const pacuda_item item{plan, external, points, ext_stride, nullptr, nullptr};
    pacuda_launch l;
    memset(&l, 0, sizeof l);
    l.abi_version = PACMAN_RECURRENCE_ABI;
    l.struct_size = sizeof l;
    l.kernel = kernel;
    Schedule sc;
    if (const int st = schedule(&item, 1, &l, sc)) return st;
    sc.items[0].amp_re = are;
    sc.items[0].amp_im = aim;
    sc.items[0].amp_stride = stride;
    return launch_schedule(sc, 1, kernel, s, nullptr);
}

}  // namespace pacuda

// ---- the evaluation without allocation ---------------------------------------------------------
//
// pacuda_evaluate_static runs the same launch as pacuda_evaluate with every buffer it needs
// beyond the caller's inputs and outputs (the work counter, the item table, the units' first
// indices and the wavefronts' slots) in one file-scope device array, so it never allocates.
// Layout of the array: [counter, 256 B][items, STATIC_ITEMS x ItemDev][unit starts,
// STATIC_ITEMS + 1 u32][slots, the rest]. Each GPU holds its own copy; its address comes from
// hipGetSymbolAddress once per device and is the only name host code uses for it.
namespace {

constexpr size_t STATIC_BYTES = 256ull << 20;
constexpr uint32_t STATIC_ITEMS = 4096;
constexpr size_t STATIC_ITEMS_AT = 256;
constexpr size_t STATIC_STARTS_AT = STATIC_ITEMS_AT + ((STATIC_ITEMS * sizeof(ItemDev) + 255) & ~(size_t)255);
constexpr size_t STATIC_SLOTS_AT = STATIC_STARTS_AT + (((STATIC_ITEMS + 1) * sizeof(uint32_t) + 255) & ~(size_t)255);
static_assert(STATIC_SLOTS_AT < STATIC_BYTES / 64, "the static scratch is mostly slots");

static __device__ __align__(256) double scratch[STATIC_BYTES / sizeof(double)];

constexpr int STATIC_DEVICES = 64;
std::mutex static_lock;
char *static_base[STATIC_DEVICES];  // per device: the array's device address (nullptr: not yet)

// the device address of the current device's array, looked up once per device
char *static_area(int *device) {
    if (hip_failed(hipGetDevice(device), "device") || *device < 0 || *device >= STATIC_DEVICES) return nullptr;
    if (!static_base[*device]) {
        void *p = nullptr;
        if (hip_failed(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "static scratch address")) return nullptr;
        static_base[*device] = (char *)p;
    }
    return static_base[*device];
}

}  // namespace

extern "C" int pacuda_evaluate_static(const pacuda_item *items, uint32_t count, const pacuda_launch *l) {
    Schedule sc;
    int st = schedule(items, count, l, sc);
    if (st) return st;
    if (l->kernel_ms) {
        fprintf(stderr, "pacuda: pacuda_evaluate_static does not time its launch (kernel_ms must be NULL)\n");
        return PACMAN_RECURRENCE_INVALID;
    }
    if (count > STATIC_ITEMS) {
        fprintf(stderr, "pacuda: pacuda_evaluate_static takes at most %u items (%u given)\n", STATIC_ITEMS, count);
        return PACMAN_RECURRENCE_INVALID;
    }
    if (count == 0 || sc.units == 0) return PACMAN_RECURRENCE_OK;
    // the resident workgroups whose slots fit: at most the launch's, at least one (a slot per
    // wavefront, or per workgroup in a cooperative launch)
    const size_t slot_bytes = (size_t)sc.slot * sizeof(double);
    const uint64_t fit = (STATIC_BYTES - STATIC_SLOTS_AT) / std::max<size_t>(slot_bytes, 1);
    const uint32_t slots_per_block = sc.slots_per_block();
    if (fit < slots_per_block) {
        fprintf(stderr,
                "pacuda: pacuda_evaluate_static: %u slots of %.1f MB exceed the static scratch of %zu MB\n",
                slots_per_block, slot_bytes / 1e6, STATIC_BYTES >> 20);
        return PACMAN_RECURRENCE_INVALID;
    }
    if (sc.slots > fit) {
        sc.grid = (uint32_t)(fit / slots_per_block);
        sc.waves = (uint64_t)sc.grid * (sc.block / LANES);
        sc.slots = (uint64_t)sc.grid * slots_per_block;
    }
    std::lock_guard<std::mutex> guard(static_lock);
    int device = 0;
    char *base = static_area(&device);
    if (!base) return PACMAN_RECURRENCE_DEVICE_ERROR;
    uint32_t *counter = (uint32_t *)base;
    ItemDev *table = (ItemDev *)(base + STATIC_ITEMS_AT);
    uint32_t *starts = (uint32_t *)(base + STATIC_STARTS_AT);
    double *slots = (double *)(base + STATIC_SLOTS_AT);
    // (the null stream when the caller gives none: the library's own stream is created on first
    // use, which this entry point must not do)
    hipStream_t s = (hipStream_t)l->stream;
    // the previous call's launch may still read the table: wait for the stream, then copy
    // synchronously (the host copies go out of scope on return)
    if (hip_failed(hipStreamSynchronize(s), "static items") ||
        hip_failed(hipMemcpy(table, sc.items.data(), count * sizeof(ItemDev), hipMemcpyHostToDevice), "static items") ||
        hip_failed(hipMemcpy(starts, sc.starts.data(), (count + 1) * sizeof(uint32_t), hipMemcpyHostToDevice),
                   "static items"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (!launch_fused(l->kernel, sc, s, table, starts, count, counter, slots) ||
        hip_failed(hipGetLastError(), "static fused launch"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}

// A sweep: one batch of items with its own item table, counter and scratch, captured once into a
// HIP graph (the counter reset kernel and the fused kernel) and replayed by pacuda_sweep_run.
struct pacuda_sweep {
    Schedule sc;
    uint32_t kernel = 0, count = 0;
    ItemDev *items = nullptr;
    uint32_t *starts = nullptr, *counter = nullptr;
    double *scratch = nullptr;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
};

extern "C" void pacuda_sweep_destroy(pacuda_sweep *w) {
    if (!w) return;
    (void)hipDeviceSynchronize();
    if (w->exec) (void)hipGraphExecDestroy(w->exec);
    if (w->graph) (void)hipGraphDestroy(w->graph);
    if (w->capture) (void)hipStreamDestroy(w->capture);
    if (w->start) (void)hipEventDestroy(w->start);
    if (w->stop) (void)hipEventDestroy(w->stop);
    for (void *p : {(void *)w->items, (void *)w->starts, (void *)w->counter, (void *)w->scratch})
        if (p) (void)hipFree(p);
    delete w;
}

extern "C" int pacuda_sweep_create(const pacuda_item *items, uint32_t count, const pacuda_launch *l, pacuda_sweep **out) {
    if (!out) return PACMAN_RECURRENCE_INVALID;
    *out = nullptr;
    pacuda_sweep *w = new pacuda_sweep();
    int st = schedule(items, count, l, w->sc);
    if (st || count == 0 || w->sc.units == 0) {
        delete w;
        return st ? st : PACMAN_RECURRENCE_INVALID;
    }
    w->kernel = l->kernel;
    w->count = count;
    const size_t need = (size_t)w->sc.slots * w->sc.slot;
    bool ok = !hip_failed(hipMalloc((void **)&w->items, count * sizeof(ItemDev)), "sweep") &&
              !hip_failed(hipMalloc((void **)&w->starts, (count + 1) * sizeof(uint32_t)), "sweep") &&
              !hip_failed(hipMalloc((void **)&w->counter, 64), "sweep") &&
              !hip_failed(hipMalloc((void **)&w->scratch, std::max<size_t>(need, 1) * sizeof(double)), "sweep") &&
              !hip_failed(hipMemcpy(w->items, w->sc.items.data(), count * sizeof(ItemDev), hipMemcpyHostToDevice), "sweep") &&
              !hip_failed(hipMemcpy(w->starts, w->sc.starts.data(), (count + 1) * sizeof(uint32_t), hipMemcpyHostToDevice),
                          "sweep") &&
              !hip_failed(hipEventCreate(&w->start), "sweep") && !hip_failed(hipEventCreate(&w->stop), "sweep") &&
              !hip_failed(hipStreamCreateWithFlags(&w->capture, hipStreamNonBlocking), "sweep") &&
              lds_ready(w->kernel, w->sc) &&
              !hip_failed(hipStreamBeginCapture(w->capture, hipStreamCaptureModeThreadLocal), "sweep capture");
    if (ok) {
        const bool launched = launch_fused(w->kernel, w->sc, w->capture, w->items, w->starts, count, w->counter,
                                           w->scratch);
        ok = !hip_failed(hipStreamEndCapture(w->capture, &w->graph), "sweep capture") && launched &&
             !hip_failed(hipGraphInstantiate(&w->exec, w->graph, nullptr, nullptr, 0), "sweep graph");
    }
    if (!ok) {
        pacuda_sweep_destroy(w);
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    }
    *out = w;
    return PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_sweep_run(pacuda_sweep *w, void *stream, float *kernel_ms) {
    if (!w) return PACMAN_RECURRENCE_INVALID;
    hipStream_t s = stream ? (hipStream_t)stream : runner_stream();
    if (!s) return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (kernel_ms) (void)hipEventRecord(w->start, s);
    if (hip_failed(hipGraphLaunch(w->exec, s), "sweep launch")) return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (kernel_ms) {
        float ms = 0;
        if (hip_failed(hipEventRecord(w->stop, s), "event") || hip_failed(hipEventSynchronize(w->stop), "sweep") ||
            hip_failed(hipEventElapsedTime(&ms, w->start, w->stop), "event"))
            return PACMAN_RECURRENCE_DEVICE_ERROR;
        *kernel_ms = ms;
    }
    return PACMAN_RECURRENCE_OK;
}

extern "C" void *pacuda_device_alloc(size_t bytes) {
    void *p = nullptr;
    if (hip_failed(hipMalloc(&p, bytes ? bytes : 1), "alloc")) return nullptr;
    return p;
}
extern "C" void pacuda_device_free(void *p) {
    if (p) (void)hipFree(p);
}
extern "C" int pacuda_copy_to_device(void *dst, const void *src, size_t bytes) {
    hipStream_t s = runner_stream();
    if (!s || hip_failed(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "copy") ||
        hip_failed(hipStreamSynchronize(s), "copy"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}
extern "C" int pacuda_copy_to_host(void *dst, const void *src, size_t bytes) {
    hipStream_t s = runner_stream();
    if (!s || hip_failed(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "copy") ||
        hip_failed(hipStreamSynchronize(s), "copy"))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    return PACMAN_RECURRENCE_OK;
}
extern "C" int pacuda_synchronize(void) {
    return hip_failed(hipDeviceSynchronize(), "synchronize") ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK;
}
extern "C" int pacuda_device_compute_units(void) { return compute_units(); }
