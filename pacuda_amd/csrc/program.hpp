// SPDX-License-Identifier: 0BSD
//
// The instruction format the host's plan compiler (plan_compiler.hpp) emits, and the
// interpreter loop (device side) that one lane runs for its point.
//
// Device: the instruction stream is the same for every lane of a wavefront, so each
// instruction is fetched with scalar loads into SGPRs and dispatched by a uniform branch; the
// arena values are vector loads of one plane for 64 consecutive points.
#pragma once
#include <stdint.h>

#include <vector>

#include "device.hpp"
#include "plane_kernels.hpp"

namespace pacuda {

enum { K_CLEAR = 1, K_CONTRIB, K_PROP, K_IDENTITY, K_CLOSURE, K_SOURCE, K_SYMJIT, K_BARRIER, K_FETCH };
// a symjit binding's input code: bits 0-2 the kind below, bit 3 which parent or form, bit 4 the
// imaginary part, bits 5 and up the component (or the scale slot) (runner/recurrence_runner.cu:
// 582, 631-684)
enum { SB_PARENT, SB_MOMENTUM, SB_DESTINATION, SB_SCALE, SB_FACTOR };
// F_HOLD: the row's result stays in the accumulators for the next row instead of being stored;
// F_FROMACC: an in-place row (propagator, identity) takes its input from the accumulators
// instead of loading it; F_ZERO: a closure that is the first to touch its amplitude starts from
// zero instead of loading the zero the arena holds (all three set only by the fused path's
// passes, never on v1 programs)
enum { F_INIT = 1, F_FIRST = 2, F_LAST = 4, F_HOLD = 8, F_FROMACC = 16, F_ZERO = 32 };

// w[0] = kind | op << 8 | flags << 16; contributions: w[1] scale slot, w[2] parent 0, w[3]
// parent 1, w[4] destination, w[5] form 0, w[6] form 1, w[7] factor; clear: w[1] base, w[2]
// count; propagator: w[1] base, w[2] form, w[3] factor; identity: w[1] base, w[2] count, w[3]
// factor; closure: w[1] parent 0, w[2] parent 1, w[3] amplitude, w[4] factor, w[5] first
// component factor, w[6] count; source: w[1] destination, w[2] form, w[3] factor, w[4] mass
// slot, w[5] template; symjit: w[1] parent 0, w[2] parent 1, w[3] destination, w[4] form 0 |
// form 1 << 16, w[5] factor, w[6] the binding's offset in the binding table (kernel, inputs,
// output pairs, direct, then the input codes and the output components); barrier (cooperative
// programs only): no operands, the wavefront waits for its workgroup; fetch (cooperative
// programs only, once per program): the wavefront takes the workgroup's next unit off the queue.
struct alignas(32) Instr {
    uint32_t w[8];
};

// the tables a program indexes (the extents a compiled program records per table)
enum { N_CURRENT, N_AMPLITUDE, N_FORM, N_SCALE, N_FACTOR, N_COUNT };
// true if a call with these table sizes keeps every index of the program in bounds
inline bool program_fits(const uint64_t *need, uint64_t currents, uint64_t amplitudes, uint64_t forms, uint64_t scales,
                         uint64_t factors) {
    return need[N_CURRENT] <= currents && need[N_AMPLITUDE] <= amplitudes && need[N_FORM] <= forms &&
           need[N_SCALE] <= scales && need[N_FACTOR] <= factors;
}


// ---- the interpreter (device) ----------------------------------------------------------------
// A is the arena (PlaneArena for the v1 path, SlotArena for the fused path). The fused programs
// carry known-zero component masks in spare bits (A::zero_masks): contribution: w[0] bits 22-27
// parent 0, bits 28-31 parent 1, w[1] bits 16-21 the destination's old value, w[7] bits 16-21
// the stored value (the scale slot and factor are bits 0-15 then); propagator: w[4] = load |
// store << 8; identity: w[4] = load | store << 16; closure: w[7] = parent 0 | parent 1 << 8;
// source: w[6] = store. Bit k: component k is zero for every input.
template <class A> PCD uint32_t field16(uint32_t w) { return A::zero_masks ? (w & 0xffffu) : w; }
template <class A> PCD uint32_t zbits(uint32_t w, int shift, uint32_t width) {
    return A::zero_masks ? (w >> shift) & ((1u << width) - 1u) : 0u;
}

// A contribution row in the policy's vertex arithmetic T (K::acc): the parents widened to T
// (exactly), the forms and the constants as T, the destination's sum in T, and the stored value
// rounded to R once.
template <class K, int OP, class A>
PCD void contribution(const A &a, const Instr &o, int flags, Cx<typename K::acc> *acc) {
    typedef typename K::real R;
    typedef typename K::den Den;
    typedef typename K::acc T;
    constexpr int c0 = Shape<OP>::c0, c1 = Shape<OP>::c1, cd = Shape<OP>::cd;
    const uint32_t dst = o.w[4], slot = field16<A>(o.w[1]), x = field16<A>(o.w[7]);
    const Cx<T> scale =
        mul(cx(ldc<T>(a.sre(), slot), ldc<T>(a.sim(), slot)), cx(ldc<T>(a.fre(), x), ldc<T>(a.fim(), x)));
    Cx<R> l[c0], r[c1];
    Cx<T> wl[c0], wr[c1], v[cd];
    T q0[4], q1[4];
    a.template curs<R, c0>(o.w[2], l, zbits<A>(o.w[0], 22, 6));
    a.template curs<R, c1>(o.w[3], r, zbits<A>(o.w[0], 28, 4));
#pragma unroll
    for (int k = 0; k < c0; k++) wl[k] = cvt<T>(l[k]);
#pragma unroll
    for (int k = 0; k < c1; k++) wr[k] = cvt<T>(r[k]);
    if constexpr (OP == 6) {
        double f0[4], f1[4];
        a.form(o.w[5], f0);
        a.form(o.w[6], f1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            q0[k] = Form<T, Den>::of(f0[k]);
            q1[k] = Form<T, Den>::of(f1[k]);
        }
    }
    contribution_values<OP>(wl, wr, q0, q1, v);
    if (!(flags & F_FIRST)) {
#pragma unroll
        for (int k = 0; k < cd; k++) acc[k] = add(acc[k], mul(v[k], scale));
    } else if (flags & F_INIT) {
#pragma unroll
        for (int k = 0; k < cd; k++) acc[k] = mul(v[k], scale);
    } else {
        Cx<R> old[cd];
        a.template curs<R, cd>(dst, old, zbits<A>(o.w[1], 16, 6));
#pragma unroll
        for (int k = 0; k < cd; k++) acc[k] = add(cvt<T>(old[k]), mul(v[k], scale));
    }
    if ((flags & F_LAST) && !(flags & F_HOLD)) {
        Cx<R> out[cd];
#pragma unroll
        for (int k = 0; k < cd; k++) out[k] = cvt<R>(acc[k]);
        a.template puts<R, cd>(dst, out, zbits<A>(o.w[7], 16, 6));
    }
}

// A propagator row in T: an input taken from the accumulators is first rounded to R (the value
// the stored current would have), so forwarding never changes a result.
template <class K, int OP, class A>
PCD void propagator(const A &a, const Instr &o, int flags, Cx<typename K::acc> *acc) {
    typedef typename K::real R;
    typedef typename K::den Den;
    typedef typename K::acc T;
    constexpr int m = OP == 9 ? 4 : 2;
    const uint32_t base = o.w[1];
    T p[4];
    Den q[4];
    double f[4];
    a.form(o.w[2], f);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        q[k] = from_raw<Den>(f[k]);
        p[k] = Form<T, Den>::of(f[k]);
    }
    const T reciprocal = ToReal<T, Den>::of(reciprocal_minkowski_norm(q[0], q[1], q[2], q[3]));
    Cx<R> c[m];
    Cx<T> wc[m], v[m];
    if (flags & F_FROMACC) {
#pragma unroll
        for (int k = 0; k < m; k++) c[k] = cvt<R>(acc[k]);
    } else {
        a.template curs<R, m>(base, c, zbits<A>(o.w[4], 0, 4));
    }
#pragma unroll
    for (int k = 0; k < m; k++) wc[k] = cvt<T>(c[k]);
    propagator_values<OP>(wc, p, reciprocal, cx(ldc<T>(a.fre(), o.w[3]), ldc<T>(a.fim(), o.w[3])), v);
    if (flags & F_HOLD) {
#pragma unroll
        for (int k = 0; k < m; k++) acc[k] = v[k];
    } else {
        Cx<R> out[m];
#pragma unroll
        for (int k = 0; k < m; k++) out[k] = cvt<R>(v[k]);
        a.template puts<R, m>(base, out, zbits<A>(o.w[4], 8, 4));
    }
}

// a current, form or constant value as the f64 a symjit kernel takes (df64: hi + lo in f64)
PCD double as_f64(double x) { return x; }
PCD double as_f64(float x) { return (double)x; }
PCD double as_f64(D x) { return (double)x.hi + (double)x.lo; }

// A symjit row (op 13), as RustiCol's executors run it (runner/recurrence_runner.cu:631-684): the
// kernel's inputs gathered in f64 by the binding's codes, the generated kernel, then each output
// pair (re, im) times the row's exact factor (re fr - im fi, re fi + im fr, each operation rounded
// once) added to the destination component unless the row initialises it -- or stored as it is
// when the binding is direct and the factor is exactly (1, +0) -- and narrowed to the kernel's
// current type.
// (inlined, with the gather and scatter loops unrolled to the kernels' limits, so that the in/out
// arrays have constant indices and stay in registers)
template <class K, class A> __device__ __attribute__((always_inline)) inline void symjit_row(const A &a, const Instr &o,
                                                                                           int flags) {
    typedef typename K::real R;
    typedef typename K::den Den;
    typedef typename K::acc T;  // the scales and factor as the policy's acc type (R but in KComp)
    const uint32_t *b = a.bind() + o.w[6];
    const uint32_t kernel = cnst(b)[0], inputs = cnst(b)[1], pairs = cnst(b)[2], direct = cnst(b)[3];
    const uint32_t dst = o.w[3];
    const double fr = as_f64(ldc<T>(a.fre(), o.w[5])), fi = as_f64(ldc<T>(a.fim(), o.w[5]));
    double in[PACUDA_MAX_IN], out[PACUDA_MAX_OUT];
#pragma unroll
    for (uint32_t k = 0; k < PACUDA_MAX_IN; k++) {
        if (k >= inputs) {
            in[k] = 0.0;
            continue;
        }
        const uint32_t code = cnst(b)[4 + k], which = (code >> 3) & 1, imag = (code >> 4) & 1, index = code >> 5;
        const uint32_t kind = code & 7;
        double v;
        if (kind == SB_PARENT || kind == SB_DESTINATION) {
            const Cx<R> x = a.template cur<R>((kind == SB_PARENT ? (which ? o.w[2] : o.w[1]) : dst) + index);
            v = as_f64(imag ? x.im : x.re);
        } else if (kind == SB_MOMENTUM) {
            const uint32_t q = (which ? o.w[4] >> 16 : o.w[4] & 0xffff) * 4 + index;
            double raw[4];
            a.form(q / 4, raw);
            v = as_f64(from_raw<Den>(raw[q % 4]));
        } else if (kind == SB_SCALE) {
            v = as_f64(ldc<T>(a.sre(), index));
        } else {
            v = imag ? fi : fr;
        }
        in[k] = v;
    }
    const bool simd = PACUDA_SIMD_LANES > 0 &&
                      a.point() < a.points() / (PACUDA_SIMD_LANES > 0 ? PACUDA_SIMD_LANES : 1) * PACUDA_SIMD_LANES;
    pacuda_kernel(kernel, simd, in, out);
    const bool as_is = direct && __builtin_bit_cast(uint64_t, fr) == __builtin_bit_cast(uint64_t, 1.0) &&
                       __builtin_bit_cast(uint64_t, fi) == 0;
#pragma unroll
    for (uint32_t j = 0; j < PACUDA_MAX_OUT / 2; j++) {
        if (j >= pairs) break;
        const double re = out[2 * j], im = out[2 * j + 1];
        const uint32_t plane = dst + cnst(b)[4 + inputs + j];
        double sr = re, si = im;
        if (!as_is) {
            sr = re * fr - im * fi;
            si = re * fi + im * fr;
            if (!(flags & F_INIT)) {
                const Cx<R> d = a.template cur<R>(plane);
                sr = as_f64(d.re) + sr;
                si = as_f64(d.im) + si;
            }
        }
        a.template put<R>(plane, cx(narrow<R>(sr), narrow<R>(si)));
    }
}

// One lane runs `count` instructions for its point (runner/recurrence_runner.cu:787-838).
template <class R, class A> __device__ inline void source_row(const Instr &o, const A &a);

// SYMJIT: whether the program may hold symjit rows. The symjit row is an out-of-line call with
// stack arrays, so kernels instantiated with it allocate registers worse everywhere; programs
// without symjit rows run an instance without the branch. COOP: the program is one wavefront's
// share of a cooperative program, whose barrier rows end its phases (every wavefront of the
// workgroup runs the same number of them).
template <class K, bool SYMJIT = false, bool COOP = false, class A>
__device__ inline void interpret(const Instr *__restrict__ prog, uint32_t count, const A &a) {
    typedef typename K::real R;
    typedef typename K::amp Amp;
    typedef typename K::acc T;
    typedef Cx<R> C;
    Cx<T> acc[6];
    if (count == 0) return;
    // the next instruction is fetched while this one runs (its scalar load overlaps the row's
    // vector loads instead of preceding them)
    Instr next = cnst(prog)[0];
    for (uint32_t n = 0; n < count; n++) {
        const Instr o = next;
        next = cnst(prog)[n + 1 < count ? n + 1 : n];
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
        const int flags = K::forwarding && A::forwarding ? (int)(o.w[0] >> 16) : (int)(o.w[0] >> 16) & (F_INIT | F_FIRST | F_LAST);
        if (kind == K_CONTRIB) {
            switch (op) {
            case 6: contribution<K, 6>(a, o, flags, acc); break;
            case 4: contribution<K, 4>(a, o, flags, acc); break;
            case 5: contribution<K, 5>(a, o, flags, acc); break;
            case 2: contribution<K, 2>(a, o, flags, acc); break;
            default: contribution<K, 3>(a, o, flags, acc); break;
            }
        } else if (kind == K_PROP) {
            if (op == 9) propagator<K, 9>(a, o, flags, acc);
            else if (op == 7) propagator<K, 7>(a, o, flags, acc);
            else propagator<K, 8>(a, o, flags, acc);
        } else if (kind == K_CLOSURE) {
            // (the closure's products in T: the compensated policy's are df64)
            T rfr = ldc<T>(a.fre(), o.w[4]), rfi = ldc<T>(a.fim(), o.w[4]);
            Amp dr, di;
            if (flags & F_ZERO) {
                dr = Amp(0.0);
                di = Amp(0.0);
            } else {
                const Cx<Amp> d = a.template amp<Amp>(o.w[3]);
                dr = d.re;
                di = d.im;
            }
            const uint32_t zl = zbits<A>(o.w[7], 0, 8), zr = zbits<A>(o.w[7], 8, 8);
            if (o.w[6] == 4) {  // four-plane currents: whole-object accesses
                C l[4], r[4];
                a.template curs<R, 4>(o.w[1], l, zl);
                a.template curs<R, 4>(o.w[2], r, zr);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    Cx<T> v = closure_increment(cvt<T>(l[k]), cvt<T>(r[k]), ldc<T>(a.fre(), o.w[5] + k),
                                                ldc<T>(a.fim(), o.w[5] + k), rfr, rfi);
                    dr = dr + widen(v.re);
                    di = di + widen(v.im);
                }
            } else {
                for (uint32_t k = 0; k < o.w[6]; k++) {
                    Cx<T> v = closure_increment(cvt<T>(a.template cur<R>(o.w[1] + k, (zl >> (k & 7)) & (k < 8))),
                                                cvt<T>(a.template cur<R>(o.w[2] + k, (zr >> (k & 7)) & (k < 8))),
                                                ldc<T>(a.fre(), o.w[5] + k), ldc<T>(a.fim(), o.w[5] + k), rfr, rfi);
                    dr = dr + widen(v.re);
                    di = di + widen(v.im);
                }
            }
            a.template put_amp<Amp>(o.w[3], cx(dr, di));
        } else if (kind == K_IDENTITY) {
            // (in T; an input from the accumulators rounded to R first, as in the propagator)
            T fr = ldc<T>(a.fre(), o.w[3]), fi = ldc<T>(a.fim(), o.w[3]);
            const uint32_t count = o.w[2], zin = zbits<A>(o.w[4], 0, 16), zout = zbits<A>(o.w[4], 16, 16);
            if (count == 4) {  // four-plane currents: whole-object accesses
                C u[4], out[4];
                Cx<T> v[4];
                if (flags & F_FROMACC) {
#pragma unroll
                    for (int k = 0; k < 4; k++) u[k] = cvt<R>(acc[k]);
                } else {
                    a.template curs<R, 4>(o.w[1], u, zin);
                }
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = identity_value(fr, fi, cvt<T>(u[k]));
                if (flags & F_HOLD) {
#pragma unroll
                    for (int k = 0; k < 4; k++) acc[k] = v[k];
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++) out[k] = cvt<R>(v[k]);
                    a.template puts<R, 4>(o.w[1], out, zout);
                }
            } else if (flags & (F_FROMACC | F_HOLD)) {  // at most 6 components (the accumulators)
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    if ((uint32_t)k < count) {
                        C u;
                        if (flags & F_FROMACC) u = cvt<R>(acc[k]);
                        else u = a.template cur<R>(o.w[1] + k, (zin >> k) & 1);
                        const Cx<T> v = identity_value(fr, fi, cvt<T>(u));
                        if (flags & F_HOLD) acc[k] = v;
                        else a.template put<R>(o.w[1] + k, cvt<R>(v), (zout >> k) & 1);
                    }
                }
            } else {
                for (uint32_t k = 0; k < count; k++) {
                    const uint32_t zi = k < 16 ? (zin >> k) & 1 : 0, zo = k < 16 ? (zout >> k) & 1 : 0;
                    const Cx<T> v = identity_value(fr, fi, cvt<T>(a.template cur<R>(o.w[1] + k, zi)));
                    a.template put<R>(o.w[1] + k, cvt<R>(v), zo);
                }
            }
        } else if (kind == K_SYMJIT) {
            if constexpr (SYMJIT) symjit_row<K>(a, o, flags);
        } else if (kind == K_BARRIER) {
            if constexpr (COOP) {
                __syncthreads();
                if (o.w[1]) __builtin_amdgcn_s_setprio(2);
                else __builtin_amdgcn_s_setprio(0);
            }
        } else if (kind == K_FETCH) {
            if constexpr (COOP) a.fetch();
        } else if (kind == K_SOURCE) {  // a source row made again (programs of the f64 kernel), or a cooperative program's
            if constexpr (COOP || (__is_same(R, double) && __is_same(typename K::den, double))) source_row<R>(o, a);
        } else if (kind == K_CLEAR) {
            if (o.w[2] == 4) {
                const C z[4] = {cx(R(0.0), R(0.0)), cx(R(0.0), R(0.0)), cx(R(0.0), R(0.0)), cx(R(0.0), R(0.0))};
                a.template puts<R, 4>(o.w[1], z);
            } else {
                for (uint32_t k = 0; k < o.w[2]; k++) a.template put<R>(o.w[1] + k, cx(R(0.0), R(0.0)));
            }
        }
    }
}

// One source instruction, in f64 from an f64 form, scale and factor, then narrowed to the
// kernel's current layout (runner/recurrence_runner.cu:766-785).
template <class R, class A> __device__ inline void source_row(const Instr &o, const A &a) {
    double p[4];
    a.form(o.w[2], p);
    double fr = ldc<double>(a.fre(), o.w[3]), fi = ldc<double>(a.fim(), o.w[3]), mass = ldc<double>(a.sre(), o.w[4]);
    W v[4];
    const int c = source_values(o.w[5], mass, p, v);
    Cx<R> out[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {  // c <= 4: a constant bound keeps v in registers
        if (k < c) {
            W s = cx(fr * v[k].re - fi * v[k].im, fr * v[k].im + fi * v[k].re);
            out[k] = cx(narrow<R>(s.re), narrow<R>(s.im));
        }
    }
    const uint32_t z = zbits<A>(o.w[6], 0, 4);
    if (c == 4) {  // four-plane currents: a whole-object store
        a.template puts<R, 4>(o.w[1], out, z);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k < c) a.template put<R>(o.w[1] + k, out[k], (z >> k) & 1);
    }
}

// the leading source instructions
template <class R, class A>
__device__ inline void run_sources_point(const Instr *__restrict__ prog, uint32_t count, const A &a) {
    for (uint32_t n = 0; n < count; n++) source_row<R>(cnst(prog)[n], a);
}

}  // namespace pacuda
