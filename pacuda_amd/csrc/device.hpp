// SPDX-License-Identifier: 0BSD
//
// Device arithmetic of the recurrence runner for gfx950: df64 pairs, complex helpers, the
// precision policies, the row operations (ops 2-11) and the external wavefunctions (op 12).
//
// Every function keeps the reference's operation order (runner/recurrence_runner.cu:72-542) so
// the f64 kernel is bit-identical to RustiCol's arithmetic; the build passes -ffp-contract=off
// and the pragma below repeats it, so no product or sum is fused unless an fma is written.
// Each value of a row is a register of the lane that owns the point; nothing here crosses
// lanes (the fused path's arena keeps some currents in the wavefront's LDS region).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define PCD __device__ __forceinline__
// the row formulas and complex helpers also run on the host, over the zero lattice of the
// fused path's structural-zero pass
#define PCHD __host__ __device__ __forceinline__

namespace pacuda {

// ---- df64: (hi, lo) float pairs -------------------------------------------------------------
// Joldes, Muller, Popescu, "Tight and rigorous error bounds for basic building blocks of
// double-word arithmetic", ACM TOMS 44(2), 2017: AccurateDWPlusDW (Alg. 6), DWTimesDW3
// (Alg. 12), DWDivDW2 (Alg. 17) with DWTimesFP3 (Alg. 9); the reference's choice
// (runner/recurrence_runner.cu:72-126).
struct D {
    float hi, lo;
    D() = default;
    PCD D(float h, float l) : hi(h), lo(l) {}
    PCD explicit D(double x) : hi((float)x), lo(0.0f) {}  // exact constants only (0, +-1, 2)
};

PCD void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
PCD void fast_two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    e = b - (s - a);
}
PCD D operator+(D x, D y) {
    float sh, sl, th, tl, vh, vl;
    two_sum(x.hi, y.hi, sh, sl);
    two_sum(x.lo, y.lo, th, tl);
    fast_two_sum(sh, sl + th, vh, vl);
    D z;
    fast_two_sum(vh, tl + vl, z.hi, z.lo);
    return z;
}
PCD D operator-(D x) { return D(-x.hi, -x.lo); }
PCD D operator-(D x, D y) { return x + (-y); }
PCD D operator*(D x, D y) {
    float ch = x.hi * y.hi;
    float cl1 = __builtin_fmaf(x.hi, y.hi, -ch);
    float tl0 = x.lo * y.lo;
    float tl1 = __builtin_fmaf(x.hi, y.lo, tl0);
    float cl2 = __builtin_fmaf(x.lo, y.hi, tl1);
    D z;
    fast_two_sum(ch, cl1 + cl2, z.hi, z.lo);
    return z;
}
PCD D operator/(D x, D y) {
    float th = x.hi / y.hi;
    float rh = y.hi * th;
    float rl1 = __builtin_fmaf(y.hi, th, -rh);
    float rl3 = __builtin_fmaf(y.lo, th, rl1);
    float r_hi, r_lo;
    fast_two_sum(rh, rl3, r_hi, r_lo);
    float pi_h = x.hi - r_hi;
    float delta = pi_h + (x.lo - r_lo);
    D z;
    fast_two_sum(th, delta / y.hi, z.hi, z.lo);
    return z;
}

// ---- complex values ------------------------------------------------------------------------
template <class R> struct Cx { R re, im; };
template <class R> PCHD Cx<R> cx(R re, R im) { Cx<R> c; c.re = re; c.im = im; return c; }
template <class R> PCHD Cx<R> add(Cx<R> a, Cx<R> b) { return cx(a.re + b.re, a.im + b.im); }
template <class R> PCHD Cx<R> sub(Cx<R> a, Cx<R> b) { return cx(a.re - b.re, a.im - b.im); }
template <class R> PCHD Cx<R> neg(Cx<R> a) { return cx(-a.re, -a.im); }
template <class R> PCHD Cx<R> mul(Cx<R> a, Cx<R> b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
template <class R> PCHD Cx<R> mul_i(Cx<R> a) { return cx(-a.im, a.re); }
template <class R> PCHD Cx<R> mul_real(Cx<R> a, R s) { return cx(a.re * s, a.im * s); }

// ---- uniform tables ----------------------------------------------------------------------------
// A pointer into the constant address space: a load at a wavefront-uniform address becomes a
// scalar load (SGPRs, scalar cache). Only for tables no kernel writes (instructions, plans,
// scales, factors, reduction tables).
// (The host pass of the HIP compiler sees the plain pointer: these functions only run on the device.)
#if defined(__HIP_DEVICE_COMPILE__)
#define PACUDA_CONST __attribute__((address_space(4)))
#else
#define PACUDA_CONST
#endif
template <class T> PCD const PACUDA_CONST T *cnst(const T *p) { return (const PACUDA_CONST T *)p; }

// a uniform constant (scale or factor) as R: doubles, or df64/f32 values in a double's 8 bytes
template <class R> PCD R ldc(const double *p, size_t at);
template <> PCD double ldc<double>(const double *p, size_t at) { return cnst(p)[at]; }
template <> PCD D ldc<D>(const double *p, size_t at) {
    float2 v = cnst(reinterpret_cast<const float2 *>(p))[at];
    return D(v.x, v.y);
}
template <> PCD float ldc<float>(const double *p, size_t at) { return cnst(reinterpret_cast<const float2 *>(p))[at].x; }

// ---- arena values: doubles, or df64/f32 values in the 8 bytes of a double --------------------
template <class R> PCD R ldr(const double *p, size_t at);
template <> PCD double ldr<double>(const double *p, size_t at) { return p[at]; }
template <> PCD D ldr<D>(const double *p, size_t at) {
    float2 v = reinterpret_cast<const float2 *>(p)[at];
    return D(v.x, v.y);
}
template <> PCD float ldr<float>(const double *p, size_t at) { return reinterpret_cast<const float2 *>(p)[at].x; }
PCD void str(double *p, size_t at, double v) { p[at] = v; }
PCD void str(double *p, size_t at, D v) { reinterpret_cast<float2 *>(p)[at] = make_float2(v.hi, v.lo); }
PCD void str(double *p, size_t at, float v) { reinterpret_cast<float2 *>(p)[at] = make_float2(v, 0.0f); }

// ---- raw 8-byte values: a double, or a df64 pair / f32 value in a double's 8 bytes -----------
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float flt2 __attribute__((ext_vector_type(2)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <class R> PCD R from_raw(double raw);
template <> PCD double from_raw<double>(double raw) { return raw; }
template <> PCD D from_raw<D>(double raw) {
    const flt2 v = __builtin_bit_cast(flt2, raw);
    return D(v.x, v.y);
}
template <> PCD float from_raw<float>(double raw) { return __builtin_bit_cast(flt2, raw).x; }
PCD double to_raw(double v) { return v; }
PCD double to_raw(D v) { return __builtin_bit_cast(double, flt2{v.hi, v.lo}); }
PCD double to_raw(float v) { return __builtin_bit_cast(double, flt2{v, 0.0f}); }

// Precision policies: R the values, Den the propagator denominators and their forms, Amp the
// amplitude accumulation (runner/recurrence_runner.cu:178-190), Acc the vertex arithmetic and
// the accumulation of a destination's contributions (R in the reference's four policies).
// forwarding: whether the kernel honours F_HOLD / F_FROMACC (every kernel does; a kernel without
// it would run the plan's program without forwarding and LDS residency)
struct KF64 {
    typedef double real; typedef double den; typedef double amp; typedef double acc;
    static constexpr bool forwarding = true;
};
struct KDF64 {
    typedef D real; typedef D den; typedef D amp; typedef D acc;
    static constexpr bool forwarding = true;
};
struct KMixed {
    typedef float real; typedef D den; typedef D amp; typedef float acc;
    static constexpr bool forwarding = true;
};
struct KMixed64 {
    typedef float real; typedef double den; typedef D amp; typedef float acc;
    static constexpr bool forwarding = true;
};
// The compensated mixed policy (no counterpart in the reference): f32 currents and df64
// denominators and amplitude sums as in KMixed, and the arithmetic of every row -- the vertices
// (ops 2-6) and the sum of a destination's contributions, propagators, identity rows and the
// closures' products -- in df64 from the f32 currents, the df64 forms and the df64 constants, so
// that a current is rounded to f32 once, when it is stored (a forwarded one is rounded as if it
// had been stored). Symjit rows run in f64 as in every policy.
struct KComp {
    typedef float real; typedef D den; typedef D amp; typedef D acc;
    static constexpr bool forwarding = true;
};
// The compensated policy with f64 arithmetic (no counterpart in the reference): KComp's f32
// currents and single rounding per stored current, with every row's arithmetic, the forms, the
// constants, the denominators and the amplitude sums in f64 -- native on this GPU (f64 vector
// operations at half the f32 rate), where df64 takes about ten f32 operations per operation.
// It equals KComp to df64's accuracy.
struct KComp64 {
    typedef float real; typedef double den; typedef double amp; typedef double acc;
    static constexpr bool forwarding = true;
};
template <class K> constexpr bool compensated() { return !__is_same(typename K::real, typename K::acc); }

// conversions between a policy's R and Acc: exact widening, and the rounding of a df64 value
// to the nearest f32 (fl(hi + lo))
template <class T, class S> struct Conv { static PCD T of(S x) { return x; } };
template <> struct Conv<D, float> { static PCD D of(float x) { return D(x, 0.0f); } };
template <> struct Conv<float, D> { static PCD float of(D x) { return x.hi + x.lo; } };
template <class T, class S> PCD Cx<T> cvt(Cx<S> c) { return cx(Conv<T, S>::of(c.re), Conv<T, S>::of(c.im)); }

PCD float to_float(float x) { return x; }
PCD float to_float(double x) { return (float)x; }
PCD float to_float(D x) { return x.hi + x.lo; }

// a form component as the kernel's R, from its raw value (f64, or df64 after the df64 kernels'
// in-place conversion)
template <class R, class Den> struct Form { static PCD R of(double raw) { return from_raw<R>(raw); } };
template <> struct Form<float, double> { static PCD float of(double raw) { return (float)raw; } };
template <> struct Form<float, D> { static PCD float of(double raw) { return from_raw<float>(raw); } };

template <class R, class S> struct ToReal { static PCD R of(S x) { return x; } };
template <class S> struct ToReal<float, S> { static PCD float of(S x) { return to_float(x); } };

PCD D widen(float x) { return D(x, 0.0f); }
PCD double widen(double x) { return x; }
PCD D widen(D x) { return x; }

template <class R> PCD R narrow(double x);
template <> PCD double narrow<double>(double x) { return x; }
template <> PCD float narrow<float>(double x) { return (float)x; }
template <> PCD D narrow<D>(double x) {
    float hi = (float)x;
    return D(hi, (float)(x - (double)hi));
}

// ---- arenas ------------------------------------------------------------------------------------
// The interpreter reaches its values through an arena: cur / curs / put / puts (current
// components by plane), form (the four raw components of a momentum form), amp / put_amp
// (amplitude planes), and the constant tables sre, sim, fre, fim.

// The v1 path's arena: the reference's layout (runner/recurrence_runner.cu:850-876, 1083-1104),
// value (plane q, point i) at q * ld + i of each array, ld = the caller's stride; the mixed
// kernels' currents in separate f32 planes.
struct Arena {
    double *cre, *cim, *are, *aim;
    float *fcre, *fcim;  // mixed kernels: the currents as f32 planes
    const double *mom;
    const double *sre, *sim, *fre, *fim;
    const uint32_t *bind;  // the program's symjit bindings (none: nullptr)
    size_t ld;
    uint32_t points;  // the call's points
};

template <class R> struct Cur {
    static PCD Cx<R> ld(const Arena &a, size_t at) { return cx(ldr<R>(a.cre, at), ldr<R>(a.cim, at)); }
    static PCD void st(const Arena &a, size_t at, Cx<R> v) { str(a.cre, at, v.re); str(a.cim, at, v.im); }
};
template <> struct Cur<float> {
    static PCD Cx<float> ld(const Arena &a, size_t at) { return cx(a.fcre[at], a.fcim[at]); }
    static PCD void st(const Arena &a, size_t at, Cx<float> v) { a.fcre[at] = v.re; a.fcim[at] = v.im; }
};

// one lane's view of the v1 arena: point i (v1 programs carry no forwarding flags)
struct PlaneArena {
    static constexpr bool forwarding = false;
    static constexpr bool zero_masks = false;  // v1 programs carry no zero masks
    Arena a;
    size_t i;
    template <class R> PCD Cx<R> cur(uint32_t p, uint32_t = 0) const { return Cur<R>::ld(a, (size_t)p * a.ld + i); }
    template <class R> PCD void put(uint32_t p, Cx<R> v, uint32_t = 0) const { Cur<R>::st(a, (size_t)p * a.ld + i, v); }
    template <class R, int N> PCD void curs(uint32_t p, Cx<R> *v, uint32_t = 0) const {
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = cur<R>(p + k);
    }
    template <class R, int N> PCD void puts(uint32_t p, const Cx<R> *v, uint32_t = 0) const {
#pragma unroll
        for (int k = 0; k < N; k++) put<R>(p + k, v[k]);
    }
    PCD void form(uint32_t f, double *raw) const {
#pragma unroll
        for (int k = 0; k < 4; k++) raw[k] = a.mom[((size_t)f * 4 + k) * a.ld + i];
    }
    template <class T> PCD Cx<T> amp(uint32_t q) const {
        return cx(ldr<T>(a.are, (size_t)q * a.ld + i), ldr<T>(a.aim, (size_t)q * a.ld + i));
    }
    template <class T> PCD void put_amp(uint32_t q, Cx<T> v) const {
        str(a.are, (size_t)q * a.ld + i, v.re);
        str(a.aim, (size_t)q * a.ld + i, v.im);
    }
    PCD const double *sre() const { return a.sre; }
    PCD const double *sim() const { return a.sim; }
    PCD const double *fre() const { return a.fre; }
    PCD const double *fim() const { return a.fim; }
    PCD const uint32_t *bind() const { return a.bind; }
    PCD size_t point() const { return i; }
    PCD size_t points() const { return a.points; }
};

// A complex current component as one vector: (re, im) doubles, (re.hi, re.lo, im.hi, im.lo) for
// df64, (re, im) floats for the mixed kernels -- 16, 16 and 8 bytes per lane.
template <class R> struct Packed;
template <> struct Packed<double> {
    typedef dbl2 T;
    static PCD Cx<double> get(T v) { return cx(v.x, v.y); }
    static PCD T set(Cx<double> c) { return T{c.re, c.im}; }
};
template <> struct Packed<D> {
    typedef flt4 T;
    static PCD Cx<D> get(T v) { return cx(D(v.x, v.y), D(v.z, v.w)); }
    static PCD T set(Cx<D> c) { return T{c.re.hi, c.re.lo, c.im.hi, c.im.lo}; }
};
template <> struct Packed<float> {
    typedef flt2 T;
    static PCD Cx<float> get(T v) { return cx(v.x, v.y); }
    static PCD T set(Cx<float> c) { return T{c.re, c.im}; }
};

// Current planes at or above LDS_PLANE are on chip, at slot s = plane - LDS_PLANE: in the
// wavefront's LDS region for s below the launch's LDS capacity, else (s - capacity) in the lane's
// register tier; only the fused path's passes write such indices.
constexpr uint32_t LDS_PLANE = 1u << 24;
// A slot plane p with an on-chip copy at slot s: flags | ((s - p) & 63) << 24 | p, so that the
// field plus c names component c in both places (s + c < 64). DUAL_FILL: a read loads from the
// slot and keeps a copy on chip, a write goes on chip; DUAL_THROUGH: a read reads on chip (unless
// DUAL_FILL), a write goes on chip and to the slot.
constexpr uint32_t DUAL_FILL = 1u << 31, DUAL_THROUGH = 1u << 30, DUAL = DUAL_FILL | DUAL_THROUGH;
PCHD uint32_t dual_chip(uint32_t f) { return ((f & (LDS_PLANE - 1)) + (f >> 24)) & 63; }
PCHD uint32_t dual_field(uint32_t flags, uint32_t p, uint32_t s) { return flags | ((s - p) & 63) << 24 | p; }
// the register tier: REG_PLANES planes in units of one four-plane current
constexpr uint32_t REG_PLANES = 16;
// ... and up to FORM_REGS momentum forms: a form field of an instruction with bits 24 and up set
// names the register slot ((field >> 24) - 1) holding form (field & 0xffffff), when the kernel
// keeps that many forms (else the form is read from the slot)
constexpr uint32_t FORM_REGS = 6;
constexpr uint32_t FORM_SLOT_SHIFT = 24;

// Raw buffer accesses: a wavefront-uniform resource (base, size), a per-lane byte offset
// (VGPR) and a uniform byte offset (SGPR), so a lane's access needs no address arithmetic of
// its own. Word 3 of the resource is the gfx9-family raw-buffer format (32-bit data format).
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int BUFFER_FORMAT = 0x00020000;
template <int BYTES> struct Buffer;
template <> struct Buffer<16> {
    template <class T> static PCD T load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
    }
    template <class T> static PCD void store(T x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, x), r, v, s, 0);
    }
};
template <> struct Buffer<8> {
    template <class T> static PCD T load(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
    }
    template <class T> static PCD void store(T x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, v, s, 0);
    }
};

// The register tier of one lane: REG_PLANES / 4 four-plane currents, selected by a uniform
// switch (the unit index is a scalar) so every element keeps a fixed register.
template <class R> struct RegFile {
    Cx<R> u0[4], u1[4], u2[4], u3[4];
    PCD void get(uint32_t unit, Cx<R> *v) const {
        switch (unit) {
        case 0: for (int k = 0; k < 4; k++) v[k] = u0[k]; break;
        case 1: for (int k = 0; k < 4; k++) v[k] = u1[k]; break;
        case 2: for (int k = 0; k < 4; k++) v[k] = u2[k]; break;
        default: for (int k = 0; k < 4; k++) v[k] = u3[k]; break;
        }
    }
    PCD void set(uint32_t unit, const Cx<R> *v) {
        switch (unit) {
        case 0: for (int k = 0; k < 4; k++) u0[k] = v[k]; break;
        case 1: for (int k = 0; k < 4; k++) u1[k] = v[k]; break;
        case 2: for (int k = 0; k < 4; k++) u2[k] = v[k]; break;
        default: for (int k = 0; k < 4; k++) u3[k] = v[k]; break;
        }
    }
};
static_assert(REG_PLANES == 16, "RegFile holds four units");

// The register-resident forms of one lane (four raw doubles each), selected like RegFile.
// N register-resident forms of one lane, selected by a uniform slot (unrolled compares: a
// runtime index into the array would put it in scratch memory)
template <int N> struct FormRegs {
    double f[N][4];
    PCD void get(uint32_t s, double *raw) const {
#pragma unroll
        for (int q = 0; q < N; q++)
            if (s == (uint32_t)q) {
#pragma unroll
                for (int k = 0; k < 4; k++) raw[k] = f[q][k];
            }
    }
    PCD void set(uint32_t s, const double *raw) {
#pragma unroll
        for (int q = 0; q < N; q++)
            if (s == (uint32_t)q) {
#pragma unroll
                for (int k = 0; k < 4; k++) f[q][k] = raw[k];
            }
    }
};

// The fused path's arena: one wavefront's scratch slot for 64 points, each plane 64 lanes of one
// vector: currents as Packed<R> (PLANE bytes per plane), then amplitudes as (re, im) raw pairs
// (1 KB per plane), then forms as four raw doubles (2 KB per form); LDS-resident currents as
// Packed<R> in the wavefront's LDS region. A lane's access is one buffer load or store at
// (lane offset, uniform plane offset). Every operand of a row is one whole object of the LDS
// pass, so one residency test serves all its components.
// REG: the lane has a register tier; an on-chip slot at or above lds_cap is in it (only whole
// four-plane currents go there, so only the four-plane accessors look). NF: the lane also holds
// the first NF register-resident forms (a form field marked with a later slot, or any marked
// field when NF is 0, is read from the slot's copy). COOP: the arena of a cooperative kernel,
// whose slot and LDS region belong to the workgroup (one unit, shared by its wavefronts): a form
// field with FORM_LDS set names the LDS planes from (field & (FORM_LDS - 1)) on that hold its 2 KB
// (a lane's components 0-1 at 16 bytes per lane, then 2-3 1 KB further), an amplitude field at or
// above LDS_PLANE the planes from (field - LDS_PLANE) on that hold its 1 KB (16 bytes per lane).
constexpr uint32_t FORM_LDS = 1u << 23;
template <class R, bool REG = false, int NF = 0, bool COOP = false> struct SlotArena {
    typedef typename Packed<R>::T V;
    typedef Buffer<sizeof(V)> B;
    static constexpr bool forwarding = true;
    static constexpr bool zero_masks = true;  // a masked component is neither loaded nor stored in the slot
    static constexpr uint32_t PLANE = 64 * sizeof(V);  // bytes per current plane
    __amdgpu_buffer_rsrc_t slot;
    uint32_t amp_base, form_base;  // byte offsets of the amplitude and form regions in the slot
    const double *sre_, *sim_, *fre_, *fim_;
    uint32_t idx;    // the lane (0-63)
    uint32_t lds;    // byte address of the lane's value in LDS plane 0 of the wavefront's region
    uint32_t lds_cap;     // on-chip slots below this are LDS planes
    RegFile<R> *regs;     // the register tier (REG only)
    FormRegs<NF ? NF : 1> *fregs;  // the register-resident forms (NF > 0 only)

    static PCD V *lds_at(uint32_t off) {
        extern __shared__ __attribute__((aligned(16))) char pacuda_lds[];
        return reinterpret_cast<V *>(pacuda_lds + off);
    }
    // z = 1: a known zero, so the access is put out of the slot's range (the load returns +0 and
    // the store is dropped, with no memory traffic)
    static constexpr uint32_t OUT_OF_RANGE = 0x80000000u;
    PCD V mem_ld(uint32_t p, uint32_t z = 0) const {
        return B::template load<V>(slot, idx * (uint32_t)sizeof(V) + z * OUT_OF_RANGE, p * PLANE);
    }
    PCD void mem_st(uint32_t p, V v, uint32_t z = 0) const {
        B::store(v, slot, idx * (uint32_t)sizeof(V) + z * OUT_OF_RANGE, p * PLANE);
    }

    // single planes: the slot or LDS (the register tier holds whole four-plane currents only)
    template <class Q> PCD Cx<R> cur(uint32_t p, uint32_t z = 0) const {
        static_assert(__is_same(Q, R), "one value type per slot");
        if (p >= LDS_PLANE) {
            if (!COOP && (p & DUAL)) {  // (cooperative programs have no dual fields)
                V *c = lds_at(lds + dual_chip(p) * PLANE);
                if (!(p & DUAL_FILL)) return Packed<R>::get(*c);
                const V x = mem_ld(p & (LDS_PLANE - 1), z);
                *c = x;
                return Packed<R>::get(x);
            }
            return Packed<R>::get(*lds_at(lds + (p - LDS_PLANE) * PLANE));
        }
        return Packed<R>::get(mem_ld(p, z));
    }
    template <class Q> PCD void put(uint32_t p, Cx<R> v, uint32_t z = 0) const {
        static_assert(__is_same(Q, R), "one value type per slot");
        if (p >= LDS_PLANE) {
            if (!COOP && (p & DUAL)) {  // (cooperative programs have no dual fields)
                *lds_at(lds + dual_chip(p) * PLANE) = Packed<R>::set(v);
                if (p & DUAL_THROUGH) mem_st(p & (LDS_PLANE - 1), Packed<R>::set(v), z);
                return;
            }
            *lds_at(lds + (p - LDS_PLANE) * PLANE) = Packed<R>::set(v);
        } else {
            mem_st(p, Packed<R>::set(v), z);
        }
    }
    // whole objects on chip at slot s: the register tier from lds_cap up, else LDS
    template <int N> PCD void chip_get(uint32_t s, Cx<R> *v) const {
        if constexpr (REG && N == 4) {
            if (s >= lds_cap) {
                regs->get((s - lds_cap) / 4, v);
                return;
            }
        }
        const uint32_t o = lds + s * PLANE;
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = Packed<R>::get(*lds_at(o + k * PLANE));
    }
    template <int N> PCD void chip_set(uint32_t s, const Cx<R> *v) const {
        if constexpr (REG && N == 4) {
            if (s >= lds_cap) {
                regs->set((s - lds_cap) / 4, v);
                return;
            }
        }
        const uint32_t o = lds + s * PLANE;
#pragma unroll
        for (int k = 0; k < N; k++) *lds_at(o + k * PLANE) = Packed<R>::set(v[k]);
    }
    // zmask bit k: component k is a known zero (the LDS and register tiers hold it anyway)
    template <class Q, int N> PCD void curs(uint32_t p, Cx<R> *v, uint32_t zmask = 0) const {
        static_assert(__is_same(Q, R), "one value type per slot");
        if (p >= LDS_PLANE) {
            if (!COOP && (p & DUAL)) {  // (cooperative programs have no dual fields)
                if (p & DUAL_FILL) {  // (copies go to LDS: the register tier takes objects on writes)
                    const uint32_t o = lds + dual_chip(p) * PLANE;
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        const V x = mem_ld((p & (LDS_PLANE - 1)) + k, (zmask >> k) & 1);
                        *lds_at(o + k * PLANE) = x;
                        v[k] = Packed<R>::get(x);
                    }
                } else {
                    chip_get<N>(dual_chip(p), v);
                }
                return;
            }
            chip_get<N>(p - LDS_PLANE, v);
        } else {
#pragma unroll
            for (int k = 0; k < N; k++) v[k] = Packed<R>::get(mem_ld(p + k, (zmask >> k) & 1));
        }
    }
    template <class Q, int N> PCD void puts(uint32_t p, const Cx<R> *v, uint32_t zmask = 0) const {
        static_assert(__is_same(Q, R), "one value type per slot");
        if (p >= LDS_PLANE) {
            if (!COOP && (p & DUAL)) {  // (cooperative programs have no dual fields)
                chip_set<N>(dual_chip(p), v);
                if (p & DUAL_THROUGH) {
#pragma unroll
                    for (int k = 0; k < N; k++)
                        mem_st((p & (LDS_PLANE - 1)) + k, Packed<R>::set(v[k]), (zmask >> k) & 1);
                }
                return;
            }
            chip_set<N>(p - LDS_PLANE, v);
        } else {
#pragma unroll
            for (int k = 0; k < N; k++) mem_st(p + k, Packed<R>::set(v[k]), (zmask >> k) & 1);
        }
    }
    // form field f: a register slot when bits 24 and up are set (REG kernels), else a slot form
    PCD void form(uint32_t f, double *raw) const {
        if constexpr (COOP) {
            if (f & FORM_LDS) {
                const uint32_t o = (f & (FORM_LDS - 1)) * PLANE + idx * 16u;
                const dbl2 x = *reinterpret_cast<const dbl2 *>(lds_at(o));
                const dbl2 y = *reinterpret_cast<const dbl2 *>(lds_at(o + 1024u));
                raw[0] = x.x; raw[1] = x.y; raw[2] = y.x; raw[3] = y.y;
                return;
            }
        }
        if constexpr (NF > 0) {
            if ((f >> FORM_SLOT_SHIFT) && (f >> FORM_SLOT_SHIFT) <= (uint32_t)NF) {
                fregs->get((f >> FORM_SLOT_SHIFT) - 1, raw);
                return;
            }
        }
        f &= (1u << FORM_SLOT_SHIFT) - 1;
        const dbl2 x = Buffer<16>::load<dbl2>(slot, idx * 32u, form_base + f * 2048u);
        const dbl2 y = Buffer<16>::load<dbl2>(slot, idx * 32u + 16u, form_base + f * 2048u);
        raw[0] = x.x; raw[1] = x.y; raw[2] = y.x; raw[3] = y.y;
    }
    PCD void put_form(uint32_t f, const double *raw) const {
        if constexpr (COOP) {
            if (f & FORM_LDS) {
                const uint32_t o = (f & (FORM_LDS - 1)) * PLANE + idx * 16u;
                *reinterpret_cast<dbl2 *>(lds_at(o)) = dbl2{raw[0], raw[1]};
                *reinterpret_cast<dbl2 *>(lds_at(o + 1024u)) = dbl2{raw[2], raw[3]};
                return;
            }
        }
        Buffer<16>::store(dbl2{raw[0], raw[1]}, slot, idx * 32u, form_base + f * 2048u);
        Buffer<16>::store(dbl2{raw[2], raw[3]}, slot, idx * 32u + 16u, form_base + f * 2048u);
    }
    template <class T> PCD Cx<T> amp(uint32_t q) const {
        if constexpr (COOP) {
            if (q >= LDS_PLANE) {
                const dbl2 v = *reinterpret_cast<const dbl2 *>(lds_at((q - LDS_PLANE) * PLANE + idx * 16u));
                return cx(from_raw<T>(v.x), from_raw<T>(v.y));
            }
        }
        const dbl2 v = Buffer<16>::load<dbl2>(slot, idx * 16u, amp_base + q * 1024u);
        return cx(from_raw<T>(v.x), from_raw<T>(v.y));
    }
    template <class T> PCD void put_amp(uint32_t q, Cx<T> v) const {
        if constexpr (COOP) {
            if (q >= LDS_PLANE) {
                *reinterpret_cast<dbl2 *>(lds_at((q - LDS_PLANE) * PLANE + idx * 16u)) = dbl2{to_raw(v.re), to_raw(v.im)};
                return;
            }
        }
        Buffer<16>::store(dbl2{to_raw(v.re), to_raw(v.im)}, slot, idx * 16u, amp_base + q * 1024u);
    }
    PCD const double *sre() const { return sre_; }
    PCD const double *sim() const { return sim_; }
    PCD const double *fre() const { return fre_; }
    PCD const double *fim() const { return fim_; }
    const uint32_t *bind_;  // the plan's symjit bindings
    uint32_t point_, points_;  // the lane's point in its item, the item's points
    // (COOP) the work counter and the LDS word that receives the workgroup's next unit
    uint32_t *counter;
    volatile uint32_t *next_unit;
    PCD void fetch() const {
        if (idx == 0) *next_unit = atomicAdd(counter, 1u);
    }
    PCD const uint32_t *bind() const { return bind_; }
    PCD uint32_t point() const { return point_; }
    PCD uint32_t points() const { return points_; }
};

template <class R> PCD R reciprocal_minkowski_norm(R p0, R p1, R p2, R p3) {
    return R(1.0) / (p0 * p0 - p1 * p1 - p2 * p2 - p3 * p3);
}
template <class R> PCHD Cx<R> minkowski_dot(const Cx<R> *l, const Cx<R> *r) {
    return sub(sub(sub(mul(l[0], r[0]), mul(l[1], r[1])), mul(l[2], r[2])), mul(l[3], r[3]));
}
template <class R> PCHD Cx<R> minkowski_real_dot(const Cx<R> *l, const R *q) {
    return sub(sub(sub(mul_real(l[0], q[0]), mul_real(l[1], q[1])), mul_real(l[2], q[2])), mul_real(l[3], q[3]));
}

// components read from parent 0 / parent 1 and written, per contribution op
template <int OP> struct Shape {
    static constexpr int c0 = OP == 4 ? 6 : (OP == 5 || OP == 6) ? 4 : 2;
    static constexpr int c1 = 4;
    static constexpr int cd = (OP == 4 || OP == 6) ? 4 : OP == 5 ? 6 : 2;
};

// ops 2-6 before the scale (convert/CONVENTIONS.md:71-79 in runner/recurrence_runner.cu:244-282's order)
template <int OP, class R>
PCHD void contribution_values(const Cx<R> *l, const Cx<R> *r, const R *q0, const R *q1, Cx<R> *o) {
    if constexpr (OP == 2) {
        o[0] = add(add(sub(neg(mul_i(mul(l[0], r[3]))), mul_i(mul(l[1], r[1]))), mul(l[1], r[2])), mul_i(mul(l[0], r[0])));
        o[1] = add(add(sub(neg(mul(l[0], r[2])), mul_i(mul(l[0], r[1]))), mul_i(mul(l[1], r[0]))), mul_i(mul(l[1], r[3])));
    } else if constexpr (OP == 3) {
        o[0] = add(add(add(neg(mul(l[1], r[2])), mul_i(mul(l[0], r[0]))), mul_i(mul(l[0], r[3]))), mul_i(mul(l[1], r[1])));
        o[1] = add(add(add(neg(mul_i(mul(l[1], r[3]))), mul(l[0], r[2])), mul_i(mul(l[0], r[1]))), mul_i(mul(l[1], r[0])));
    } else if constexpr (OP == 4) {
        o[0] = add(add(mul(l[0], r[1]), mul(l[1], r[2])), mul(l[2], r[3]));
        o[1] = add(add(mul(l[0], r[0]), mul(l[3], r[2])), mul(l[4], r[3]));
        o[2] = add(sub(mul(l[1], r[0]), mul(l[3], r[1])), mul(l[5], r[3]));
        o[3] = sub(sub(mul(l[2], r[0]), mul(l[4], r[1])), mul(l[5], r[2]));
    } else if constexpr (OP == 5) {
        o[0] = sub(mul(l[0], r[1]), mul(l[1], r[0]));
        o[1] = sub(mul(l[0], r[2]), mul(l[2], r[0]));
        o[2] = sub(mul(l[0], r[3]), mul(l[3], r[0]));
        o[3] = sub(mul(l[1], r[2]), mul(l[2], r[1]));
        o[4] = sub(mul(l[1], r[3]), mul(l[3], r[1]));
        o[5] = sub(mul(l[2], r[3]), mul(l[3], r[2]));
    } else {
        Cx<R> vd = minkowski_dot(l, r);
        Cx<R> ld1 = minkowski_real_dot(l, q1);
        Cx<R> rd0 = minkowski_real_dot(r, q0);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            Cx<R> mt = mul_real(vd, q0[k] - q1[k]);
            Cx<R> ct = mul_real(sub(mul(ld1, r[k]), mul(rd0, l[k])), R(2.0));
            o[k] = add(mt, ct);
        }
    }
}

// ops 7-9 (runner/recurrence_runner.cu:284-314)
template <int OP, class R>
PCHD void propagator_values(const Cx<R> *c, const R *p, R reciprocal, Cx<R> row_factor, Cx<R> *o) {
    if constexpr (OP == 9) {
        Cx<R> scale = mul_real(mul(row_factor, cx(R(0.0), R(-1.0))), reciprocal);
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = mul(c[k], scale);
    } else {
        Cx<R> scale = mul_real(row_factor, reciprocal);
        Cx<R> o0, o1;
        if constexpr (OP == 7) {
            o0 = add(add(sub(neg(mul_i(mul_real(c[0], p[3]))), mul_i(mul_real(c[1], p[1]))), mul_real(c[1], p[2])),
                     mul_i(mul_real(c[0], p[0])));
            o1 = add(add(sub(neg(mul_real(c[0], p[2])), mul_i(mul_real(c[0], p[1]))), mul_i(mul_real(c[1], p[0]))),
                     mul_i(mul_real(c[1], p[3])));
        } else {
            o0 = add(add(add(neg(mul_real(c[1], p[2])), mul_i(mul_real(c[0], p[0]))), mul_i(mul_real(c[0], p[3]))),
                     mul_i(mul_real(c[1], p[1])));
            o1 = add(add(add(neg(mul_i(mul_real(c[1], p[3]))), mul_real(c[0], p[2])), mul_i(mul_real(c[0], p[1]))),
                     mul_i(mul_real(c[1], p[0])));
        }
        o[0] = mul(o0, scale);
        o[1] = mul(o1, scale);
    }
}

template <class R> PCHD Cx<R> identity_value(R fr, R fi, Cx<R> v) { return cx(fr * v.re - fi * v.im, fr * v.im + fi * v.re); }

template <class R> PCHD Cx<R> closure_increment(Cx<R> l, Cx<R> r, R cr, R ci, R rfr, R rfi) {
    R scale_re = cr * rfr - ci * rfi;
    R scale_im = cr * rfi + ci * rfr;
    R product_re = l.re * r.re - l.im * r.im;
    R product_im = l.re * r.im + l.im * r.re;
    return cx(product_re * scale_re - product_im * scale_im, product_re * scale_im + product_im * scale_re);
}

// ---- external wavefunctions in f64 (op 12; runner/recurrence_runner.cu:332-542, legend
// convert/CONVENTIONS.md:52-61). An incoming leg has a negative energy. ----------------------
typedef Cx<double> W;
PCD W c64(double re, double im) { return cx(re, im); }
PCD W scale_w(W c, double s) { return cx(c.re * s, c.im * s); }
PCD double fortran_sign(double v, double s) { return __builtin_copysign(__builtin_fabs(v), s); }
// max/min returning the other operand when one is NaN; on equal operands the second
PCD double rmax(double a, double b) { return a != a ? b : b != b ? a : a > b ? a : b; }
PCD double rmin(double a, double b) { return a != a ? b : b != b ? a : a < b ? a : b; }

// the shared (a, b) pair of the massless Weyl/Dirac spinors; `anti` the antiparticle signs
PCD void weyl_pair(const double *m, int helicity, bool anti, W &chi1, W &chi2) {
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    double h = (double)helicity;
    if (energy > 0.0) {
        double s = (px == 0.0 && py == 0.0 && pz < 0.0) ? 0.0 : sqrt(rmax(energy + pz, 0.0));
        if (anti && !(px == 0.0 && py == 0.0 && pz < 0.0)) s = -s;
        chi1 = c64(s, 0.0);
        if (s == 0.0) chi2 = c64(-h * sqrt(2.0 * energy), 0.0);
        else chi2 = anti ? c64(-h * px / s, py / s) : c64(h * px / s, -py / s);
        return;
    }
    double s = (px == 0.0 && py == 0.0 && pz > 0.0) ? 0.0 : sqrt(rmax(-(energy + pz), 0.0));
    if (!anti && !(px == 0.0 && py == 0.0 && pz > 0.0)) s = -s;
    chi1 = c64(s, 0.0);
    if (s == 0.0) chi2 = c64(-h * sqrt(2.0 * fabs(energy)), 0.0);
    else chi2 = anti ? c64(h * (-px) / s, -py / s) : c64(-h * (-px) / s, py / s);
}

PCD void ext_weyl(const double *m, int helicity, int chirality, bool anti, W *o) {
    W a, b;
    weyl_pair(m, helicity, anti, a, b);
    W z = c64(0.0, 0.0);
    bool incoming = !(m[0] > 0.0);
    int hp = incoming ? -1 : 1;  // the helicity of (a, b) order for a particle with chirality +1
    if (helicity == hp && chirality == 1) { if (anti) { o[0] = b; o[1] = a; } else { o[0] = a; o[1] = b; } return; }
    if (helicity == -hp && chirality == -1) { if (anti) { o[0] = a; o[1] = b; } else { o[0] = b; o[1] = a; } return; }
    o[0] = z; o[1] = z;
}

PCD void ext_dirac_massless(const double *m, int helicity, bool anti, W *o) {
    W a, b;
    weyl_pair(m, helicity, anti, a, b);
    W z = c64(0.0, 0.0);
    bool incoming = !(m[0] > 0.0);
    if (!anti) {
        bool upper = incoming ? (-helicity == 1) : (helicity == 1);
        if (upper) { o[0] = a; o[1] = b; o[2] = z; o[3] = z; }
        else { o[0] = z; o[1] = z; o[2] = b; o[3] = a; }
    } else {
        bool lower = incoming ? (helicity == 1) : (-helicity == 1);
        if (lower) { o[0] = z; o[1] = z; o[2] = a; o[3] = b; }
        else { o[0] = b; o[1] = a; o[2] = z; o[3] = z; }
    }
}

PCD void ext_dirac_massive(const double *m, int helicity, double mass, bool anti, W *o) {
    if (fabs(mass) < 1.0e-8) { ext_dirac_massless(m, helicity, anti, o); return; }
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    int nsf = anti ? (energy > 0.0 ? -1 : 1) : (energy > 0.0 ? 1 : -1);
    int nh = nsf * helicity;
    double pp = fabs(sqrt(px * px + py * py + pz * pz));
    double omega1 = sqrt(fabs(energy) + pp);
    double omega2 = mass / omega1;
    double sf1 = (double)(1 + nsf + (1 - nsf) * nh) * 0.5;
    double sf2 = (double)(1 + nsf - (1 - nsf) * nh) * 0.5;
    int ip = (3 + nh) / 2 - 1, im = (3 - nh) / 2 - 1;
    double om_ip = ip == 0 ? omega1 : omega2, om_im = im == 0 ? omega1 : omega2;
    double sfo0 = sf1 * om_ip, sfo1 = sf2 * om_im;
    double spx = energy > 0.0 ? px : -px, spy = energy > 0.0 ? py : -py, spz = energy > 0.0 ? pz : -pz;
    double pp3 = rmax(pp + spz, 0.0);
    W chi1 = pp == 0.0 ? c64(1.0, 0.0) : c64(sqrt(pp3 * 0.5 / pp), 0.0);
    W chi2;
    if (pp3 == 0.0 || pp == 0.0) chi2 = c64(-(double)nh, 0.0);
    else {
        double denom = sqrt(2.0 * pp * pp3);
        chi2 = c64((double)nh * spx / denom, (anti ? spy : -spy) / denom);
    }
    W chi_im = im == 0 ? chi1 : chi2, chi_ip = ip == 0 ? chi1 : chi2;
    double first = anti ? sfo0 : sfo1, second = anti ? sfo1 : sfo0;
    o[0] = scale_w(chi_im, first);
    o[1] = scale_w(chi_ip, first);
    o[2] = scale_w(chi_im, second);
    o[3] = scale_w(chi_ip, second);
}

PCD void ext_gluon(const double *m, int helicity, W *o) {
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    double sqh = sqrt(0.5);
    bool out = energy > 0.0;
    double hel = out ? (double)helicity : (double)(-helicity), pp = out ? energy : -energy;
    double pt = sqrt(px * px + py * py);
    W wf3 = c64(hel * pt / pp * sqh, 0.0), wf1, wf2;
    if (pt != 0.0) {
        if (out) {
            double pzpt = pz / (pp * pt) * sqh * hel;
            wf1 = c64(-px * pzpt, -py / pt * sqh);
            wf2 = c64(-py * pzpt, px / pt * sqh);
        } else {
            double pzpt = -pz / (pp * pt) * sqh * hel;
            wf1 = c64(px * pzpt, py / pt * sqh);
            wf2 = c64(py * pzpt, -px / pt * sqh);
        }
    } else {
        wf1 = c64(-hel * sqh, 0.0);
        wf2 = c64(0.0, out ? fortran_sign(sqh, pz) : -fortran_sign(sqh, pz));
    }
    o[0] = c64(0.0, 0.0); o[1] = wf1; o[2] = wf2; o[3] = wf3;
}

PCD void ext_massive_vector(const double *m0, int helicity, double mass, W *o) {
    double m[4] = {m0[0], m0[1], m0[2], m0[3]};
    if (m[0] < 0.0) {
        for (int k = 0; k < 4; k++) m[k] = -m[k];
        helicity = -helicity;
    }
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    double sqh = sqrt(0.5);
    double hel = (double)helicity;
    double nsvahl = (double)(helicity < 0 ? -helicity : helicity);
    double pt2 = px * px + py * py;
    double pp = rmin(energy, sqrt(pt2 + pz * pz));
    double pt = rmin(pp, sqrt(pt2));
    double hel0 = 1.0 - fabs(hel);
    if (pp == 0.0) {
        o[0] = c64(0.0, 0.0); o[1] = c64(-hel * sqh, 0.0); o[2] = c64(0.0, nsvahl * sqh); o[3] = c64(hel0, 0.0);
        return;
    }
    double emp = energy / (mass * pp);
    W wf0 = c64(hel0 * pp / mass, 0.0);
    W wf3 = c64(hel0 * pz * emp + hel * pt / pp * sqh, 0.0);
    W wf1, wf2;
    if (pt != 0.0) {
        double pzpt = pz / (pp * pt) * sqh * hel;
        wf1 = c64(hel0 * px * emp - px * pzpt, -nsvahl * py / pt * sqh);
        wf2 = c64(hel0 * py * emp - py * pzpt, nsvahl * px / pt * sqh);
    } else {
        wf1 = c64(-hel * sqh, 0.0);
        wf2 = c64(0.0, nsvahl * fortran_sign(sqh, pz));
    }
    o[0] = wf0; o[1] = wf1; o[2] = wf2; o[3] = wf3;
}

// the wavefunction of a 10-bit source template: family &7, orientation >>3&3, helicity
// (>>5&7)-2, chirality (>>8&3)-1; returns its component count
__device__ inline int source_values(uint32_t code, double mass, const double *p, W *o) {
    int family = code & 7, orientation = (code >> 3) & 3;
    int helicity = (int)((code >> 5) & 7) - 2, chirality = (int)((code >> 8) & 3) - 1;
    if (family == 0) { o[0] = c64(1.0, 0.0); return 1; }
    if (family == 1) { ext_weyl(p, helicity, chirality, orientation != 0, o); return 2; }
    if (family == 2) { ext_dirac_massive(p, helicity, mass, orientation == 1, o); return 4; }
    if (mass == 0.0) ext_gluon(p, helicity, o);
    else ext_massive_vector(p, helicity, mass, o);
    return 4;
}

}  // namespace pacuda
