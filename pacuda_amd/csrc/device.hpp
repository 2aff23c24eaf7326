// SPDX-License-Identifier: 0BSD
//
// Device arithmetic of the recurrence runner for gfx950: df64 pairs, complex helpers, the
// precision policies, the row operations (ops 2-11) and the external wavefunctions (op 12).
//
// Every function keeps the reference's operation order (runner/recurrence_runner.cu:72-542) so
// the f64 kernel is bit-identical to RustiCol's arithmetic; the build passes -ffp-contract=off
// and the pragma below repeats it, so no product or sum is fused unless an fma is written.
// Each value of a row is a register of the lane that owns the point; nothing here touches LDS
// or crosses lanes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define PCD __device__ __forceinline__

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
template <class R> PCD Cx<R> cx(R re, R im) { Cx<R> c; c.re = re; c.im = im; return c; }
template <class R> PCD Cx<R> add(Cx<R> a, Cx<R> b) { return cx(a.re + b.re, a.im + b.im); }
template <class R> PCD Cx<R> sub(Cx<R> a, Cx<R> b) { return cx(a.re - b.re, a.im - b.im); }
template <class R> PCD Cx<R> neg(Cx<R> a) { return cx(-a.re, -a.im); }
template <class R> PCD Cx<R> mul(Cx<R> a, Cx<R> b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
template <class R> PCD Cx<R> mul_i(Cx<R> a) { return cx(-a.im, a.re); }
template <class R> PCD Cx<R> mul_real(Cx<R> a, R s) { return cx(a.re * s, a.im * s); }

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

// The arena as one lane sees it: value (plane q, point i) at q * ld + i of each array. The v1
// path passes the caller's layout (ld = stride); the fused path passes one wave's slice of its
// scratch (ld = 64, i = lane).
struct Arena {
    double *cre, *cim, *are, *aim;
    float *fcre, *fcim;  // mixed kernels: the currents as f32 planes
    const double *mom;
    const double *sre, *sim, *fre, *fim;
    size_t ld;
    uint32_t lds_base;  // the wavefront's LDS region, in elements of the current type (fused path)
};

// Current planes at or above LDS_PLANE live in the wavefront's LDS region (slot plane - LDS_PLANE:
// its real part at lds_base + 2 s ld, its imaginary part ld further); only the fused path's LDS
// pass writes such indices, so v1 programs never take that branch.
constexpr uint32_t LDS_PLANE = 1u << 24;

// Precision policies: R the values, Den the propagator denominators and their forms, Amp the
// amplitude accumulation (runner/recurrence_runner.cu:178-190).
// forwarding: whether the kernel honours F_HOLD / F_FROMACC (every kernel does; a kernel without
// it would run the plan's program without forwarding and LDS residency)
struct KF64 { typedef double real; typedef double den; typedef double amp; static constexpr bool forwarding = true; };
struct KDF64 { typedef D real; typedef D den; typedef D amp; static constexpr bool forwarding = true; };
struct KMixed { typedef float real; typedef D den; typedef D amp; static constexpr bool forwarding = true; };
struct KMixed64 { typedef float real; typedef double den; typedef D amp; static constexpr bool forwarding = true; };

PCD float to_float(float x) { return x; }
PCD float to_float(double x) { return (float)x; }
PCD float to_float(D x) { return x.hi + x.lo; }

template <class R, class Den> struct Form { static PCD R ld(const double *m, size_t at) { return ldr<R>(m, at); } };
template <> struct Form<float, double> { static PCD float ld(const double *m, size_t at) { return (float)m[at]; } };
template <> struct Form<float, D> { static PCD float ld(const double *m, size_t at) { return ldr<float>(m, at); } };

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

template <class R> struct Cur {
    static PCD Cx<R> ld(const Arena &a, size_t at) { return cx(ldr<R>(a.cre, at), ldr<R>(a.cim, at)); }
    static PCD void st(const Arena &a, size_t at, Cx<R> v) { str(a.cre, at, v.re); str(a.cim, at, v.im); }
    // by plane index (slot or LDS) for lane i
    static PCD Cx<R> ldp(const Arena &a, uint32_t plane, size_t i) {
        if (plane >= LDS_PLANE) {
            extern __shared__ double lds_currents[];
            const size_t at = a.lds_base + (size_t)(plane - LDS_PLANE) * 2 * a.ld + i;
            return cx(ldr<R>(lds_currents, at), ldr<R>(lds_currents, at + a.ld));
        }
        return ld(a, (size_t)plane * a.ld + i);
    }
    static PCD void stp(const Arena &a, uint32_t plane, size_t i, Cx<R> v) {
        if (plane >= LDS_PLANE) {
            extern __shared__ double lds_currents[];
            const size_t at = a.lds_base + (size_t)(plane - LDS_PLANE) * 2 * a.ld + i;
            str(lds_currents, at, v.re);
            str(lds_currents, at + a.ld, v.im);
            return;
        }
        st(a, (size_t)plane * a.ld + i, v);
    }
};
template <> struct Cur<float> {
    static PCD Cx<float> ld(const Arena &a, size_t at) { return cx(a.fcre[at], a.fcim[at]); }
    static PCD void st(const Arena &a, size_t at, Cx<float> v) { a.fcre[at] = v.re; a.fcim[at] = v.im; }
    static PCD Cx<float> ldp(const Arena &a, uint32_t plane, size_t i) {
        if (plane >= LDS_PLANE) {
            extern __shared__ float lds_currents_f32[];
            const size_t at = a.lds_base + (size_t)(plane - LDS_PLANE) * 2 * a.ld + i;
            return cx(lds_currents_f32[at], lds_currents_f32[at + a.ld]);
        }
        return ld(a, (size_t)plane * a.ld + i);
    }
    static PCD void stp(const Arena &a, uint32_t plane, size_t i, Cx<float> v) {
        if (plane >= LDS_PLANE) {
            extern __shared__ float lds_currents_f32[];
            const size_t at = a.lds_base + (size_t)(plane - LDS_PLANE) * 2 * a.ld + i;
            lds_currents_f32[at] = v.re;
            lds_currents_f32[at + a.ld] = v.im;
            return;
        }
        st(a, (size_t)plane * a.ld + i, v);
    }
};

template <class R> PCD R reciprocal_minkowski_norm(R p0, R p1, R p2, R p3) {
    return R(1.0) / (p0 * p0 - p1 * p1 - p2 * p2 - p3 * p3);
}
template <class R> PCD Cx<R> minkowski_dot(const Cx<R> *l, const Cx<R> *r) {
    return sub(sub(sub(mul(l[0], r[0]), mul(l[1], r[1])), mul(l[2], r[2])), mul(l[3], r[3]));
}
template <class R> PCD Cx<R> minkowski_real_dot(const Cx<R> *l, const R *q) {
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
PCD void contribution_values(const Cx<R> *l, const Cx<R> *r, const R *q0, const R *q1, Cx<R> *o) {
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
PCD void propagator_values(const Cx<R> *c, const R *p, R reciprocal, Cx<R> row_factor, Cx<R> *o) {
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

template <class R> PCD Cx<R> identity_value(R fr, R fi, Cx<R> v) { return cx(fr * v.re - fi * v.im, fr * v.im + fi * v.re); }

template <class R> PCD Cx<R> closure_increment(Cx<R> l, Cx<R> r, R cr, R ci, R rfr, R rfi) {
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
