// SPDX-License-Identifier: 0BSD
//
// refcpu: the CPU oracle of the recurrence runner. TEST INFRASTRUCTURE ONLY: it is imported by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the checker and the
// CPU baseline; the product (pacuda_amd/lib/librecurrence.so) never loads it.
//
// It restates, operation for operation, the reference runner's arithmetic and call semantics:
//   df64 arithmetic                      runner/recurrence_runner.cu:72-126 (Joldes, Muller,
//                                        Popescu, ACM TOMS 44(2) 2017, Alg. 6, 9, 12, 17)
//   complex helpers, Minkowski dots      :128-142, :173-176, :220-227
//   precision policy (f64/df64/mixed)    :178-218, :544-552
//   contribution ops 2-6                 :231-282 (legend convert/CONVENTIONS.md:71-79)
//   propagators 7-9, identity 10,
//   closure 11                           :284-330
//   external wavefunctions (op 12)       :332-542 (legend convert/CONVENTIONS.md:52-61)
//   contribution/propagator rows         :686-743
//   momentum forms                       :745-764
//   sources kernel                       :766-785
//   interpreter                          :787-838
//   conversions                          :840-880
//   decode, sort, program                :914-1054
//   the v1 call (run_locked)             :1290-1487
//   reductions and sampling              :1489-1723
//   lock and entry points                :1725-1865
// Built with -O2 -ffp-contract=off (no contraction: every product and sum is rounded on its
// own, as the reference's -fmad=false build); fma/fmaf are explicit.
//
// Exported: refcpu_v1_* (the same contracts as pacman_recurrence_v1_*, on host memory standing in
// for the device), refcpu_test_sources, refcpu_evaluate (the whole plan run per point from the
// external momenta to the per-point total: the CPU baseline). Built with -DREFCPU_ABI_NAMES the
// v1 functions are also exported under their pacman_recurrence_v1_* names, so the reference's
// runner/rectest.c can drive the oracle.

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "pacman_recurrence.h"
// the runner's generated symjit kernels (the functions its op-13 rows call)
#include "../pacuda_amd/csrc/plane_kernels.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---- df64: a double emulated by a pair of floats (hi, lo) ----------------------------------

struct D {
    float hi, lo;
    D() {}
    D(float h, float l) : hi(h), lo(l) {}
    explicit D(double x) : hi((float)x), lo(0.0f) {}  // exact constants 0, +-1, 2
};

inline void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
inline void fast_two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    e = b - (s - a);
}
// AccurateDWPlusDW (JMP 2017 Alg. 6)
inline D operator+(D x, D y) {
    float sh, sl, th, tl, vh, vl;
    two_sum(x.hi, y.hi, sh, sl);
    two_sum(x.lo, y.lo, th, tl);
    fast_two_sum(sh, sl + th, vh, vl);
    D z;
    fast_two_sum(vh, tl + vl, z.hi, z.lo);
    return z;
}
inline D operator-(D x) { return D(-x.hi, -x.lo); }
inline D operator-(D x, D y) { return x + (-y); }
// DWTimesDW3 (JMP 2017 Alg. 12)
inline D operator*(D x, D y) {
    float ch = x.hi * y.hi;
    float cl1 = fmaf(x.hi, y.hi, -ch);
    float tl0 = x.lo * y.lo;
    float tl1 = fmaf(x.hi, y.lo, tl0);
    float cl2 = fmaf(x.lo, y.hi, tl1);
    D z;
    fast_two_sum(ch, cl1 + cl2, z.hi, z.lo);
    return z;
}
// DWDivDW2 (JMP 2017 Alg. 17) with DWTimesFP3 (Alg. 9)
inline D operator/(D x, D y) {
    float th = x.hi / y.hi;
    float rh = y.hi * th;
    float rl1 = fmaf(y.hi, th, -rh);
    float rl3 = fmaf(y.lo, th, rl1);
    float r_hi, r_lo;
    fast_two_sum(rh, rl3, r_hi, r_lo);
    float pi_h = x.hi - r_hi;
    float delta = pi_h + (x.lo - r_lo);
    D z;
    fast_two_sum(th, delta / y.hi, z.hi, z.lo);
    return z;
}

// ---- complex numbers over R ---------------------------------------------------------------

template <class R> struct Cx { R re, im; };
template <class R> inline Cx<R> cx(R re, R im) { Cx<R> c; c.re = re; c.im = im; return c; }
template <class R> inline Cx<R> add(Cx<R> a, Cx<R> b) { return cx(a.re + b.re, a.im + b.im); }
template <class R> inline Cx<R> sub(Cx<R> a, Cx<R> b) { return cx(a.re - b.re, a.im - b.im); }
template <class R> inline Cx<R> neg(Cx<R> a) { return cx(-a.re, -a.im); }
template <class R> inline Cx<R> mul(Cx<R> a, Cx<R> b) {
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class R> inline Cx<R> mul_i(Cx<R> a) { return cx(-a.im, a.re); }
template <class R> inline Cx<R> mul_real(Cx<R> a, R s) { return cx(a.re * s, a.im * s); }

// ---- the "device" memory: host arrays in the device byte layout --------------------------
// A df64 value occupies the 8 bytes of the double it replaces as (float hi, float lo).

template <class R> inline R ldr(const double *base, size_t at);
template <> inline double ldr<double>(const double *base, size_t at) { return base[at]; }
template <> inline D ldr<D>(const double *base, size_t at) {
    float v[2];
    memcpy(v, base + at, 8);
    return D(v[0], v[1]);
}
template <> inline float ldr<float>(const double *base, size_t at) {
    float v[2];
    memcpy(v, base + at, 8);
    return v[0];
}
inline void str(double *base, size_t at, double v) { base[at] = v; }
inline void str(double *base, size_t at, D v) {
    float f[2] = {v.hi, v.lo};
    memcpy(base + at, f, 8);
}
inline void str(double *base, size_t at, float v) {
    float f[2] = {v, 0.0f};
    memcpy(base + at, f, 8);
}

struct Arena {
    double *cre, *cim, *are, *aim;
    float *fcre, *fcim;
    const double *mom;
    const double *sre, *sim, *fre, *fim;
    size_t ld;
};

// the precision policies (runner/recurrence_runner.cu:178-190); acc: the vertex arithmetic and
// the accumulation of a destination's contributions, the current type R in the reference's four
struct KF64 { typedef double real; typedef double den; typedef double amp; typedef double acc; };
struct KDF64 { typedef D real; typedef D den; typedef D amp; typedef D acc; };
struct KMixed { typedef float real; typedef D den; typedef D amp; typedef float acc; };
struct KMixed64 { typedef float real; typedef double den; typedef D amp; typedef float acc; };
// the library's compensated mixed policy (no counterpart in the reference): KMixed with the
// arithmetic of every row (vertices and the sum of a destination's contributions, propagators,
// identity rows, the closures' products) in df64 from the f32 currents, the df64 forms and the
// df64 constants; a current is rounded to f32 (fl(hi + lo)) once, when it is stored, and a
// forwarded one as if it had been stored
struct KComp { typedef float real; typedef D den; typedef D amp; typedef D acc; };
// the library's compensated policy with f64 arithmetic (no counterpart in the reference): the
// same f32 currents and roundings with every other value and operation in f64, which KComp's
// df64 follows to about 2^-44 per operation
struct KComp64 { typedef float real; typedef double den; typedef double amp; typedef double acc; };

inline float to_float(float x) { return x; }
inline float to_float(double x) { return (float)x; }
inline float to_float(D x) { return x.hi + x.lo; }

template <class R, class Den> inline R ld_form(const double *mom, size_t at) { return ldr<R>(mom, at); }
template <> inline float ld_form<float, double>(const double *mom, size_t at) { return (float)mom[at]; }
template <> inline float ld_form<float, D>(const double *mom, size_t at) { return ldr<float>(mom, at); }

template <class R, class S> struct ToReal { static R of(S x) { return x; } };
template <class S> struct ToReal<float, S> { static float of(S x) { return to_float(x); } };

inline D widen(float x) { return D(x, 0.0f); }
inline double widen(double x) { return x; }
inline D widen(D x) { return x; }

// R <-> acc: the exact widening of an f32 to df64, and a df64 rounded to the nearest f32
template <class T, class S> struct Conv { static T of(S x) { return x; } };
template <> struct Conv<D, float> { static D of(float x) { return D(x, 0.0f); } };
template <> struct Conv<float, D> { static float of(D x) { return x.hi + x.lo; } };
template <class T, class S> inline Cx<T> cvt(Cx<S> c) { return cx(Conv<T, S>::of(c.re), Conv<T, S>::of(c.im)); }

template <class R> inline R narrow(double x);
template <> inline double narrow<double>(double x) { return x; }
template <> inline float narrow<float>(double x) { return (float)x; }
template <> inline D narrow<D>(double x) {
    float hi = (float)x;
    return D(hi, (float)(x - (double)hi));
}

template <class R> inline Cx<R> ld_cur(const Arena &a, size_t at) { return cx(ldr<R>(a.cre, at), ldr<R>(a.cim, at)); }
template <> inline Cx<float> ld_cur<float>(const Arena &a, size_t at) { return cx(a.fcre[at], a.fcim[at]); }
template <class R> inline void st_cur(const Arena &a, size_t at, Cx<R> v) {
    str(a.cre, at, v.re);
    str(a.cim, at, v.im);
}
template <> inline void st_cur<float>(const Arena &a, size_t at, Cx<float> v) {
    a.fcre[at] = v.re;
    a.fcim[at] = v.im;
}

template <class R> inline R reciprocal_minkowski_norm(R p0, R p1, R p2, R p3) {
    return R(1.0) / (p0 * p0 - p1 * p1 - p2 * p2 - p3 * p3);
}
template <class R> inline Cx<R> minkowski_dot(const Cx<R> *l, const Cx<R> *r) {
    return sub(sub(sub(mul(l[0], r[0]), mul(l[1], r[1])), mul(l[2], r[2])), mul(l[3], r[3]));
}
template <class R> inline Cx<R> minkowski_real_dot(const Cx<R> *l, const R *q) {
    return sub(sub(sub(mul_real(l[0], q[0]), mul_real(l[1], q[1])), mul_real(l[2], q[2])), mul_real(l[3], q[3]));
}

void contribution_shape(int op, int *c0, int *c1, int *cd) {
    *c0 = op == 4 ? 6 : (op == 5 || op == 6) ? 4 : 2;
    *c1 = 4;
    *cd = (op == 4 || op == 6) ? 4 : op == 5 ? 6 : 2;
}

// ops 2-6: the values before the scale (convert/CONVENTIONS.md:71-79, in the reference's order)
template <class R>
void contribution_values(int op, const Cx<R> *l, const Cx<R> *r, const R *q0, const R *q1, Cx<R> *o) {
    if (op == 2) {
        o[0] = add(add(sub(neg(mul_i(mul(l[0], r[3]))), mul_i(mul(l[1], r[1]))), mul(l[1], r[2])), mul_i(mul(l[0], r[0])));
        o[1] = add(add(sub(neg(mul(l[0], r[2])), mul_i(mul(l[0], r[1]))), mul_i(mul(l[1], r[0]))), mul_i(mul(l[1], r[3])));
    } else if (op == 3) {
        o[0] = add(add(add(neg(mul(l[1], r[2])), mul_i(mul(l[0], r[0]))), mul_i(mul(l[0], r[3]))), mul_i(mul(l[1], r[1])));
        o[1] = add(add(add(neg(mul_i(mul(l[1], r[3]))), mul(l[0], r[2])), mul_i(mul(l[0], r[1]))), mul_i(mul(l[1], r[0])));
    } else if (op == 4) {
        o[0] = add(add(mul(l[0], r[1]), mul(l[1], r[2])), mul(l[2], r[3]));
        o[1] = add(add(mul(l[0], r[0]), mul(l[3], r[2])), mul(l[4], r[3]));
        o[2] = add(sub(mul(l[1], r[0]), mul(l[3], r[1])), mul(l[5], r[3]));
        o[3] = sub(sub(mul(l[2], r[0]), mul(l[4], r[1])), mul(l[5], r[2]));
    } else if (op == 5) {
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
        for (int k = 0; k < 4; k++) {
            Cx<R> mt = mul_real(vd, q0[k] - q1[k]);
            Cx<R> ct = mul_real(sub(mul(ld1, r[k]), mul(rd0, l[k])), R(2.0));
            o[k] = add(mt, ct);
        }
    }
}

template <class R>
int propagator_values(int op, const Cx<R> *c, const R *p, R reciprocal, Cx<R> row_factor, Cx<R> *o) {
    if (op == 9) {
        Cx<R> scale = mul_real(mul(row_factor, cx(R(0.0), R(-1.0))), reciprocal);
        for (int k = 0; k < 4; k++) o[k] = mul(c[k], scale);
        return 4;
    }
    Cx<R> scale = mul_real(row_factor, reciprocal);
    Cx<R> o0, o1;
    if (op == 7) {
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
    return 2;
}

template <class R> inline Cx<R> identity_value(R fr, R fi, Cx<R> v) {
    return cx(fr * v.re - fi * v.im, fr * v.im + fi * v.re);
}

template <class R> inline Cx<R> closure_increment(Cx<R> l, Cx<R> r, R cr, R ci, R rfr, R rfi) {
    R scale_re = cr * rfr - ci * rfi;
    R scale_im = cr * rfi + ci * rfr;
    R product_re = l.re * r.re - l.im * r.im;
    R product_im = l.re * r.im + l.im * r.re;
    return cx(product_re * scale_re - product_im * scale_im, product_re * scale_im + product_im * scale_re);
}

// ---- external wavefunctions (op 12), f64 (runner/recurrence_runner.cu:332-542) ------------

typedef Cx<double> W;
inline W c64(double re, double im) { return cx(re, im); }
inline W scale_w(W c, double s) { return cx(c.re * s, c.im * s); }
inline double fortran_sign(double value, double sign_source) { return copysign(fabs(value), sign_source); }
// max/min returning the other operand when one is NaN; on equal operands (+0 vs -0) the second
inline double rmax(double a, double b) { return a != a ? b : b != b ? a : a > b ? a : b; }
inline double rmin(double a, double b) { return a != a ? b : b != b ? a : a < b ? a : b; }

void ext_quark_weyl_array(const double *m, int helicity, int chirality, W *o) {
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    if (energy > 0.0) {
        double s = (px == 0.0 && py == 0.0 && pz < 0.0) ? 0.0 : sqrt(rmax(energy + pz, 0.0));
        W chi1 = c64(s, 0.0);
        W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * energy), 0.0) : c64((double)helicity * px / s, -py / s);
        if (helicity == 1 && chirality == 1) { o[0] = chi1; o[1] = chi2; return; }
        if (helicity == -1 && chirality == -1) { o[0] = chi2; o[1] = chi1; return; }
        o[0] = c64(0.0, 0.0); o[1] = c64(0.0, 0.0);
        return;
    }
    double s = (px == 0.0 && py == 0.0 && pz > 0.0) ? 0.0 : -sqrt(rmax(-(energy + pz), 0.0));
    W chi1 = c64(s, 0.0);
    W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * fabs(energy)), 0.0) : c64(-(double)helicity * (-px) / s, py / s);
    if (helicity == -1 && chirality == 1) { o[0] = chi1; o[1] = chi2; }
    else if (helicity == 1 && chirality == -1) { o[0] = chi2; o[1] = chi1; }
    else { o[0] = c64(0.0, 0.0); o[1] = c64(0.0, 0.0); }
}

void ext_antiquark_weyl_array(const double *m, int helicity, int chirality, W *o) {
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    if (energy > 0.0) {
        double s = (px == 0.0 && py == 0.0 && pz < 0.0) ? 0.0 : -sqrt(rmax(energy + pz, 0.0));
        W chi1 = c64(s, 0.0);
        W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * energy), 0.0) : c64(-(double)helicity * px / s, py / s);
        if (helicity == 1 && chirality == 1) { o[0] = chi2; o[1] = chi1; return; }
        if (helicity == -1 && chirality == -1) { o[0] = chi1; o[1] = chi2; return; }
        o[0] = c64(0.0, 0.0); o[1] = c64(0.0, 0.0);
        return;
    }
    double s = (px == 0.0 && py == 0.0 && pz > 0.0) ? 0.0 : sqrt(rmax(-(energy + pz), 0.0));
    W chi1 = c64(s, 0.0);
    W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * fabs(energy)), 0.0) : c64((double)helicity * (-px) / s, -py / s);
    if (helicity == -1 && chirality == 1) { o[0] = chi2; o[1] = chi1; }
    else if (helicity == 1 && chirality == -1) { o[0] = chi1; o[1] = chi2; }
    else { o[0] = c64(0.0, 0.0); o[1] = c64(0.0, 0.0); }
}

void ext_quark_dirac(const double *m, int helicity, W *o) {
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    W z = c64(0.0, 0.0);
    if (energy > 0.0) {
        double s = (px == 0.0 && py == 0.0 && pz < 0.0) ? 0.0 : sqrt(rmax(energy + pz, 0.0));
        W chi1 = c64(s, 0.0);
        W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * energy), 0.0) : c64((double)helicity * px / s, -py / s);
        if (helicity == 1) { o[0] = chi1; o[1] = chi2; o[2] = z; o[3] = z; return; }
        o[0] = z; o[1] = z; o[2] = chi2; o[3] = chi1;
        return;
    }
    double s = (px == 0.0 && py == 0.0 && pz > 0.0) ? 0.0 : -sqrt(rmax(-(energy + pz), 0.0));
    W chi1 = c64(s, 0.0);
    W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * fabs(energy)), 0.0) : c64(-(double)helicity * (-px) / s, py / s);
    if (-helicity == 1) { o[0] = chi1; o[1] = chi2; o[2] = z; o[3] = z; }
    else { o[0] = z; o[1] = z; o[2] = chi2; o[3] = chi1; }
}

void ext_antiquark_dirac(const double *m, int helicity, W *o) {
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    W z = c64(0.0, 0.0);
    if (energy > 0.0) {
        double s = (px == 0.0 && py == 0.0 && pz < 0.0) ? 0.0 : -sqrt(rmax(energy + pz, 0.0));
        W chi1 = c64(s, 0.0);
        W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * energy), 0.0) : c64(-(double)helicity * px / s, py / s);
        if (-helicity == 1) { o[0] = z; o[1] = z; o[2] = chi1; o[3] = chi2; return; }
        o[0] = chi2; o[1] = chi1; o[2] = z; o[3] = z;
        return;
    }
    double s = (px == 0.0 && py == 0.0 && pz > 0.0) ? 0.0 : sqrt(rmax(-(energy + pz), 0.0));
    W chi1 = c64(s, 0.0);
    W chi2 = s == 0.0 ? c64(-(double)helicity * sqrt(2.0 * fabs(energy)), 0.0) : c64((double)helicity * (-px) / s, -py / s);
    if (helicity == 1) { o[0] = z; o[1] = z; o[2] = chi1; o[3] = chi2; }
    else { o[0] = chi2; o[1] = chi1; o[2] = z; o[3] = z; }
}

void ext_dirac_massive(const double *m, int helicity, double mass, bool antiquark, W *o) {
    if (fabs(mass) < 1.0e-8) {
        if (antiquark) ext_antiquark_dirac(m, helicity, o);
        else ext_quark_dirac(m, helicity, o);
        return;
    }
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    int nsf = antiquark ? (energy > 0.0 ? -1 : 1) : (energy > 0.0 ? 1 : -1);
    int nh = nsf * helicity;
    double pp = fabs(sqrt(px * px + py * py + pz * pz));
    double omega1 = sqrt(fabs(energy) + pp);
    double omega2 = mass / omega1;
    double omega[2] = {omega1, omega2};
    double sf1 = (double)(1 + nsf + (1 - nsf) * nh) * 0.5;
    double sf2 = (double)(1 + nsf - (1 - nsf) * nh) * 0.5;
    int ip = (3 + nh) / 2 - 1, im = (3 - nh) / 2 - 1;
    double sfomeg[2] = {sf1 * omega[ip], sf2 * omega[im]};
    double spx = energy > 0.0 ? px : -px, spy = energy > 0.0 ? py : -py, spz = energy > 0.0 ? pz : -pz;
    double pp3 = rmax(pp + spz, 0.0);
    W chi1 = pp == 0.0 ? c64(1.0, 0.0) : c64(sqrt(pp3 * 0.5 / pp), 0.0);
    W chi2;
    if (pp3 == 0.0 || pp == 0.0) chi2 = c64(-(double)nh, 0.0);
    else {
        double denom = sqrt(2.0 * pp * pp3);
        chi2 = c64((double)nh * spx / denom, (antiquark ? spy : -spy) / denom);
    }
    W chi[2] = {chi1, chi2};
    double first = antiquark ? sfomeg[0] : sfomeg[1], second = antiquark ? sfomeg[1] : sfomeg[0];
    o[0] = scale_w(chi[im], first);
    o[1] = scale_w(chi[ip], first);
    o[2] = scale_w(chi[im], second);
    o[3] = scale_w(chi[ip], second);
}

void ext_gluon(const double *m, int helicity, W *o) {
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    double sqh = sqrt(0.5);
    if (energy > 0.0) {
        double hel = (double)helicity, pp = energy;
        double pt = sqrt(px * px + py * py);
        W wf3 = c64(hel * pt / pp * sqh, 0.0), wf1, wf2;
        if (pt != 0.0) {
            double pzpt = pz / (pp * pt) * sqh * hel;
            wf1 = c64(-px * pzpt, -py / pt * sqh);
            wf2 = c64(-py * pzpt, px / pt * sqh);
        } else {
            wf1 = c64(-hel * sqh, 0.0);
            wf2 = c64(0.0, fortran_sign(sqh, pz));
        }
        o[0] = c64(0.0, 0.0); o[1] = wf1; o[2] = wf2; o[3] = wf3;
        return;
    }
    double hel = (double)(-helicity), pp = -energy;
    double pt = sqrt(px * px + py * py);
    W wf3 = c64(hel * pt / pp * sqh, 0.0), wf1, wf2;
    if (pt != 0.0) {
        double pzpt = -pz / (pp * pt) * sqh * hel;
        wf1 = c64(px * pzpt, py / pt * sqh);
        wf2 = c64(py * pzpt, -px / pt * sqh);
    } else {
        wf1 = c64(-hel * sqh, 0.0);
        wf2 = c64(0.0, -fortran_sign(sqh, pz));
    }
    o[0] = c64(0.0, 0.0); o[1] = wf1; o[2] = wf2; o[3] = wf3;
}

void ext_massive_vector(const double *m0, int helicity, double mass, W *o) {
    double m[4] = {m0[0], m0[1], m0[2], m0[3]};
    if (m[0] < 0.0) {
        for (int k = 0; k < 4; k++) m[k] = -m[k];
        helicity = -helicity;
    }
    double energy = m[0], px = m[1], py = m[2], pz = m[3];
    double sqh = sqrt(0.5);
    double hel = (double)helicity;
    double nsvahl = (double)abs(helicity);
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

int source_values(uint32_t code, double mass, const double *p, W *o) {
    int family = code & 7, orientation = (code >> 3) & 3;
    int helicity = (int)((code >> 5) & 7) - 2, chirality = (int)((code >> 8) & 3) - 1;
    if (family == 0) { o[0] = c64(1.0, 0.0); return 1; }
    if (family == 1) {
        if (orientation == 0) ext_quark_weyl_array(p, helicity, chirality, o);
        else ext_antiquark_weyl_array(p, helicity, chirality, o);
        return 2;
    }
    if (family == 2) { ext_dirac_massive(p, helicity, mass, orientation == 1, o); return 4; }
    if (mass == 0.0) ext_gluon(p, helicity, o);
    else ext_massive_vector(p, helicity, mass, o);
    return 4;
}

// ---- rows: decode, sort, program (runner/recurrence_runner.cu:914-1054) -------------------

enum { K_CLEAR = 1, K_CONTRIB, K_PROP, K_IDENTITY, K_CLOSURE, K_SOURCE, K_SYMJIT };
// F_HOLD / F_FROMACC: forwarding through the accumulators, as the library's fused path emits it
// (a run keeps its result for the next in-place row, which takes it instead of loading); the
// reference's programs never carry them. Used by refcpu_evaluate_program.
enum { F_INIT = 1, F_FIRST = 2, F_LAST = 4, F_HOLD = 8, F_FROMACC = 16 };

struct Row {
    int op;
    uint32_t f[8];
    std::vector<uint32_t> data;
};

std::vector<Row> decode(const uint64_t *words, uint32_t count, const uint8_t *width) {
    std::vector<Row> rows;
    for (uint32_t n = 0; n < count; n++) {
        uint64_t w = words[n];
        Row r;
        memset(r.f, 0, sizeof r.f);
        if ((w & 15) == 14) {
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
            rows.push_back(r);
            continue;
        }
        auto take = [&](int bits) {
            uint32_t v = (uint32_t)(w & ((1ull << bits) - 1));
            w >>= bits;
            return v;
        };
        r.op = (int)take(4);
        const int P = width[W_P], A = width[W_A], F = width[W_F], X = width[W_X], S = width[W_S], K = width[W_K];
        std::vector<int> fields;
        if (r.op == 1) fields = {P, P};
        else if (r.op >= 2 && r.op <= 6) fields = {1, S, P, P, P, F, F, X};
        else if (r.op >= 7 && r.op <= 9) fields = {P, F, X};
        else if (r.op == 10) fields = {P, K, X};
        else if (r.op == 11) fields = {P, P, A, X, X, K};
        else if (r.op == 12) fields = {P, F, X, S, 10};
        else if (r.op == 13) fields = {1, P, P, P, F, F, X};
        for (size_t k = 0; k < fields.size(); k++) r.f[k] = take(fields[k]);
        rows.push_back(r);
    }
    return rows;
}

void sort_contribution_runs(std::vector<Row> &rows) {
    size_t n = 0;
    while (n < rows.size()) {
        if (rows[n].op < 2 || rows[n].op > 6) { n++; continue; }
        size_t end = n;
        while (end < rows.size() && rows[end].op >= 2 && rows[end].op <= 6) end++;
        std::map<uint32_t, uint32_t> dst_base;
        bool safe = true;
        for (size_t k = n; k < end && safe; k++) {
            int c0, c1, cd;
            contribution_shape(rows[k].op, &c0, &c1, &cd);
            for (int q = 0; q < cd; q++) {
                auto it = dst_base.emplace(rows[k].f[4] + q, rows[k].f[4]);
                if (it.first->second != rows[k].f[4]) safe = false;
            }
        }
        for (size_t k = n; k < end && safe; k++) {
            int c0, c1, cd;
            contribution_shape(rows[k].op, &c0, &c1, &cd);
            for (int q = 0; q < c0 && safe; q++) safe = !dst_base.count(rows[k].f[2] + q);
            for (int q = 0; q < c1 && safe; q++) safe = !dst_base.count(rows[k].f[3] + q);
        }
        if (safe)
            std::stable_sort(rows.begin() + n, rows.begin() + end, [](const Row &x, const Row &y) { return x.f[4] < y.f[4]; });
        n = end;
    }
}

// An executable row: kind, op, flags and the fields in word order; z: the known-zero component
// masks of a fused-path instruction (0 for the reference's own programs): z[0] the first loaded
// operand (contribution parent 0, the in-place input, closure parent 0), z[1] contribution or
// closure parent 1, z[2] a contribution's old destination value, z[3] the stored value. A masked
// load reads +0 and a masked store is skipped, as the fused kernels do in the slot.
// dual[k] / chip[k]: plane field f[k] of a fused program names a slot plane with a copy on chip
// (the library's DUAL_FILL: a read loads the slot and keeps the copy, a write goes on chip;
// DUAL_THROUGH: reads on chip, writes on chip and to the slot); zero in the reference's programs
enum { DUAL_FILL = 1u << 31, DUAL_THROUGH = 1u << 30 };
// bind: a symjit row's binding (kernel, inputs, output pairs, direct, input codes, output
// components), copied from the op-14 row before it
constexpr int BIND_WORDS = 4 + PACUDA_MAX_IN + PACUDA_MAX_OUT / 2;
struct Op {
    int kind, op, flags;
    uint32_t f[8];
    uint32_t z[4];
    uint32_t dual[8], chip[8];
    uint32_t bind[BIND_WORDS];
};

// a current component of plane field k (component c, known-zero bit b of mask z) read or written
// as the op says where it lives
template <class R> Cx<R> ld_cur_z(const Arena &a, size_t at, uint32_t z, int k);
template <class R> Cx<R> ld_field(const Arena &a, const Op &o, int k, uint32_t c, uint32_t z, int b, size_t i) {
    if (o.dual[k] & DUAL_FILL) {
        const Cx<R> v = ld_cur_z<R>(a, (o.f[k] + c) * a.ld + i, z, b);
        st_cur<R>(a, (o.chip[k] + c) * a.ld + i, v);
        return v;
    }
    return ld_cur_z<R>(a, ((o.dual[k] ? o.chip[k] : o.f[k]) + c) * a.ld + i, z, b);
}
template <class R> void st_field(const Arena &a, const Op &o, int k, uint32_t c, uint32_t z, int b, Cx<R> v, size_t i) {
    const bool masked = (z >> b) & 1;
    if (o.dual[k]) {  // on chip every component is kept (as the kernels keep it)
        st_cur<R>(a, (o.chip[k] + c) * a.ld + i, v);
        if ((o.dual[k] & DUAL_THROUGH) && !masked) st_cur<R>(a, (o.f[k] + c) * a.ld + i, v);
    } else if (!masked) {
        st_cur<R>(a, (o.f[k] + c) * a.ld + i, v);
    }
}

bool program(const std::vector<Row> &rows, std::vector<Op> &ops, uint32_t *sources) {
    const Row *binding = nullptr;  // the last binding (op 14) before a symjit row (op 13)
    for (size_t n = 0; n < rows.size(); n++) {
        const Row &r = rows[n];
        if (r.op == 14) {
            binding = &r;
            continue;
        }
        Op o;
        memset(&o, 0, sizeof o);
        o.op = r.op;
        if (r.op == 13) {
            // a binding of a kernel the runner carries, within its input and output limits
            // (runner/recurrence_runner.cu:1142-1151)
            if (!binding || binding->f[0] >= PACUDA_KERNEL_COUNT || binding->f[1] > PACUDA_MAX_IN ||
                2 * binding->f[2] > PACUDA_MAX_OUT) {
                fprintf(stderr, "refcpu: a symjit row without a valid binding\n");
                return false;
            }
            o.kind = K_SYMJIT;
            o.flags = r.f[0] ? F_INIT : 0;
            memcpy(o.f, r.f, sizeof o.f);
            for (int k = 0; k < 4; k++) o.bind[k] = binding->f[k];
            for (size_t k = 0; k < binding->data.size(); k++) o.bind[4 + k] = binding->data[k];
            ops.push_back(o);
            continue;
        }
        if (r.op >= 2 && r.op <= 6) {
            bool first = n == 0 || rows[n - 1].op < 2 || rows[n - 1].op > 6 || rows[n - 1].f[4] != r.f[4];
            bool last = n + 1 == rows.size() || rows[n + 1].op < 2 || rows[n + 1].op > 6 || rows[n + 1].f[4] != r.f[4];
            o.kind = K_CONTRIB;
            o.flags = (r.f[0] ? F_INIT : 0) | (first ? F_FIRST : 0) | (last ? F_LAST : 0);
        } else {
            o.kind = r.op == 1 ? K_CLEAR : r.op <= 9 ? K_PROP : r.op == 10 ? K_IDENTITY : r.op == 11 ? K_CLOSURE : K_SOURCE;
            if (r.op < 1 || r.op > 12) {
                fprintf(stderr, "refcpu: op %d\n", r.op);
                return false;
            }
        }
        memcpy(o.f, r.f, sizeof o.f);
        ops.push_back(o);
    }
    uint32_t s = 0;
    while (s < ops.size() && ops[s].kind == K_SOURCE) s++;
    for (size_t n = s; n < ops.size(); n++)
        if (ops[n].kind == K_SOURCE) {
            fprintf(stderr, "refcpu: a source row follows other rows\n");
            return false;
        }
    *sources = s;
    return true;
}

// ---- execution on a block of points [i0, i1) ----------------------------------------------

void fill_forms(const double *external, size_t ext_points, const uint32_t *form_terms, uint32_t form_count,
                const uint32_t *slots, const double *coefficients, double *mom, size_t ld, size_t i0, size_t i1,
                size_t ext_offset) {
    for (uint32_t f = 0; f < form_count; f++) {
        uint32_t first = form_terms[2 * f], n = form_terms[2 * f + 1];
        for (int k = 0; k < 4; k++)
            for (size_t i = i0; i < i1; i++) {
                double v = 0.0;
                for (uint32_t t = 0; t < n; t++) {
                    double term = coefficients[first + t] * external[((size_t)slots[first + t] * 4 + k) * ext_points + ext_offset + i];
                    v = t == 0 ? 0.0 + term : v + term;
                }
                mom[(f * 4 + k) * ld + i] = v;
            }
    }
}

template <class R>
void run_sources(const Op *ops, uint32_t count, const Arena &a, size_t i0, size_t i1) {
    for (uint32_t n = 0; n < count; n++) {
        const Op &o = ops[n];
        uint32_t dst = o.f[0], form = o.f[1], factor = o.f[2], mslot = o.f[3], code = o.f[4];
        double fr = a.fre[factor], fi = a.fim[factor], mass = a.sre[mslot];
        for (size_t i = i0; i < i1; i++) {
            double p[4];
            for (int k = 0; k < 4; k++) p[k] = a.mom[(form * 4 + k) * a.ld + i];
            W v[4];
            int c = source_values(code, mass, p, v);
            for (int k = 0; k < c; k++) {
                W s = cx(fr * v[k].re - fi * v[k].im, fr * v[k].im + fi * v[k].re);
                st_field<R>(a, o, 0, k, o.z[3], k, cx(narrow<R>(s.re), narrow<R>(s.im)), i);
            }
        }
    }
}

inline double f64_of(double x) { return x; }
inline double f64_of(float x) { return (double)x; }
inline double f64_of(D x) { return (double)x.hi + (double)x.lo; }

// A symjit row (runner/recurrence_runner.cu:631-684) per point: the binding's inputs as f64
// (kind 0 a parent's component, 1 a form component, 2 a destination component, 3 the real part of
// a scale, else the row factor; bit 3 which parent or form, bit 4 the imaginary part, bits 5 and
// up the component or slot), the kernel, then each output pair times the factor (re fr - im fi,
// re fi + im fr) added to the destination component unless the row initialises it, or stored
// unscaled when the binding is direct and the factor is exactly (1, +0); narrowed to R. (The
// scales and factor are read as the policy's acc type: R in the reference's policies.)
template <class K> void symjit_row(const Op &o, const Arena &a, size_t i0, size_t i1) {
    typedef typename K::real R;
    typedef typename K::den Den;
    typedef typename K::acc T;
    const uint32_t kernel = o.bind[0], n_in = o.bind[1], pairs = o.bind[2], direct = o.bind[3];
    const uint32_t parent[2] = {o.f[1], o.f[2]}, dst = o.f[3], form[2] = {o.f[4], o.f[5]}, x = o.f[6];
    const double fr = f64_of(ldr<T>(a.fre, x)), fi = f64_of(ldr<T>(a.fim, x));
    uint64_t frb, fib;
    memcpy(&frb, &fr, 8);
    memcpy(&fib, &fi, 8);
    const bool as_is = direct && frb == 0x3ff0000000000000ull && fib == 0;
    for (size_t i = i0; i < i1; i++) {
        double in[PACUDA_MAX_IN], out[PACUDA_MAX_OUT];
        for (uint32_t k = 0; k < n_in; k++) {
            const uint32_t c = o.bind[4 + k], kind = c & 7, which = (c >> 3) & 1, im = (c >> 4) & 1, n = c >> 5;
            double v;
            if (kind == 0 || kind == 2) {
                const Cx<R> z = ld_cur<R>(a, ((kind == 0 ? parent[which] : dst) + n) * a.ld + i);
                v = f64_of(im ? z.im : z.re);
            } else if (kind == 1) {
                v = f64_of(ldr<Den>(a.mom, ((size_t)form[which] * 4 + n) * a.ld + i));
            } else if (kind == 3) {
                v = f64_of(ldr<T>(a.sre, n));
            } else {
                v = im ? fi : fr;
            }
            in[k] = v;
        }
        pacuda_kernel(kernel, false, in, out);
        for (uint32_t j = 0; j < pairs; j++) {
            const double re = out[2 * j], imv = out[2 * j + 1];
            const size_t at = (dst + o.bind[4 + n_in + j]) * a.ld + i;
            double sr, si;
            if (as_is) {
                sr = re;
                si = imv;
            } else {
                sr = re * fr - imv * fi;
                si = re * fi + imv * fr;
                if (!(o.flags & F_INIT)) {
                    const Cx<R> d = ld_cur<R>(a, at);
                    sr = f64_of(d.re) + sr;
                    si = f64_of(d.im) + si;
                }
            }
            st_cur<R>(a, at, cx(narrow<R>(sr), narrow<R>(si)));
        }
    }
}

template <class R> Cx<R> ld_cur_z(const Arena &a, size_t at, uint32_t z, int k) {
    return ((z >> k) & 1) ? cx(R(0.0), R(0.0)) : ld_cur<R>(a, at);
}

template <class K>
void run_program(const Op *ops, uint32_t count, const Arena &a, size_t i0, size_t i1) {
    typedef typename K::real R;
    typedef typename K::den Den;
    typedef typename K::amp Amp;
    typedef typename K::acc T;
    typedef Cx<R> C;
    typedef Cx<T> CT;
    size_t npts = i1 - i0;
    // the destination accumulators of a run of rows (F_FIRST .. F_LAST), per point
    std::vector<CT> acc(6 * npts);
    for (uint32_t n = 0; n < count; n++) {
        const Op &o = ops[n];
        if (o.kind == K_CONTRIB) {
            int c0, c1, cd;
            contribution_shape(o.op, &c0, &c1, &cd);
            uint32_t slot = o.f[1], p0 = o.f[2], p1 = o.f[3], dst = o.f[4], f0 = o.f[5], f1 = o.f[6], x = o.f[7];
            CT scale = mul(cx(ldr<T>(a.sre, slot), ldr<T>(a.sim, slot)), cx(ldr<T>(a.fre, x), ldr<T>(a.fim, x)));
            for (size_t i = i0; i < i1; i++) {
                CT l[6], r[4], v[6];
                T q0[4], q1[4];
                for (int k = 0; k < c0; k++) l[k] = cvt<T>(ld_field<R>(a, o, 2, k, o.z[0], k, i));
                for (int k = 0; k < c1; k++) r[k] = cvt<T>(ld_field<R>(a, o, 3, k, o.z[1], k, i));
                if (o.op == 6)
                    for (int k = 0; k < 4; k++) {
                        q0[k] = ld_form<T, Den>(a.mom, (f0 * 4 + k) * a.ld + i);
                        q1[k] = ld_form<T, Den>(a.mom, (f1 * 4 + k) * a.ld + i);
                    }
                contribution_values(o.op, l, r, q0, q1, v);
                CT *ac = &acc[6 * (i - i0)];
                for (int k = 0; k < cd; k++) {
                    CT c = mul(v[k], scale);
                    if (!(o.flags & F_FIRST)) ac[k] = add(ac[k], c);
                    else if (o.flags & F_INIT) ac[k] = c;
                    else ac[k] = add(cvt<T>(ld_field<R>(a, o, 4, k, o.z[2], k, i)), c);
                }
                if ((o.flags & F_LAST) && !(o.flags & F_HOLD))
                    for (int k = 0; k < cd; k++)
                        st_field<R>(a, o, 4, k, o.z[3], k, cvt<R>(ac[k]), i);
            }
        } else if (o.kind == K_CLEAR) {
            for (uint32_t k = 0; k < o.f[1]; k++)
                for (size_t i = i0; i < i1; i++) st_field<R>(a, o, 0, k, 0, 0, cx(R(0.0), R(0.0)), i);
        } else if (o.kind == K_PROP) {
            // (in T; an input taken from the accumulators is first rounded to R, the value a
            // stored current would have, so forwarding never changes a result)
            int m = o.op == 9 ? 4 : 2;
            uint32_t base = o.f[0], form = o.f[1], x = o.f[2];
            CT rf = cx(ldr<T>(a.fre, x), ldr<T>(a.fim, x));
            for (size_t i = i0; i < i1; i++) {
                T p[4];
                Den q[4];
                for (int k = 0; k < 4; k++) {
                    size_t at = (form * 4 + k) * a.ld + i;
                    q[k] = ldr<Den>(a.mom, at);
                    p[k] = ld_form<T, Den>(a.mom, at);
                }
                T reciprocal = ToReal<T, Den>::of(reciprocal_minkowski_norm(q[0], q[1], q[2], q[3]));
                CT c[4], v[4];
                CT *ac = &acc[6 * (i - i0)];
                for (int k = 0; k < m; k++)
                    c[k] = cvt<T>((o.flags & F_FROMACC) ? cvt<R>(ac[k]) : ld_field<R>(a, o, 0, k, o.z[0], k, i));
                propagator_values(o.op, c, p, reciprocal, rf, v);
                for (int k = 0; k < m; k++) {
                    if (o.flags & F_HOLD) ac[k] = v[k];
                    else st_field<R>(a, o, 0, k, o.z[3], k, cvt<R>(v[k]), i);
                }
            }
        } else if (o.kind == K_IDENTITY) {
            T fr = ldr<T>(a.fre, o.f[2]), fi = ldr<T>(a.fim, o.f[2]);
            for (uint32_t k = 0; k < o.f[1]; k++)
                for (size_t i = i0; i < i1; i++) {
                    CT *ac = &acc[6 * (i - i0)];
                    const int b = k < 16 ? (int)k : 31;  // (no mask bit beyond 16 components)
                    CT v = identity_value(
                        fr, fi, cvt<T>((o.flags & F_FROMACC) ? cvt<R>(ac[k]) : ld_field<R>(a, o, 0, k, o.z[0], b, i)));
                    if (o.flags & F_HOLD) ac[k] = v;
                    else st_field<R>(a, o, 0, k, o.z[3], b, cvt<R>(v), i);
                }
        } else if (o.kind == K_CLOSURE) {
            uint32_t p0 = o.f[0], p1 = o.f[1], amp = o.f[2], x = o.f[3], first = o.f[4], cnt = o.f[5];
            T rfr = ldr<T>(a.fre, x), rfi = ldr<T>(a.fim, x);
            for (size_t i = i0; i < i1; i++) {
                size_t at = amp * a.ld + i;
                Amp dr = ldr<Amp>(a.are, at), di = ldr<Amp>(a.aim, at);
                for (uint32_t k = 0; k < cnt; k++) {
                    const int b = k < 8 ? (int)k : 31;
                    CT v = closure_increment(cvt<T>(ld_field<R>(a, o, 0, k, o.z[0], b, i)),
                                             cvt<T>(ld_field<R>(a, o, 1, k, o.z[1], b, i)),
                                             ldr<T>(a.fre, first + k), ldr<T>(a.fim, first + k), rfr, rfi);
                    dr = dr + widen(v.re);
                    di = di + widen(v.im);
                }
                str(a.are, at, dr);
                str(a.aim, at, di);
            }
        } else if (o.kind == K_SOURCE) {  // a source row made again (the library's f64 programs)
            run_sources<R>(&o, 1, a, i0, i1);
        } else if (o.kind == K_SYMJIT) {
            symjit_row<K>(o, a, i0, i1);
        }
    }
}

// ---- conversions (runner/recurrence_runner.cu:840-880) ------------------------------------

void to_df64(double *p, size_t count) {
    for (size_t k = 0; k < count; k++) str(p, k, narrow<D>(p[k]));
}
void to_df64_planes(double *p, size_t planes, size_t stride, size_t points) {
    for (size_t q = 0; q < planes; q++)
        for (size_t i = 0; i < points; i++) {
            size_t at = q * stride + i;
            str(p, at, narrow<D>(p[at]));
        }
}
void to_f64_planes(double *p, size_t planes, size_t stride, size_t points) {
    for (size_t q = 0; q < planes; q++)
        for (size_t i = 0; i < points; i++) {
            size_t at = q * stride + i;
            D v = ldr<D>(p, at);
            p[at] = (double)v.hi + (double)v.lo;
        }
}

// ---- the v1 call on host memory (runner/recurrence_runner.cu:886-1487) --------------------

enum { KERNEL_DF64 = 0, KERNEL_F64 = 1, KERNEL_MIXED = 2, KERNEL_MIXED64 = 3 };
// the library's fused path also takes its compensated mixed policies (KComp, KComp64)
enum { KERNEL_COMPENSATED = 4, KERNEL_COMPENSATED64 = 5 };
struct Options {
    int kernel = KERNEL_DF64;
    bool df64 = true, df64_forms = true;
    int block = 128;
    bool sort = true;
};

struct Prepared {
    std::vector<Op> ops;
    uint32_t sources = 0;
};

struct Dev {
    std::vector<double> cre, cim, are, aim, mom, small;
    std::vector<float> fcre, fcim;
    bool kept = false, kept_df64 = false;
    size_t kept_stride = 0, kept_points = 0, kept_planes = 0;
    std::map<uint64_t, Prepared> programs;
    Prepared scratch;
    Options options;
    bool runs_started = false;
};
Dev dev;

// a grown buffer loses its contents (filled with NaN here, so no test can rely on them)
template <class T> void grow(std::vector<T> &v, size_t n) {
    if (n <= v.size()) return;
    v.assign(n, (T)NAN);
}

bool copy_runs(double *dre, double *dim, double *hre, double *him, const uint32_t *runs, uint32_t count,
               uint32_t planes, size_t stride, size_t points, bool to_device) {
    for (uint32_t r = 0; r < count; r++) {
        uint32_t first = runs[2 * r], n = runs[2 * r + 1];
        if (first > planes || n > planes - first) return false;
        for (uint32_t q = first; q < first + n; q++) {
            size_t off = (size_t)q * stride;
            if (to_device) {
                memcpy(dre + off, hre + off, points * 8);
                memcpy(dim + off, him + off, points * 8);
            } else {
                memcpy(hre + off, dre + off, points * 8);
                memcpy(him + off, dim + off, points * 8);
            }
        }
    }
    return true;
}

void convert_runs(bool to_df, double *re, double *im, const uint32_t *runs, uint32_t count, size_t stride, size_t points) {
    for (uint32_t r = 0; r < count; r++) {
        size_t off = (size_t)runs[2 * r] * stride, n = runs[2 * r + 1];
        if (to_df) { to_df64_planes(re + off, n, stride, points); to_df64_planes(im + off, n, stride, points); }
        else { to_f64_planes(re + off, n, stride, points); to_f64_planes(im + off, n, stride, points); }
    }
}

void convert_current_runs(bool to_f32, const uint32_t *runs, uint32_t count, size_t stride, size_t points) {
    for (uint32_t r = 0; r < count; r++) {
        size_t off = (size_t)runs[2 * r] * stride, n = runs[2 * r + 1];
        for (size_t q = 0; q < n; q++)
            for (size_t i = 0; i < points; i++) {
                size_t at = off + q * stride + i;
                if (to_f32) { dev.fcre[at] = (float)dev.cre[at]; dev.fcim[at] = (float)dev.cim[at]; }
                else { dev.cre[at] = (double)dev.fcre[at]; dev.cim[at] = (double)dev.fcim[at]; }
            }
    }
}

const Prepared *prepare(const pacman_recurrence_call *call) {
    if (call->words_key) {
        auto it = dev.programs.find(call->words_key);
        if (it != dev.programs.end()) return &it->second;
    }
    std::vector<Row> rows = decode(call->words, call->word_count, call->width);
    if (dev.options.sort) sort_contribution_runs(rows);
    Prepared p;
    if (!program(rows, p.ops, &p.sources)) return nullptr;
    if (call->words_key) return &(dev.programs[call->words_key] = p);
    dev.scratch = p;
    return &dev.scratch;
}

template <class K> void launch_program(const Prepared &p, const Arena &a, size_t points) {
    run_program<K>(p.ops.data() + p.sources, (uint32_t)(p.ops.size() - p.sources), a, 0, points);
}

int run_locked(const pacman_recurrence_call *in) {
    dev.runs_started = true;
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
    dev.kept = false;
    bool keep = call->flags & PACMAN_RECURRENCE_KEEP_AMPLITUDES;
    size_t stride = call->stride, points = call->points;
    size_t current_doubles = (size_t)call->current_planes * stride;
    size_t amplitude_doubles = (size_t)call->amplitude_planes * stride;
    grow(dev.cre, current_doubles); grow(dev.cim, current_doubles);
    grow(dev.are, amplitude_doubles); grow(dev.aim, amplitude_doubles);
    const Options &opt = dev.options;
    bool mixed = opt.kernel == KERNEL_MIXED || opt.kernel == KERNEL_MIXED64;
    if (mixed) { grow(dev.fcre, current_doubles); grow(dev.fcim, current_doubles); }
    size_t momentum_doubles = (size_t)call->momentum_planes * stride;
    size_t small = 2 * ((size_t)call->scale_count + call->factor_count);
    grow(dev.mom, momentum_doubles ? momentum_doubles : 1);
    grow(dev.small, small ? small : 1);
    double *dsre = dev.small.data(), *dsim = dsre + call->scale_count;
    double *dfre = dsim + call->scale_count, *dfim = dfre + call->factor_count;
    bool forms_on_device = call->flags & PACMAN_RECURRENCE_FORMS_ON_DEVICE;
    if (forms_on_device) {
        uint32_t form_count = call->momentum_planes / 4;
        if (!call->external_momenta || !call->form_terms || (call->term_count && (!call->term_slots || !call->term_coefficients)))
            return PACMAN_RECURRENCE_INVALID;
        for (uint32_t f = 0; f < form_count; f++)
            if (call->form_terms[2 * f] > call->term_count || call->form_terms[2 * f + 1] > call->term_count - call->form_terms[2 * f])
                return PACMAN_RECURRENCE_INVALID;
        for (uint32_t t = 0; t < call->term_count; t++)
            if (call->term_slots[t] >= call->external_count) return PACMAN_RECURRENCE_INVALID;
        if (form_count)
            fill_forms(call->external_momenta, points, call->form_terms, form_count, call->term_slots,
                       call->term_coefficients, dev.mom.data(), stride, 0, points, 0);
    }
    if (momentum_doubles && !forms_on_device)
        for (uint32_t q = 0; q < call->momentum_planes; q++) memcpy(&dev.mom[q * stride], call->momenta + q * stride, points * 8);
    if (call->scale_count) {
        memcpy(dsre, call->scale_re, call->scale_count * 8);
        memcpy(dsim, call->scale_im, call->scale_count * 8);
    }
    if (call->factor_count) {
        memcpy(dfre, call->factor_re, call->factor_count * 8);
        memcpy(dfim, call->factor_im, call->factor_count * 8);
    }
    if (!copy_runs(dev.cre.data(), dev.cim.data(), call->current_re, call->current_im, call->upload_current,
                   call->upload_current_runs, call->current_planes, stride, points, true) ||
        !copy_runs(dev.are.data(), dev.aim.data(), call->amplitude_re, call->amplitude_im, call->upload_amplitude,
                   call->upload_amplitude_runs, call->amplitude_planes, stride, points, true))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    for (uint32_t r = 0; r < call->zero_amplitude_runs; r++) {
        uint32_t first = call->zero_amplitude[2 * r], n = call->zero_amplitude[2 * r + 1];
        if (first > call->amplitude_planes || n > call->amplitude_planes - first) return PACMAN_RECURRENCE_INVALID;
        for (uint32_t q = first; q < first + n; q++) {
            memset(&dev.are[(size_t)q * stride], 0, points * 8);
            memset(&dev.aim[(size_t)q * stride], 0, points * 8);
        }
    }
    Arena a;
    a.cre = dev.cre.data(); a.cim = dev.cim.data(); a.are = dev.are.data(); a.aim = dev.aim.data();
    a.fcre = dev.fcre.data(); a.fcim = dev.fcim.data();
    a.mom = dev.mom.data(); a.sre = dsre; a.sim = dsim; a.fre = dfre; a.fim = dfim; a.ld = stride;
    const Prepared *prepared = prepare(call);
    if (!prepared) return PACMAN_RECURRENCE_DEVICE_ERROR;
    // the sources first, from the f64 forms, scales and factors
    if (prepared->sources) {
        if (opt.kernel == KERNEL_F64) run_sources<double>(prepared->ops.data(), prepared->sources, a, 0, points);
        else if (mixed) run_sources<float>(prepared->ops.data(), prepared->sources, a, 0, points);
        else run_sources<D>(prepared->ops.data(), prepared->sources, a, 0, points);
    }
    if (mixed) convert_current_runs(true, call->upload_current, call->upload_current_runs, stride, points);
    if (opt.df64) {
        if (!mixed) convert_runs(true, dev.cre.data(), dev.cim.data(), call->upload_current, call->upload_current_runs, stride, points);
        convert_runs(true, dev.are.data(), dev.aim.data(), call->upload_amplitude, call->upload_amplitude_runs, stride, points);
        if (opt.df64_forms) to_df64_planes(dev.mom.data(), call->momentum_planes, stride, points);
        to_df64(dev.small.data(), small);
    }
    switch (opt.kernel) {
    case KERNEL_F64: launch_program<KF64>(*prepared, a, points); break;
    case KERNEL_MIXED: launch_program<KMixed>(*prepared, a, points); break;
    case KERNEL_MIXED64: launch_program<KMixed64>(*prepared, a, points); break;
    default: launch_program<KDF64>(*prepared, a, points); break;
    }
    uint32_t amplitude_runs = keep ? 0 : call->download_amplitude_runs;
    if (opt.df64) {
        convert_runs(false, dev.are.data(), dev.aim.data(), call->download_amplitude, amplitude_runs, stride, points);
        if (mixed) convert_current_runs(false, call->download_current, call->download_current_runs, stride, points);
        else convert_runs(false, dev.cre.data(), dev.cim.data(), call->download_current, call->download_current_runs, stride, points);
    }
    if (!copy_runs(dev.are.data(), dev.aim.data(), call->amplitude_re, call->amplitude_im, call->download_amplitude,
                   amplitude_runs, call->amplitude_planes, stride, points, false) ||
        !copy_runs(dev.cre.data(), dev.cim.data(), call->current_re, call->current_im, call->download_current,
                   call->download_current_runs, call->current_planes, stride, points, false))
        return PACMAN_RECURRENCE_DEVICE_ERROR;
    if (opt.df64) {
        convert_runs(true, dev.are.data(), dev.aim.data(), call->download_amplitude, amplitude_runs, stride, points);
        if (!mixed) convert_runs(true, dev.cre.data(), dev.cim.data(), call->download_current, call->download_current_runs, stride, points);
    }
    if (keep) {
        dev.kept = true;
        dev.kept_df64 = opt.df64;
        dev.kept_stride = stride;
        dev.kept_points = points;
        dev.kept_planes = call->amplitude_planes;
    }
    return PACMAN_RECURRENCE_OK;
}

// ---- reductions (runner/recurrence_runner.cu:1489-1723) -----------------------------------

inline double amp_value(const double *p, size_t at, bool df) {
    if (!df) return p[at];
    D v = ldr<D>(p, at);
    return (double)v.hi + (double)v.lo;
}

inline double reduce_terms(const uint32_t *planes, const double *weights, uint32_t first, uint32_t end, size_t i,
                           double acc) {
    for (uint32_t k = first; k < end; k++) {
        size_t at = (size_t)planes[k] * dev.kept_stride + i;
        double re = amp_value(dev.are.data(), at, dev.kept_df64), im = amp_value(dev.aim.data(), at, dev.kept_df64);
        acc = acc + weights[k] * fma(re, re, im * im);
    }
    return acc;
}

int reduce_resolved_locked(const pacman_recurrence_reduce_resolved *r) {
    if (!r || r->abi_version != PACMAN_RECURRENCE_ABI || r->struct_size != sizeof(pacman_recurrence_reduce_resolved) ||
        !r->offsets || !r->output || r->outputs == 0 || r->outputs > 65535)
        return PACMAN_RECURRENCE_INVALID;
    if (!dev.kept || r->points != dev.kept_points) return PACMAN_RECURRENCE_INVALID;
    uint32_t count = r->offsets[r->outputs];
    if (r->offsets[0] != 0 || (count && (!r->planes || !r->weights))) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t o = 0; o < r->outputs; o++)
        if (r->offsets[o] > r->offsets[o + 1]) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t k = 0; k < count; k++)
        if (r->planes[k] >= dev.kept_planes) return PACMAN_RECURRENCE_INVALID;
    if (r->points == 0) return PACMAN_RECURRENCE_OK;
    bool zero = r->flags & PACMAN_RECURRENCE_OUTPUT_ZERO;
    for (uint32_t o = 0; o < r->outputs; o++)
        for (size_t i = 0; i < r->points; i++) {
            size_t out = o * (size_t)r->output_stride + i * (size_t)r->point_stride;
            double acc = zero ? 0.0 : r->output[out];
            r->output[out] = reduce_terms(r->planes, r->weights, r->offsets[o], r->offsets[o + 1], i, acc);
        }
    return PACMAN_RECURRENCE_OK;
}

double sample_value(const pacman_recurrence_reduce_sample *r, uint32_t o, size_t i, bool factor, double s) {
    double acc = reduce_terms(r->planes, r->weights, r->offsets[o], r->offsets[o + 1], i, 0.0);
    if (factor) acc = (acc * r->numerator) / r->denominator;
    if (s != 1.0) acc = acc * s;
    return acc;
}

int reduce_sample_locked(const pacman_recurrence_reduce_sample *r) {
    if (!r || r->abi_version != PACMAN_RECURRENCE_ABI || r->struct_size != sizeof(pacman_recurrence_reduce_sample) ||
        !r->offsets || !r->helicity || !r->random || !r->total || !r->chosen || r->outputs == 0)
        return PACMAN_RECURRENCE_INVALID;
    if (!dev.kept || r->points != dev.kept_points) return PACMAN_RECURRENCE_INVALID;
    uint32_t count = r->offsets[r->outputs];
    if (r->offsets[0] != 0 || (count && (!r->planes || !r->weights))) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t o = 0; o < r->outputs; o++)
        if (r->offsets[o] > r->offsets[o + 1]) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t k = 0; k < count; k++)
        if (r->planes[k] >= dev.kept_planes) return PACMAN_RECURRENCE_INVALID;
    if (r->points == 0) return PACMAN_RECURRENCE_OK;
    bool factor = r->numerator != r->denominator;
    for (size_t i = 0; i < r->points; i++) {
        double s = r->scale ? r->scale[i] : 1.0;
        double t = 0.0;
        for (uint32_t o = 0; o < r->outputs; o++) t = t + sample_value(r, o, i, factor, s);
        double target = r->random[i] * t, cumulative = 0.0;
        uint32_t pick = 0;
        for (uint32_t o = 0; o < r->outputs; o++) {
            double v = sample_value(r, o, i, factor, s);
            if (v <= 0.0) continue;
            pick = r->helicity[o];
            cumulative = cumulative + v;
            if (cumulative > target) break;
        }
        r->total[i] = t;
        r->chosen[i] = pick;
    }
    return PACMAN_RECURRENCE_OK;
}

int reduce_locked(const pacman_recurrence_reduce *r) {
    if (!r || r->abi_version != PACMAN_RECURRENCE_ABI || r->struct_size != sizeof(pacman_recurrence_reduce) ||
        (r->count && (!r->planes || !r->weights)) || !r->output)
        return PACMAN_RECURRENCE_INVALID;
    if (!dev.kept || r->points != dev.kept_points) return PACMAN_RECURRENCE_INVALID;
    for (uint32_t k = 0; k < r->count; k++)
        if (r->planes[k] >= dev.kept_planes) return PACMAN_RECURRENCE_INVALID;
    for (size_t i = 0; i < r->points; i++) r->output[i] = reduce_terms(r->planes, r->weights, 0, r->count, i, r->output[i]);
    return PACMAN_RECURRENCE_OK;
}

std::mutex runner_lock;
thread_local bool holding = false;
void acquire() { if (!holding) { runner_lock.lock(); holding = true; } }
void release() { if (holding) { holding = false; runner_lock.unlock(); } }

}  // namespace

extern "C" {

int refcpu_v1_run(const pacman_recurrence_call *in) {
    acquire();
    int status = run_locked(in);
    if (!(status == PACMAN_RECURRENCE_OK && dev.kept)) release();
    return status;
}

int refcpu_v1_configure(const pacman_recurrence_options *o) {
    if (!o || o->abi_version != PACMAN_RECURRENCE_ABI || o->struct_size < sizeof(pacman_recurrence_options))
        return PACMAN_RECURRENCE_INVALID;
    if (o->kernel > PACMAN_RECURRENCE_KERNEL_MIXED64 || (o->block && (o->block < 32 || o->block > 1024)))
        return PACMAN_RECURRENCE_INVALID;
    acquire();
    int status = PACMAN_RECURRENCE_OK;
    Options n;
    n.kernel = (int)o->kernel;
    n.df64 = n.kernel != KERNEL_F64;
    n.df64_forms = n.kernel == KERNEL_DF64 || n.kernel == KERNEL_MIXED;
    n.block = o->block ? (int)o->block : 128;
    n.sort = !o->no_sort;
    bool same = n.kernel == dev.options.kernel && n.block == dev.options.block && n.sort == dev.options.sort;
    if (dev.runs_started && !same) status = PACMAN_RECURRENCE_INVALID;
    else dev.options = n;
    release();
    return status;
}

// Test hook: forget every device state and the configuration (a fresh process).
void refcpu_reset(void) {
    acquire();
    dev = Dev();
    release();
}

int refcpu_v1_reduce(const pacman_recurrence_reduce *r) {
    acquire();
    int s = reduce_locked(r);
    release();
    return s;
}
int refcpu_v1_reduce_resolved(const pacman_recurrence_reduce_resolved *r) {
    acquire();
    int s = reduce_resolved_locked(r);
    release();
    return s;
}
int refcpu_v1_reduce_sample(const pacman_recurrence_reduce_sample *r) {
    acquire();
    int s = reduce_sample_locked(r);
    release();
    return s;
}
uint32_t refcpu_v1_capabilities(void) {
    return PACMAN_RECURRENCE_CAP_SOURCES | PACMAN_RECURRENCE_CAP_FORMS |
           (PACUDA_KERNEL_COUNT > 0 ? PACMAN_RECURRENCE_CAP_SYMJIT : 0u);
}
int refcpu_v1_symjit_kernel(const char *sha256_hex) {
    if (!sha256_hex) return -1;
    for (int k = 0; k < PACUDA_KERNEL_COUNT; k++)
        if (!strcasecmp(pacuda_kernel_sha256[k], sha256_hex)) return k;
    return -1;
}
void *refcpu_v1_host_alloc(size_t bytes) { return malloc(bytes ? bytes : 1); }
void refcpu_v1_host_free(void *p) { free(p); }

// Test hook: generated kernel k on `in` (PACUDA_MAX_IN doubles) into `out` (PACUDA_MAX_OUT).
int refcpu_plane_kernel(uint32_t k, const double *in, double *out) {
    if (k >= PACUDA_KERNEL_COUNT) return PACMAN_RECURRENCE_INVALID;
    pacuda_kernel(k, false, in, out);
    return PACMAN_RECURRENCE_OK;
}

// Test hook: the momentum forms of `points` points from external momenta ([(slot*4+k)*points+i])
// into mom ([(f*4+k)*points+i]), as FORMS_ON_DEVICE computes them.
int refcpu_fill_forms(const double *external, uint32_t points, const uint32_t *form_terms, uint32_t form_count,
                      const uint32_t *slots, const double *coefficients, double *mom) {
    fill_forms(external, points, form_terms, form_count, slots, coefficients, mom, points, 0, points, 0);
    return PACMAN_RECURRENCE_OK;
}

int refcpu_test_sources(const uint32_t *codes, const double *masses, const double *momenta, uint32_t count,
                        double *out_re, double *out_im) {
    for (uint32_t i = 0; i < count; i++) {
        W v[4] = {c64(0.0, 0.0), c64(0.0, 0.0), c64(0.0, 0.0), c64(0.0, 0.0)};
        source_values(codes[i], masses[i], momenta + 4 * i, v);
        for (int k = 0; k < 4; k++) { out_re[4 * i + k] = v[k].re; out_im[4 * i + k] = v[k].im; }
    }
    return PACMAN_RECURRENCE_OK;
}

// The whole plan run per point, as the v1 call with FORMS_ON_DEVICE | KEEP followed by
// reduce_resolved (OUTPUT_ZERO) and the host's combination (umami/umami_pacuda.c:737-747, 883):
// v[o] = reduction of output o from 0; v = (v * num) / den when num != den; total = 0 + v[0] +
// v[1] + ... . `resolved` ([point][output]) and `total` (per point) may each be NULL. Every
// plane the program reads is written by it first (sources, clears, initialising rows); amplitude
// planes start at zero. Points are processed in blocks of `block`, `threads` OpenMP threads.
namespace {

// The whole plan run of an executable program per point (forms, sources, program, reduction,
// helicity sum), in blocks of `block` points, OpenMP over blocks.
int evaluate_ops(uint32_t kernel, const std::vector<Op> &ops, uint32_t sources, uint32_t current_planes,
                 uint32_t amplitude_planes, uint32_t form_count, const double *scale_re, const double *scale_im,
                 uint32_t scale_count, const double *factor_re, const double *factor_im, uint32_t factor_count,
                 const uint32_t *form_terms, const uint32_t *term_slots, const double *term_coefficients,
                 uint32_t outputs, const uint32_t *offsets, const uint32_t *planes, const double *weights,
                 double numerator, double denominator, const double *external, uint32_t points, double *resolved,
                 double *total, uint32_t block, int threads) {
    bool df = kernel != KERNEL_F64 && kernel != KERNEL_COMPENSATED64,
         df_forms = kernel == KERNEL_DF64 || kernel == KERNEL_MIXED || kernel == KERNEL_COMPENSATED;
    bool mixed = kernel == KERNEL_MIXED || kernel == KERNEL_MIXED64 || kernel == KERNEL_COMPENSATED ||
                 kernel == KERNEL_COMPENSATED64;
    // constants: f64 for the sources, df64 layout for the program when df
    std::vector<double> small64(2 * ((size_t)scale_count + factor_count) + 1);
    double *s_re = small64.data(), *s_im = s_re + scale_count, *f_re = s_im + scale_count, *f_im = f_re + factor_count;
    memcpy(s_re, scale_re, scale_count * 8); memcpy(s_im, scale_im, scale_count * 8);
    memcpy(f_re, factor_re, factor_count * 8); memcpy(f_im, factor_im, factor_count * 8);
    std::vector<double> smalld(small64);
    if (df) to_df64(smalld.data(), smalld.size() - 1);
    bool factor = numerator != denominator;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
#endif
    {
        size_t ld = block;
        std::vector<double> cre(current_planes * ld), cim(current_planes * ld), are(amplitude_planes * ld + 1),
            aim(amplitude_planes * ld + 1), mom((size_t)form_count * 4 * ld + 1);
        std::vector<float> fcre(mixed ? current_planes * ld : 1), fcim(mixed ? current_planes * ld : 1);
        size_t nblocks = (points + block - 1) / block;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (size_t b = 0; b < nblocks; b++) {
            size_t start = b * block, n = std::min<size_t>(block, points - start);
            fill_forms(external, points, form_terms, form_count, term_slots, term_coefficients, mom.data(), ld, 0, n, start);
            Arena a;
            a.cre = cre.data(); a.cim = cim.data(); a.are = are.data(); a.aim = aim.data();
            a.fcre = fcre.data(); a.fcim = fcim.data(); a.mom = mom.data(); a.ld = ld;
            a.sre = s_re; a.sim = s_im; a.fre = f_re; a.fim = f_im;
            std::fill(are.begin(), are.end(), 0.0);
            std::fill(aim.begin(), aim.end(), 0.0);
            if (kernel == KERNEL_F64) run_sources<double>(ops.data(), sources, a, 0, n);
            else if (mixed) run_sources<float>(ops.data(), sources, a, 0, n);
            else run_sources<D>(ops.data(), sources, a, 0, n);
            if (df_forms) to_df64_planes(mom.data(), (size_t)form_count * 4, ld, n);
            double *sd_re = smalld.data(), *sd_im = sd_re + scale_count, *fd_re = sd_im + scale_count, *fd_im = fd_re + factor_count;
            a.sre = sd_re; a.sim = sd_im; a.fre = fd_re; a.fim = fd_im;
            const Op *prog = ops.data() + sources;
            uint32_t count = (uint32_t)(ops.size() - sources);
            switch (kernel) {
            case KERNEL_F64: run_program<KF64>(prog, count, a, 0, n); break;
            case KERNEL_MIXED: run_program<KMixed>(prog, count, a, 0, n); break;
            case KERNEL_MIXED64: run_program<KMixed64>(prog, count, a, 0, n); break;
            case KERNEL_COMPENSATED: run_program<KComp>(prog, count, a, 0, n); break;
            case KERNEL_COMPENSATED64: run_program<KComp64>(prog, count, a, 0, n); break;
            default: run_program<KDF64>(prog, count, a, 0, n); break;
            }
            for (size_t i = 0; i < n; i++) {
                double t = 0.0;
                for (uint32_t o = 0; o < outputs; o++) {
                    double acc = 0.0;
                    for (uint32_t k = offsets[o]; k < offsets[o + 1]; k++) {
                        size_t at = (size_t)planes[k] * ld + i;
                        double re = amp_value(are.data(), at, df), im = amp_value(aim.data(), at, df);
                        acc = acc + weights[k] * fma(re, re, im * im);
                    }
                    if (resolved) resolved[(start + i) * outputs + o] = acc;
                    double v = acc;
                    if (factor) v = (v * numerator) / denominator;
                    t = t + v;
                }
                if (total) total[start + i] = t;
            }
        }
    }
    return PACMAN_RECURRENCE_OK;
}

}  // namespace

int refcpu_evaluate(uint32_t kernel, const uint64_t *words, uint32_t word_count, const uint8_t *width,
                    uint32_t current_planes, uint32_t amplitude_planes, uint32_t form_count,
                    const double *scale_re, const double *scale_im, uint32_t scale_count,
                    const double *factor_re, const double *factor_im, uint32_t factor_count,
                    uint32_t external_count, const uint32_t *form_terms, const uint32_t *term_slots,
                    const double *term_coefficients, uint32_t outputs, const uint32_t *offsets,
                    const uint32_t *planes, const double *weights, double numerator, double denominator,
                    const double *external, uint32_t points, double *resolved, double *total,
                    uint32_t block, int threads) {
    (void)external_count;
    if (kernel > KERNEL_COMPENSATED64 || block == 0) return PACMAN_RECURRENCE_INVALID;
    std::vector<Row> rows = decode(words, word_count, width);
    sort_contribution_runs(rows);
    std::vector<Op> ops;
    uint32_t sources = 0;
    if (!program(rows, ops, &sources)) return PACMAN_RECURRENCE_DEVICE_ERROR;
    return evaluate_ops(kernel, ops, sources, current_planes, amplitude_planes, form_count, scale_re, scale_im,
                        scale_count, factor_re, factor_im, factor_count, form_terms, term_slots, term_coefficients,
                        outputs, offsets, planes, weights, numerator, denominator, external, points, resolved, total,
                        block, threads);
}

// The same plan run from an instruction stream of the library's fused path (pacuda_plan_program:
// 8 u32 per instruction, w[0] = kind | op << 8 | flags << 16, the operands in w[1..7]); current
// plane indices >= 2^24 (on-chip slots: LDS, then the register tier) are mapped after the slot's
// `slot_planes` planes, as are the on-chip copies of dual fields (bits 30-31 set: a slot plane in
// the low 24 bits whose copy sits at on-chip slot (plane + bits 24-29) & 63, read and written as
// the op's dual flags say), form fields keep their low 24 bits (the bits above name the register
// slot a fused kernel keeps the form in), symjit rows take their bindings from `bindings` (the
// plan's binding table, pacuda_plan_bindings), and the known-zero masks in spare bits are honoured
// as the kernels honour them in the slot (see Op). Lets the tests check, on the CPU, that the library's
// program passes leave every value unchanged.
int refcpu_evaluate_program(uint32_t kernel, const uint32_t *instr, uint32_t count, uint32_t sources,
                            uint32_t slot_planes, uint32_t lds_planes, uint32_t amplitude_planes, uint32_t form_count,
                            const double *scale_re, const double *scale_im, uint32_t scale_count,
                            const double *factor_re, const double *factor_im, uint32_t factor_count,
                            const uint32_t *form_terms, const uint32_t *term_slots, const double *term_coefficients,
                            uint32_t outputs, const uint32_t *offsets, const uint32_t *planes, const double *weights,
                            double numerator, double denominator, const double *external, uint32_t points,
                            double *resolved, double *total, uint32_t block, int threads, const uint32_t *bindings,
                            uint32_t binding_words) {
    if (kernel > KERNEL_COMPENSATED64 || block == 0 || !instr) return PACMAN_RECURRENCE_INVALID;
    const uint32_t LDS = 1u << 24;
    auto plane = [&](uint32_t p) { return p >= LDS ? slot_planes + (p - LDS) : p; };
    // plane field k of op o from word w: a slot plane, an on-chip slot, or a slot plane with an
    // on-chip copy at ((p + (w >> 24)) & 63)
    auto field = [&](Op &o, int k, uint32_t w) {
        if (w >= LDS && (w & (DUAL_FILL | DUAL_THROUGH))) {
            o.f[k] = w & (LDS - 1);
            o.dual[k] = w & (DUAL_FILL | DUAL_THROUGH);
            o.chip[k] = slot_planes + (((w & (LDS - 1)) + (w >> 24)) & 63);
        } else {
            o.f[k] = plane(w);
        }
    };
    auto form = [](uint32_t f) { return f & 0xffffffu; };
    std::vector<Op> ops(count);
    for (uint32_t n = 0; n < count; n++) {
        const uint32_t *w = instr + 8 * (size_t)n;
        Op &o = ops[n];
        memset(&o, 0, sizeof o);
        o.kind = (int)(w[0] & 0xff);
        o.op = (int)((w[0] >> 8) & 0xff);
        o.flags = (int)(w[0] >> 16);
        switch (o.kind) {
        case K_CONTRIB:
            o.f[0] = (o.flags & F_INIT) ? 1 : 0;
            o.f[1] = w[1] & 0xffff;
            field(o, 2, w[2]); field(o, 3, w[3]); field(o, 4, w[4]);
            o.f[5] = form(w[5]); o.f[6] = form(w[6]); o.f[7] = w[7] & 0xffff;
            o.z[0] = (w[0] >> 22) & 63; o.z[1] = w[0] >> 28; o.z[2] = (w[1] >> 16) & 63; o.z[3] = (w[7] >> 16) & 63;
            break;
        case K_CLEAR: field(o, 0, w[1]); o.f[1] = w[2]; break;
        case K_PROP:
            field(o, 0, w[1]); o.f[1] = form(w[2]); o.f[2] = w[3];
            o.z[0] = w[4] & 15; o.z[3] = (w[4] >> 8) & 15;
            break;
        case K_IDENTITY:
            field(o, 0, w[1]); o.f[1] = w[2]; o.f[2] = w[3];
            o.z[0] = w[4] & 0xffff; o.z[3] = w[4] >> 16;
            break;
        case K_CLOSURE:
            field(o, 0, w[1]); field(o, 1, w[2]);
            o.f[2] = w[3]; o.f[3] = w[4]; o.f[4] = w[5]; o.f[5] = w[6];
            o.z[0] = w[7] & 0xff; o.z[1] = (w[7] >> 8) & 0xff;
            break;
        case K_SOURCE:
            field(o, 0, w[1]); o.f[1] = form(w[2]); o.f[2] = w[3]; o.f[3] = w[4]; o.f[4] = w[5];
            o.z[3] = w[6] & 15;
            break;
        case K_SYMJIT: {  // w[1], w[2] parents, w[3] destination, w[4] forms, w[5] factor, w[6] binding offset
            const size_t at = w[6];
            if (!bindings || at + 4 > binding_words) return PACMAN_RECURRENCE_INVALID;
            const size_t words = 4 + (size_t)bindings[at + 1] + bindings[at + 2];
            if (at + words > binding_words || words > (size_t)BIND_WORDS || bindings[at] >= PACUDA_KERNEL_COUNT)
                return PACMAN_RECURRENCE_INVALID;
            o.flags &= F_INIT;
            o.f[1] = plane(w[1]); o.f[2] = plane(w[2]); o.f[3] = plane(w[3]);
            o.f[4] = w[4] & 0xffff; o.f[5] = w[4] >> 16; o.f[6] = w[5];
            memcpy(o.bind, bindings + at, words * sizeof(uint32_t));
            break;
        }
        default: return PACMAN_RECURRENCE_INVALID;
        }
    }
    return evaluate_ops(kernel, ops, sources, slot_planes + lds_planes, amplitude_planes, form_count, scale_re, scale_im,
                        scale_count, factor_re, factor_im, factor_count, form_terms, term_slots, term_coefficients,
                        outputs, offsets, planes, weights, numerator, denominator, external, points, resolved, total,
                        block, threads);
}

#ifdef REFCPU_ABI_NAMES
int pacman_recurrence_v1_run(const pacman_recurrence_call *c) { return refcpu_v1_run(c); }
int pacman_recurrence_v1_configure(const pacman_recurrence_options *o) { return refcpu_v1_configure(o); }
int pacman_recurrence_v1_reduce(const pacman_recurrence_reduce *r) { return refcpu_v1_reduce(r); }
int pacman_recurrence_v1_reduce_resolved(const pacman_recurrence_reduce_resolved *r) { return refcpu_v1_reduce_resolved(r); }
int pacman_recurrence_v1_reduce_sample(const pacman_recurrence_reduce_sample *r) { return refcpu_v1_reduce_sample(r); }
uint32_t pacman_recurrence_v1_capabilities(void) { return refcpu_v1_capabilities(); }
int pacman_recurrence_v1_symjit_kernel(const char *s) { return refcpu_v1_symjit_kernel(s); }
void *pacman_recurrence_v1_host_alloc(size_t b) { return refcpu_v1_host_alloc(b); }
void pacman_recurrence_v1_host_free(void *p) { refcpu_v1_host_free(p); }
int pacman_recurrence_v1_test_sources(const uint32_t *c, const double *m, const double *p, uint32_t n, double *re, double *im) {
    return refcpu_test_sources(c, m, p, n, re, im);
}
#endif

}  // extern "C"
