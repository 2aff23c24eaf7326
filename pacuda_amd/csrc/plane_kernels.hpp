// Generated by pacuda-gen from plane-kernel files; do not edit.
#pragma once
#include <stdint.h>

#ifndef PACUDA_KFN
#if defined(__HIPCC__)
#define PACUDA_KFN __host__ __device__ __attribute__((always_inline)) inline
#else
#define PACUDA_KFN inline
#endif
#endif

PACUDA_KFN double pk_bits(unsigned long long u) { return __builtin_bit_cast(double, u); }

#define PACUDA_KERNEL_COUNT 3
#define PACUDA_MAX_IN 8
#define PACUDA_MAX_OUT 8
#define PACUDA_SIMD_LANES 0
static const char *const pacuda_kernel_sha256[3] = {
    "a12abb3717afe05cca614788c9e4b70208646f423f06b688aeb331c285c96721",  // pk_binding_probe
    "8ef075a08cd8b25a87242747673bfebdf14d148c4d9248b65142b1c7f5d21b02",  // pk_fermion_pair_sigma
    "ecc3beb8bd500acfad0fb8e774f6424aa58f1231f5d361121c26d5afcbc0a283",  // pk_fermion_pair_sigmabar
};

// 8 inputs, 8 outputs
PACUDA_KFN void pk_binding_probe(const double *in, double *out) {
    out[0] = in[0];
    out[1] = in[1];
    out[2] = in[2];
    out[3] = in[3];
    out[4] = in[4];
    out[5] = in[5];
    out[6] = in[6];
    out[7] = in[7];
}

// 8 inputs, 8 outputs
PACUDA_KFN void pk_fermion_pair_sigma(const double *in, double *out) {
    const double v_p00r_a = in[0] * in[4];
    const double v_p00r_b = in[1] * in[5];
    const double v_p00r = v_p00r_a - v_p00r_b;
    const double v_p00i_a = in[0] * in[5];
    const double v_p00i_b = in[1] * in[4];
    const double v_p00i = v_p00i_a + v_p00i_b;
    const double v_p11r_a = in[2] * in[6];
    const double v_p11r_b = in[3] * in[7];
    const double v_p11r = v_p11r_a - v_p11r_b;
    const double v_p11i_a = in[2] * in[7];
    const double v_p11i_b = in[3] * in[6];
    const double v_p11i = v_p11i_a + v_p11i_b;
    const double v_p01r_a = in[0] * in[6];
    const double v_p01r_b = in[1] * in[7];
    const double v_p01r = v_p01r_a - v_p01r_b;
    const double v_p01i_a = in[0] * in[7];
    const double v_p01i_b = in[1] * in[6];
    const double v_p01i = v_p01i_a + v_p01i_b;
    const double v_p10r_a = in[2] * in[4];
    const double v_p10r_b = in[3] * in[5];
    const double v_p10r = v_p10r_a - v_p10r_b;
    const double v_p10i_a = in[2] * in[5];
    const double v_p10i_b = in[3] * in[4];
    const double v_p10i = v_p10i_a + v_p10i_b;
    out[0] = v_p00r + v_p11r;
    out[1] = v_p00i + v_p11i;
    out[2] = v_p01r + v_p10r;
    out[3] = v_p01i + v_p10i;
    const double v_dr = v_p10r - v_p01r;
    const double v_di = v_p10i - v_p01i;
    out[4] = -v_di;
    out[5] = v_dr;
    out[6] = v_p00r - v_p11r;
    out[7] = v_p00i - v_p11i;
}

// 8 inputs, 8 outputs
PACUDA_KFN void pk_fermion_pair_sigmabar(const double *in, double *out) {
    const double v_p00r_a = in[0] * in[4];
    const double v_p00r_b = in[1] * in[5];
    const double v_p00r = v_p00r_a - v_p00r_b;
    const double v_p00i_a = in[0] * in[5];
    const double v_p00i_b = in[1] * in[4];
    const double v_p00i = v_p00i_a + v_p00i_b;
    const double v_p11r_a = in[2] * in[6];
    const double v_p11r_b = in[3] * in[7];
    const double v_p11r = v_p11r_a - v_p11r_b;
    const double v_p11i_a = in[2] * in[7];
    const double v_p11i_b = in[3] * in[6];
    const double v_p11i = v_p11i_a + v_p11i_b;
    const double v_p01r_a = in[0] * in[6];
    const double v_p01r_b = in[1] * in[7];
    const double v_p01r = v_p01r_a - v_p01r_b;
    const double v_p01i_a = in[0] * in[7];
    const double v_p01i_b = in[1] * in[6];
    const double v_p01i = v_p01i_a + v_p01i_b;
    const double v_p10r_a = in[2] * in[4];
    const double v_p10r_b = in[3] * in[5];
    const double v_p10r = v_p10r_a - v_p10r_b;
    const double v_p10i_a = in[2] * in[5];
    const double v_p10i_b = in[3] * in[4];
    const double v_p10i = v_p10i_a + v_p10i_b;
    out[0] = v_p00r + v_p11r;
    out[1] = v_p00i + v_p11i;
    const double v_sr = v_p01r + v_p10r;
    const double v_si = v_p01i + v_p10i;
    out[2] = -v_sr;
    out[3] = -v_si;
    const double v_er = v_p01r - v_p10r;
    const double v_ei = v_p01i - v_p10i;
    out[4] = -v_ei;
    out[5] = v_er;
    out[6] = v_p11r - v_p00r;
    out[7] = v_p11i - v_p00i;
}

PACUDA_KFN void pacuda_kernel(uint32_t k, bool simd, const double *in, double *out) {
    (void)simd;
    switch (k) {
    case 0: pk_binding_probe(in, out); break;
    case 1: pk_fermion_pair_sigma(in, out); break;
    case 2: pk_fermion_pair_sigmabar(in, out); break;
    default: (void)in; (void)out; break;
    }
}
