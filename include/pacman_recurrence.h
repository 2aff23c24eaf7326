/* SPDX-License-Identifier: 0BSD */
/*
 * librecurrence.so for MI355X (gfx950): the recurrence runner's C interface.
 *
 * The declarations keep the reference interface byte for byte (struct layouts, field names,
 * constants and entry points of runner/pacman_recurrence.h:11-150), because its callers
 * (umami/umami_pacuda.c, pacuda.py retarget, runner/rectest.c, convert/src/replay.rs) include
 * this header by name or bind these symbols by dlsym. Every entry point below replaces the
 * reference function of the same name (runner/recurrence_runner.cu:1750-1865).
 */
#ifndef PACMAN_RECURRENCE_H
#define PACMAN_RECURRENCE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PACMAN_RECURRENCE_ABI 1

/* indices of the per-plan field widths in `width` (runner/pacman_recurrence.h:13) */
enum { W_P, W_A, W_F, W_X, W_S, W_K, W_COUNT };

/*
 * One plan run (runner/pacman_recurrence.h:17-51). A plane holds `stride` doubles, one per
 * point; runs are (first plane, plane count) pairs. The device arena outlives the call; only
 * the named runs move between host and device, and only their first `points` values. (This
 * library runs a keyed call that uploads only the external momenta and downloads no currents on
 * its fused kernels, with the same amplitudes; the arena's current planes then keep what they
 * held: pacuda_v1_fused_runs in pacuda_amd.h.)
 * struct_size: 248 (all fields), 208 (ends before external_momenta) or 192 (ends before
 * zero_amplitude); fields past struct_size read as zero.
 */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    const uint64_t *words;
    uint64_t words_key;            /* nonzero: these words never change under this key */
    uint32_t word_count;
    uint8_t width[8];              /* W_P..W_K */
    double *current_re, *current_im;
    double *amplitude_re, *amplitude_im;
    const double *momenta;         /* momentum_planes planes, 4 per form (E, px, py, pz) */
    const double *scale_re, *scale_im;
    const double *factor_re, *factor_im;
    const uint32_t *upload_current, *upload_amplitude, *download_amplitude, *download_current;
    uint32_t current_planes, amplitude_planes, momentum_planes, scale_count, factor_count;
    uint32_t upload_current_runs, upload_amplitude_runs, download_amplitude_runs,
        download_current_runs;
    uint32_t stride, points;
    /* struct_size 192 ends here */
    const uint32_t *zero_amplitude;  /* amplitude runs cleared on the device instead of uploaded */
    uint32_t zero_amplitude_runs;
    uint32_t flags;                  /* PACMAN_RECURRENCE_KEEP_AMPLITUDES | _FORMS_ON_DEVICE */
    /* struct_size 208 ends here. With PACMAN_RECURRENCE_FORMS_ON_DEVICE the momentum forms are
     * built on the device: form f, component k = 0.0 + c0 * p(s0, k) + c1 * p(s1, k) + ... over
     * terms form_terms[2f] .. form_terms[2f] + form_terms[2f+1] - 1 of (term_slots,
     * term_coefficients), with p(s, k) = external_momenta[(s * 4 + k) * points + i]; `momenta`
     * is then not read. */
    const double *external_momenta;
    uint32_t external_count;
    uint32_t term_count;
    const uint32_t *form_terms;      /* (first term, term count) per form */
    const uint32_t *term_slots;
    const double *term_coefficients;
} pacman_recurrence_call;

/* the download_amplitude runs stay on the device for a following reduction */
#define PACMAN_RECURRENCE_KEEP_AMPLITUDES 1u
/* the momentum forms are built on the device from external_momenta */
#define PACMAN_RECURRENCE_FORMS_ON_DEVICE 2u

/*
 * After a KEEP run: output[i] += weights[k] * fma(re, re, im * im) for k = 0 .. count-1 in
 * order, (re, im) = amplitude plane planes[k] at point i (runner/pacman_recurrence.h:53-67).
 */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    const uint32_t *planes;
    const double *weights;
    double *output;
    uint32_t count, points;
} pacman_recurrence_reduce;

enum {
    PACMAN_RECURRENCE_OK = 0,
    PACMAN_RECURRENCE_INVALID = 2,
    PACMAN_RECURRENCE_DEVICE_ERROR = 3,
};

/*
 * After a KEEP run, several outputs (runner/pacman_recurrence.h:77-94): for output o and point i,
 * output[o * output_stride + i * point_stride] += weights[k] * fma(re, re, im * im) over
 * k = offsets[o] .. offsets[o+1]-1 in order. With PACMAN_RECURRENCE_OUTPUT_ZERO the outputs
 * start from zero and are not uploaded. The call ends the kept run.
 */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    const uint32_t *planes;
    const double *weights;
    const uint32_t *offsets;  /* outputs + 1 entries */
    double *output;
    uint32_t outputs, points, output_stride, point_stride;
    uint32_t flags;
} pacman_recurrence_reduce_resolved;
#define PACMAN_RECURRENCE_OUTPUT_ZERO 1u

/*
 * After a KEEP run, helicity sampling per point (runner/pacman_recurrence.h:97-118): v[o] is
 * output o's value from zero, times numerator / denominator when they differ, times scale[i]
 * when scale is given and is not 1; total[i] = 0 + v[0] + v[1] + ...; chosen[i] = helicity[o]
 * of the first output whose running sum of positive v exceeds random[i] * total[i], else of the
 * last output with a positive v, else 0. The call ends the kept run.
 */
typedef struct {
    uint32_t abi_version;
    uint32_t struct_size;
    const uint32_t *planes;
    const double *weights;
    const uint32_t *offsets;   /* outputs + 1 entries */
    const uint32_t *helicity;  /* per output, the helicity `chosen` reports */
    uint32_t outputs, points;
    const double *random;      /* per point, in [0, 1) */
    const double *scale;       /* per point, or NULL (1) */
    double numerator, denominator;
    double *total;             /* per point */
    uint32_t *chosen;          /* per point */
} pacman_recurrence_reduce_sample;

/* runner/recurrence_runner.cu:1750 */
int pacman_recurrence_v1_run(const pacman_recurrence_call *call);
/* runner/recurrence_runner.cu:1853 */
int pacman_recurrence_v1_reduce_resolved(const pacman_recurrence_reduce_resolved *reduce);
/* runner/recurrence_runner.cu:1846 */
int pacman_recurrence_v1_reduce_sample(const pacman_recurrence_reduce_sample *sample);
/* pinned host memory, NULL on failure (runner/recurrence_runner.cu:1836-1844) */
void *pacman_recurrence_v1_host_alloc(size_t bytes);
void pacman_recurrence_v1_host_free(void *p);
/* runner/recurrence_runner.cu:1860 */
int pacman_recurrence_v1_reduce(const pacman_recurrence_reduce *reduce);

/* capability bits: 0 sources (op 12) on the device, 1 forms on the device */
#define PACMAN_RECURRENCE_CAP_SOURCES 1u
#define PACMAN_RECURRENCE_CAP_FORMS 2u

/* The configuration: before the first run; afterwards only the same one again
 * (runner/pacman_recurrence.h:130-145). */
typedef struct {
    uint32_t abi_version;  /* PACMAN_RECURRENCE_ABI */
    uint32_t struct_size;
    uint32_t kernel;       /* PACMAN_RECURRENCE_KERNEL_* */
    uint32_t block;        /* threads per block, 32..1024 (0: 128) */
    uint32_t no_sort;      /* 1: keep the contribution rows in plan order */
    uint32_t stats;        /* 1: print statistics at exit */
} pacman_recurrence_options;
enum {
    PACMAN_RECURRENCE_KERNEL_DF64 = 0,  /* the default */
    PACMAN_RECURRENCE_KERNEL_F64 = 1,
    PACMAN_RECURRENCE_KERNEL_MIXED = 2,
    PACMAN_RECURRENCE_KERNEL_MIXED64 = 3,
};
/* runner/recurrence_runner.cu:1757 */
int pacman_recurrence_v1_configure(const pacman_recurrence_options *options);
/* bit 2: op 13/14 rows with generated symjit kernels */
#define PACMAN_RECURRENCE_CAP_SYMJIT 4u
/* runner/recurrence_runner.cu:1785 */
uint32_t pacman_recurrence_v1_capabilities(void);
/* runner/recurrence_runner.cu:1792: index of the symjit kernel with this SHA-256, or -1 */
int pacman_recurrence_v1_symjit_kernel(const char *sha256_hex);

#ifdef __cplusplus
}
#endif
#endif
