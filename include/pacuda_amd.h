/* SPDX-License-Identifier: 0BSD */
/*
 * librecurrence.so for MI355X: entry points beyond runner/pacman_recurrence.h.
 *
 *  - pacman_recurrence_v1_test_sources: the test-only source entry the reference defines outside
 *    its header (runner/recurrence_runner.cu:1811-1834), used by tests/check_sources.py:95.
 *  - the fused evaluation (capability bit 3, PACUDA_CAP_FUSED): whole plan runs, from device-
 *    resident external momenta to device-resident |M|^2 outputs, of one or many plans in one
 *    persistent launch. No reference counterpart; the reference's nearest path is a
 *    pacman_recurrence_v1_run(KEEP | FORMS_ON_DEVICE) followed by
 *    pacman_recurrence_v1_reduce_resolved(OUTPUT_ZERO) and the caller's helicity sum
 *    (umami/umami_pacuda.c:644-747, 853-892), whose results it reproduces bit for bit.
 *  - device memory helpers, so a caller without a GPU framework can hold inputs in HBM.
 */
#ifndef PACUDA_AMD_H
#define PACUDA_AMD_H
#include <stddef.h>
#include <stdint.h>

#include "pacman_recurrence.h"

#ifdef __cplusplus
extern "C" {
#endif

/* runner/recurrence_runner.cu:1811: the source wavefunctions of `count` (template code, mass,
 * momentum) items on the device; 4 components per item in out_re/out_im (unused ones 0) */
int pacman_recurrence_v1_test_sources(const uint32_t *codes, const double *masses, const double *momenta,
                                      uint32_t count, double *out_re, double *out_im);

/* bit 3 of pacman_recurrence_v1_capabilities(): pacuda_plan_create / pacuda_evaluate */
#define PACUDA_CAP_FUSED 8u

/*
 * The number of pacman_recurrence_v1_run calls of this process that ran on the fused kernels (no
 * reference counterpart; for tests and tuning). A run does when it gives the same results there:
 * a keyed call (words_key nonzero) with PACMAN_RECURRENCE_FORMS_ON_DEVICE, no current or
 * amplitude uploads, no current downloads, a program that writes every current plane it reads,
 * and closures that write only planes of the zero_amplitude runs. It is then one launch of the
 * fused kernels (forms, sources, program) that writes the amplitude planes into the device arena;
 * the arena's current planes are not written. PACUDA_V1_FUSED=0 in the environment keeps every
 * run on the v1 kernels.
 */
uint64_t pacuda_v1_fused_runs(void);

typedef struct pacuda_plan pacuda_plan;

/*
 * A recorded plan and its reduction, as one colour of a .pacuda file holds them
 * (convert/src/format.rs:15-24): the words with their widths, the table sizes, the scales and
 * factors, the momentum forms over the external legs, the reduction of amplitude planes into
 * helicity outputs, and the process group factor. Every plane the program reads must be written
 * by it first (sources, clears, initialising rows); the amplitudes start at zero.
 */
typedef struct {
    uint32_t abi_version;  /* PACMAN_RECURRENCE_ABI */
    uint32_t struct_size;  /* sizeof(pacuda_plan_desc) */
    const uint64_t *words;
    uint32_t word_count;
    uint8_t width[8];
    uint32_t current_planes, amplitude_planes, form_count;
    const double *scale_re, *scale_im;
    uint32_t scale_count;
    const double *factor_re, *factor_im;
    uint32_t factor_count;
    uint32_t external_count;
    uint32_t term_count;
    const uint32_t *form_terms;      /* (first term, term count) per form */
    const uint32_t *term_slots;
    const double *term_coefficients;
    uint32_t outputs;                /* helicity outputs */
    const uint32_t *offsets;         /* outputs + 1 */
    const uint32_t *planes;
    const double *weights;
    double numerator, denominator;   /* group factor: v * num / den when they differ */
    uint32_t no_sort;                /* 1: the contribution rows in plan order */
} pacuda_plan_desc;

/*
 * The instruction stream the fused path runs for a plan (host only: validation, the program and
 * the fused path's passes, no device work). variant 0: the program with arena compaction only,
 * 1: with forwarding through the accumulators, 2 + v: with LDS variant
 * v (10, 13, 20, 26, 40 planes; plane indices >= 2^24 are on-chip slots), 7 + v: LDS variant v
 * with a register tier of 16 planes (whole four-plane currents at slots from the variant's LDS
 * capacity on), 12 + v and 17 + v: the same two for the f64 kernel, whose programs store no
 * source current in the slot but run its source row again (kind 6 after the leading sources)
 * where a read would have loaded it. A plane field with bit 31 or 30 set names slot plane (field & 0xffffff) with an
 * on-chip copy at slot ((field & 0xffffff) + (field >> 24)) & 63: bit 31, a read loads the slot
 * and keeps the copy; bit 30, a write also stores to the slot (reads use the copy). Each
 * instruction is 8 u32: w[0] =
 * kind | op << 8 | flags << 16 (kinds 1 clear, 2 contribution, 3 propagator, 4 identity, 5
 * closure, 6 source; flags 1 init, 2 first, 4 last, 8 hold, 16 from the accumulators). out (may
 * be NULL) receives up to `capacity` instructions; *count their number, *sources the leading
 * source rows, *planes the slot's current planes, *lds_planes the on-chip slots named. For tests
 * and inspection; no reference counterpart.
 */
int pacuda_plan_program(const pacuda_plan_desc *desc, uint32_t variant, uint32_t *out, uint32_t capacity,
                        uint32_t *count, uint32_t *sources, uint32_t *planes, uint32_t *lds_planes);
/*
 * The cooperative program of a plan for cooperative shape `shape` (0: 16 wavefronts per unit, one
 * unit per CU; 1: 8 and one; 2: 8 and two; 3: 4 and four) with 16-byte currents (f64, df64), or
 * shape + 4 for the f32 currents of the mixed and compensated kernels (shapes 0 and 2). serial 0: the wavefronts' streams one after another
 * (starts, may be NULL, receives wavefronts + 1 offsets), each a share of the rows with barrier
 * rows (kind 8) between its phases, current fields at or above 2^24 in LDS, form fields with bit
 * 23 set an LDS plane pair, amplitude fields at or above 2^24 an LDS plane; serial 1: the same rows
 * in one order (phase by phase, wavefront by wavefront) with forms and amplitudes under their plan
 * indices, which pacuda_plan_program's readers run with LDS planes after the slot's. info (may be
 * NULL) receives wavefronts, phases, LDS planes, slot current planes, and the current components
 * per point accessed in the slot and in LDS. Status 2 when the plan has none. For tests and
 * inspection; no reference counterpart.
 */
int pacuda_plan_coop_program(const pacuda_plan_desc *desc, uint32_t shape, uint32_t serial, uint32_t *out,
                             uint32_t capacity, uint32_t *count, uint32_t *info, uint32_t *starts);
/*
 * The binding table a plan's programs index (symjit rows: w[6] is an offset into it; per binding
 * the kernel, inputs, output pairs, direct, then the input codes and the output components): out
 * (may be NULL) receives up to `capacity` words, *count their number. For tests and inspection
 * (the CPU oracle runs the programs of plans with symjit rows with it); no reference counterpart.
 */
int pacuda_plan_bindings(const pacuda_plan_desc *desc, uint32_t *out, uint32_t capacity, uint32_t *count);
/* 0 OK, 2 invalid description, 3 device error */
int pacuda_plan_create(const pacuda_plan_desc *desc, pacuda_plan **plan);
void pacuda_plan_destroy(pacuda_plan *plan);
/* the plan's instruction count after compilation and the scratch doubles one wavefront needs */
int pacuda_plan_info(const pacuda_plan *plan, uint32_t *instructions, uint64_t *slot_doubles);

/*
 * One batch of points of one plan. external: device, p(leg, k) of point i at
 * external[(leg * 4 + k) * ext_stride + i] (the momenta the forms read, crossing applied);
 * resolved: device [point][outputs] (each output from zero, before the group factor) or NULL;
 * total: device, per point, 0 + v[0] + v[1] + ... with v = (r * num) / den when num != den,
 * else r, or NULL.
 */
typedef struct {
    const pacuda_plan *plan;
    const double *external;
    uint32_t points, ext_stride;
    double *resolved;
    double *total;
} pacuda_item;

/*
 * The compensated mixed policy of the fused entry points (no counterpart in the reference; the
 * v1 configure call keeps the reference's four): f32 currents, df64 denominators and amplitude
 * sums as PACMAN_RECURRENCE_KERNEL_MIXED, and the arithmetic of every row (vertices, the sum of a
 * destination's contributions, propagators, identity rows, the closures' products) in df64 from
 * the f32 values, so that a current is rounded to f32 once, when it is stored.
 */
#define PACUDA_KERNEL_COMPENSATED 4u
/*
 * The same with f64 arithmetic (no counterpart in the reference): f32 currents, each rounded to
 * f32 once when stored, and the arithmetic of every row, the forms, constants, denominators and
 * amplitude sums in f64 (native on gfx950, where df64 takes about ten f32 operations for one).
 * Equal to PACUDA_KERNEL_COMPENSATED to df64's accuracy.
 */
#define PACUDA_KERNEL_COMPENSATED64 5u

typedef struct {
    uint32_t abi_version;   /* PACMAN_RECURRENCE_ABI */
    uint32_t struct_size;   /* sizeof(pacuda_launch) */
    uint32_t kernel;        /* PACMAN_RECURRENCE_KERNEL_* or PACUDA_KERNEL_COMPENSATED(64) */
    uint32_t waves_per_cu;  /* persistent wavefronts per compute unit, 0: 12 (f64), 16 (the others) */
    uint32_t block;         /* threads per workgroup, a multiple of 64 up to 256, 0: 256 */
    uint32_t unit_waves;    /* 0: one wavefront per 64-point unit; 16, 8 or 4: cooperative launch (below) */
    /* all three 0: the default launch, cooperative with 8 wavefronts per unit and 16 per CU for the
       f64, mixed, mixed64 and compensated64 kernels when every plan has a cooperative program
       (16 wavefronts per unit when that program would move more than 64 KiB per point through
       the slot on average, as g g > g g g g's f64 one does), else one wavefront per unit at the
       default waves_per_cu and block */
    void *stream;           /* hipStream_t to queue on, or NULL: the library's stream */
    float *kernel_ms;       /* not NULL: synchronise and store the fused kernel's time (HIP events) */
} pacuda_launch;

/*
 * Evaluates every item in one persistent launch: each wavefront takes 64-point units of the
 * items off a device work queue and runs forms, sources, program and reduction for them with its
 * arena in an on-device scratch slot. Results equal the v1 path bit for bit for the same kernel.
 * Cooperative launch (unit_waves 16, 8 or 4): each workgroup of unit_waves wavefronts takes one
 * unit at a time and its wavefronts share it, running the plan's cooperative program (its rows
 * split into phases between workgroup barriers, the unit's currents in the workgroup's LDS and
 * slot); waves_per_cu 0 means 16, and waves_per_cu / unit_waves units share a compute unit's LDS
 * (f64: 16 / 16, 8 / 8, 16 / 8 or 16 / 4; the other kernels 16 / 16 or 16 / 8); block is ignored.
 * The same results bit for bit. Plans with symjit rows run in the 16 / 16 and 16 / 8 shapes only.
 * Asynchronous on the stream unless kernel_ms is set. Calls may go to different streams: the
 * item table, work counter and slots are shared by all calls, so each call first waits on the
 * host until the previous call's launch has ended. 0 OK, 2 invalid, 3 device error.
 */
int pacuda_evaluate(const pacuda_item *items, uint32_t count, const pacuda_launch *launch);

/*
 * pacuda_evaluate without allocation. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The same launch and results as pacuda_evaluate, with the work counter, the
 * item table and the wavefronts' slots in one fixed 256 MiB device array of the library (each GPU
 * holds its own copy until the process ends) instead of buffers the library allocates. Limits:
 * at most 4096 items; as many resident wavefronts as the slots fit (at most the launch's, at least
 * one workgroup: a plan whose slots for one workgroup exceed the array is refused, status 2);
 * kernel_ms must be NULL; stream NULL is the null stream. The call waits for `stream` before it
 * rewrites the item table, so consecutive calls on one stream are ordered; calls that overlap on
 * one GPU from different streams or threads share the array and race.
 */
int pacuda_evaluate_static(const pacuda_item *items, uint32_t count, const pacuda_launch *launch);

/*
 * A sweep: the same batch of items evaluated again and again (a Monte Carlo loop that rewrites
 * the momenta in place between iterations). pacuda_sweep_create uploads the item table once,
 * gives the sweep its own counter and scratch, and captures the launch into a HIP graph;
 * pacuda_sweep_run replays the graph on `stream` (NULL: the library's stream), asynchronously
 * unless kernel_ms is set (then it waits and stores the graph's time). The items' buffers must
 * outlive the sweep. Two sweeps never share state; one sweep's runs on one stream are ordered.
 */
typedef struct pacuda_sweep pacuda_sweep;
int pacuda_sweep_create(const pacuda_item *items, uint32_t count, const pacuda_launch *launch, pacuda_sweep **sweep);
int pacuda_sweep_run(pacuda_sweep *sweep, void *stream, float *kernel_ms);
void pacuda_sweep_destroy(pacuda_sweep *sweep);

/* device memory for callers without a GPU framework (NULL / nonzero on failure) */
void *pacuda_device_alloc(size_t bytes);
void pacuda_device_free(void *p);
int pacuda_copy_to_device(void *dst, const void *src, size_t bytes);
int pacuda_copy_to_host(void *dst, const void *src, size_t bytes);
int pacuda_synchronize(void);
/* compute units of device 0 (the persistent grid is sized from it) */
int pacuda_device_compute_units(void);

/*
 * The surrogate of log|M|^2: a multilayer perceptron `inputs -> hidden (x layers, SiLU) -> 1`
 * with BF16 matrix-core products and FP32 accumulation, FP32 master weights and Adam, on the
 * GPU. No reference counterpart. Parameters (fp32) per hidden layer l: W_l [hidden][fan_in]
 * (fan_in = inputs for l = 0, else hidden), b_l [hidden]; then w_out [hidden], b_out.
 * hidden: 64, 128 or 256; batches are multiples of 64; x, y and target are device pointers
 * (x [batch][inputs] fp32, y and target [batch] fp32).
 */
typedef struct pacuda_mlp pacuda_mlp;
int pacuda_mlp_create(uint32_t inputs, uint32_t hidden, uint32_t layers, uint32_t max_batch, pacuda_mlp **out);
void pacuda_mlp_destroy(pacuda_mlp *mlp);
uint64_t pacuda_mlp_param_count(const pacuda_mlp *mlp);
int pacuda_mlp_set_params(pacuda_mlp *mlp, const float *host, uint64_t count);
int pacuda_mlp_get_params(const pacuda_mlp *mlp, float *host, uint64_t count);
int pacuda_mlp_get_grads(const pacuda_mlp *mlp, float *host, uint64_t count);
/* y = network(x); with `train` the activations are kept for pacuda_mlp_backward (batch up to
 * max_batch). Asynchronous on the library's stream. */
int pacuda_mlp_forward(pacuda_mlp *mlp, const float *x, uint32_t batch, float *y, uint32_t train);
/* gradients of mean((y - target)^2) after a training forward of the same batch; loss (host,
 * may be NULL) receives the loss */
int pacuda_mlp_backward(pacuda_mlp *mlp, const float *y, const float *target, uint32_t batch, float *loss);
/* one Adam step (step counts from 1) on the master weights, then the bf16 copies */
int pacuda_mlp_adam(pacuda_mlp *mlp, float lr, float beta1, float beta2, float eps, uint32_t step);
/* `steps` training steps, each forward (train), backward and Adam of one batch: batch s takes
 * x + s batch inputs and target + s batch (device). Step first_step + s (0-based) of a run of
 * total_steps uses the learning rate (float)(lr 0.5 (1 + cos(pi (first_step + s) / total_steps)))
 * and Adam step number first_step + s + 1: the same arithmetic as the three calls above made in
 * that order with those arguments. loss (host, may be NULL) receives the last batch's loss.
 * Queued on the library's stream; returns after the last step when loss is set. */
int pacuda_mlp_train_steps(pacuda_mlp *mlp, const float *x, const float *target, uint32_t batch, uint32_t steps,
                           double lr, uint32_t first_step, uint32_t total_steps, float beta1, float beta2, float eps,
                           float *loss);
/* waits for the queued work and copies the loss of the last backward (or training step) */
int pacuda_mlp_last_loss(pacuda_mlp *mlp, float *loss);

/*
 * RAMBO phase space of 2 -> legs-2 massless partons in f64 on the device (convert/src/record.rs:
 * 17-52 on a counter-based stream: point `first + index[i]` of stream `seed`, or `first + i` when
 * index is NULL), written as the runner's external momenta (when external is not NULL)
 * external[(l * 4 + k) * stride + i] = crossing_sign[l] * p[crossing_src[l]][k];
 * pass[i] = 1 when the outgoing partons pass pT > pt_min, |eta| < eta_max, Delta R > dr_min;
 * features (or NULL): [points][3 (legs - 2) + 2] floats in [0, 1], the surrogate's inputs;
 * log_prior (or NULL): log(sum_{a<b} s_ab^4 / prod_l s_{l,l+1}), s_ab = |2 p_a.p_b| over the
 * legs in crossing order (cyclic): the pure-gluon MHV Parke-Taylor structure of the ordering.
 * All pointers but the crossing tables are device pointers.
 */
int pacuda_rambo(uint64_t seed, uint64_t first, uint32_t points, uint32_t legs, double ecm,
                 const uint32_t *crossing_src, const double *crossing_sign, double *external, uint32_t stride,
                 float *features, uint8_t *pass, double pt_min, double eta_max, double dr_min, double *log_prior,
                 const uint32_t *index);

/*
 * The reference's RAMBO stream (convert/src/record.rs:17-52), on the host: xorshift64 from
 * `state` (nonzero; record.rs takes seed | 1), points of `legs` massless partons at sqrt(s) = ecm
 * as [point][leg][E, px, py, pz] (beams first), drawn until `points` pass the cuts (pT > pt_min,
 * |eta| < eta_max, Delta R > dr_min between outgoing partons; pt_min <= 0: no cuts); `drawn`
 * (may be NULL) gets the number drawn. 0 OK, 2 invalid.
 */
int pacuda_rambo_record(uint64_t state, uint32_t legs, uint32_t points, double ecm, double pt_min, double eta_max,
                        double dr_min, double *momenta, uint64_t *drawn);
/* out[r * count + j] = external[r * stride + index[j]] for rows r < rows; features likewise
 * ([j][width] <- [index[j]][width]) when features is not NULL (device pointers) */
int pacuda_gather(const double *external, uint32_t stride, uint32_t rows, const uint32_t *index, uint32_t count,
                  double *out, const float *features, uint32_t width, float *features_out);

/*
 * Unweighting on the device (no reference counterpart; all pointers are device pointers,
 * asynchronous on the library's stream).
 * pacuda_surrogate_weight: s[i] = exp(mean + scale * y[i] + log_prior[i]) where pass[i] (or
 *   every i when pass is NULL), else 0; y or log_prior may be NULL (a zero term).
 * pacuda_weight_stats: adds the count of each w[i] > 0 to hist[bin of log w[i]] (bins over
 *   [log_lo, log_hi), clamped at both ends, bins <= PACUDA_MAX_BINS; hist may be NULL) and
 *   writes stats[3 b + {0, 1, 2}] = max, sum, sum of squares of w over the b-th of
 *   PACUDA_STATS_BLOCKS contiguous ranges (reduce them in order on the host).
 * pacuda_select: the indices i < count (ascending) that pass: flags[i] != 0 (when flags is not
 *   NULL) and, when w is not NULL, w[i] > 0 and u * threshold < w[i], u the uniform number
 *   (id + 1) of the counter-based stream `seed`, id = first + (ids ? ids[i] : i). Writes
 *   index[0..*selected) and, when values is not NULL, values[k] = w[index[k]]. index and
 *   values must hold count entries.
 */
#define PACUDA_STATS_BLOCKS 256
#define PACUDA_MAX_BINS 8192
int pacuda_surrogate_weight(const float *y, const double *log_prior, const uint8_t *pass, uint32_t count, double mean,
                            double scale, double *s);
int pacuda_weight_stats(const double *w, uint32_t count, double log_lo, double log_hi, uint32_t bins, uint32_t *hist,
                        double *stats);
int pacuda_select(const double *w, const uint8_t *flags, const uint32_t *ids, uint64_t first, uint32_t count,
                  double threshold, uint64_t seed, uint32_t *index, double *values, uint32_t *selected);

#ifdef __cplusplus
}
#endif
#endif
