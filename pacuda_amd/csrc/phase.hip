// SPDX-License-Identifier: 0BSD
//
// Phase space on the GPU in f64: RAMBO points of 2 -> n-2 massless partons from a counter-based
// stream, the analysis cuts, the surrogate's input features, and a gather of selected points.
//
// RAMBO follows convert/src/record.rs:17-52 (per outgoing parton c = 2u - 1, phi = 2 pi u',
// E = -ln(u'' u'''), then the boost and scale to sqrt(s); beams (E/2, 0, 0, +-E/2)). Uniform k
// of stream `seed` is ((mix(mix(seed + g) + (k + 1) g) >> 11) + 0.5) / 2^53, mix the splitmix64
// finaliser and g = 0x9E3779B97F4A7C15: the same numbers pacuda_amd/phase_space.py draws, so a
// point is reproducible on the host up to the last bits of log/sin/cos.
//
// One lane per point; momenta go straight into the runner's layout
// external[(leg * 4 + k) * stride + i] with the crossing applied.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <vector>

#include "pacman_recurrence.h"
#include "pacuda_amd.h"
#include "runtime.hpp"

using namespace pacuda;

namespace {

constexpr int MAX_LEGS = 8;
constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform(uint64_t state, uint64_t k) {
    return ((double)(mix(state + (k + 1) * GAMMA) >> 11) + 0.5) / 9007199254740992.0;
}

struct Crossing {
    uint32_t src[MAX_LEGS];
    double sign[MAX_LEGS];
};

// LEGS is a compile-time count so that every per-parton array stays in registers (the leg
// permutation of the crossing is applied by unrolled selects, not by indexing)
template <int LEGS>
__global__ void __launch_bounds__(256) rambo_kernel(uint64_t state, uint64_t first, uint32_t points, double ecm,
                                                    Crossing cr, double *external, uint32_t stride, float *features,
                                                    uint8_t *pass, double pt_min, double eta_max, double dr_min,
                                                    double *log_prior, const uint32_t *index) {
    constexpr int M = LEGS - 2;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= points) return;
    const uint64_t base = (first + (index ? index[i] : i)) * (uint64_t)M * 4;
    double q[M][4], u[M][4];
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
    for (int j = 0; j < M; j++) {
#pragma unroll
        for (int r = 0; r < 4; r++) u[j][r] = uniform(state, base + 4 * j + r);
        const double c = 2.0 * u[j][0] - 1.0, s = sqrt(1.0 - c * c), f = 2.0 * M_PI * u[j][1];
        const double e = -log(u[j][2] * u[j][3]);
        double sf, cf;
        sincos(f, &sf, &cf);
        q[j][0] = e; q[j][1] = e * s * cf; q[j][2] = e * s * sf; q[j][3] = e * c;
        t0 += q[j][0]; t1 += q[j][1]; t2 += q[j][2]; t3 += q[j][3];
    }
    const double mass = sqrt(t0 * t0 - t1 * t1 - t2 * t2 - t3 * t3);
    const double b[3] = {-t1 / mass, -t2 / mass, -t3 / mass};
    const double x = ecm / mass, g = t0 / mass, a = 1.0 / (1.0 + g);
    double p[LEGS][4];
    p[0][0] = ecm / 2; p[0][1] = 0; p[0][2] = 0; p[0][3] = ecm / 2;
    p[1][0] = ecm / 2; p[1][1] = 0; p[1][2] = 0; p[1][3] = -ecm / 2;
#pragma unroll
    for (int j = 0; j < M; j++) {
        const double bq = b[0] * q[j][1] + b[1] * q[j][2] + b[2] * q[j][3];
        p[j + 2][0] = x * (g * q[j][0] + bq);
#pragma unroll
        for (int k = 0; k < 3; k++) p[j + 2][k + 1] = x * (q[j][k + 1] + b[k] * (q[j][0] + a * bq));
    }
    // cuts on the outgoing partons
    double eta[LEGS], phi[LEGS], pt_frac_min = 1.0;
    bool ok = true;
#pragma unroll
    for (int j = 2; j < LEGS; j++) {
        const double px = p[j][1], py = p[j][2], pz = p[j][3];
        const double pt = sqrt(px * px + py * py), pp = sqrt(px * px + py * py + pz * pz);
        eta[j] = 0.5 * log((pp + pz) / (pp - pz));
        phi[j] = atan2(py, px);
        ok = ok && pt > pt_min && fabs(eta[j]) < eta_max;
        pt_frac_min = fmin(pt_frac_min, pt / (ecm / 2));
    }
    double dr2_min = 1e30;
#pragma unroll
    for (int j = 2; j < LEGS; j++)
#pragma unroll
        for (int k = j + 1; k < LEGS; k++) {
            double dphi = fabs(phi[j] - phi[k]);
            if (dphi > M_PI) dphi = 2 * M_PI - dphi;
            const double deta = eta[j] - eta[k], dr2 = deta * deta + dphi * dphi;
            ok = ok && dr2 > dr_min * dr_min;
            dr2_min = fmin(dr2_min, dr2);
        }
    if (pass) pass[i] = ok ? 1 : 0;
    // the plan's legs: xl[l] = p[src[l]] (unrolled selects on the uniform crossing table)
    double xl[LEGS][4];
#pragma unroll
    for (int l = 0; l < LEGS; l++) {
#pragma unroll
        for (int k = 0; k < 4; k++) xl[l][k] = p[0][k];
#pragma unroll
        for (int s = 1; s < LEGS; s++)
            if (cr.src[l] == (uint32_t)s)
#pragma unroll
                for (int k = 0; k < 4; k++) xl[l][k] = p[s][k];
    }
    if (external)
#pragma unroll
        for (int l = 0; l < LEGS; l++)
#pragma unroll
            for (int k = 0; k < 4; k++) external[((size_t)l * 4 + k) * stride + i] = cr.sign[l] * xl[l][k];
    if (log_prior) {
        // log(sum_{a<b} s_ab^4 / prod_l s_{l,l+1}) over the plan's legs in order (cyclic),
        // s_ab = |2 p_a.p_b|: the helicity-summed MHV Parke-Taylor structure of the ordering (the
        // product of the adjacent invariants stays within f64's range: each lies between about
        // 10^1 and s = 10^6 GeV^2 for points that pass the cuts)
        double num = 0.0, den = 1.0;
#pragma unroll
        for (int a = 0; a < LEGS; a++)
#pragma unroll
            for (int c = a + 1; c < LEGS; c++) {
                const double sab = fabs(2.0 * (xl[a][0] * xl[c][0] - xl[a][1] * xl[c][1] - xl[a][2] * xl[c][2] -
                                               xl[a][3] * xl[c][3]));
                const double s2 = sab * sab;
                num += s2 * s2;
                if (c == a + 1 || (a == 0 && c == LEGS - 1)) den *= sab;
            }
        log_prior[i] = log(num) - log(den);
    }
    if (features) {
        // per outgoing parton: the uniforms of its direction and the CDF of its energy product
        // (t - t ln t for t = u u'), then the smallest pT over sqrt(s)/2 and the smallest Delta R / 4
        float *f = features + (size_t)i * (3 * M + 2);
#pragma unroll
        for (int j = 0; j < M; j++) {
            const double t = u[j][2] * u[j][3];
            f[3 * j] = (float)u[j][0];
            f[3 * j + 1] = (float)u[j][1];
            f[3 * j + 2] = (float)(t - t * log(t));
        }
        f[3 * M] = (float)pt_frac_min;
        f[3 * M + 1] = (float)fmin(1.0, sqrt(dr2_min) / 4.0);
    }
}

// out[(leg * 4 + k) * count + j] = external[(leg * 4 + k) * stride + index[j]]
__global__ void gather_kernel(const double *external, uint32_t stride, uint32_t rows, const uint32_t *index, uint32_t count,
                              double *out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const uint32_t src = index[j];
    for (uint32_t r = 0; r < rows; r++) out[(size_t)r * count + j] = external[(size_t)r * stride + src];
}

__global__ void gather_features(const float *features, uint32_t width, const uint32_t *index, uint32_t count, float *out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const uint32_t src = index[j];
    for (uint32_t r = 0; r < width; r++) out[(size_t)j * width + r] = features[(size_t)src * width + r];
}

uint64_t seed_state(uint64_t seed) {
    uint64_t z = seed + GAMMA;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" int pacuda_rambo(uint64_t seed, uint64_t first, uint32_t points, uint32_t legs, double ecm,
                            const uint32_t *crossing_src, const double *crossing_sign, double *external, uint32_t stride,
                            float *features, uint8_t *pass, double pt_min, double eta_max, double dr_min,
                            double *log_prior, const uint32_t *index) {
    if (legs < 3 || legs > MAX_LEGS || !crossing_src || !crossing_sign || (external && stride < points))
        return PACMAN_RECURRENCE_INVALID;
    Crossing cr;
    for (uint32_t l = 0; l < legs; l++) {
        if (crossing_src[l] >= legs) return PACMAN_RECURRENCE_INVALID;
        cr.src[l] = crossing_src[l];
        cr.sign[l] = crossing_sign[l];
    }
    if (points == 0) return PACMAN_RECURRENCE_OK;
    const dim3 grid((points + 255) / 256);
    const uint64_t st = seed_state(seed);
    hipStream_t s = runner_stream();
    switch (legs) {
#define PACUDA_RAMBO(L)                                                                                                   \
    case L:                                                                                                               \
        rambo_kernel<L><<<grid, 256, 0, s>>>(st, first, points, ecm, cr, external, stride, features, pass, pt_min, eta_max, \
                                             dr_min, log_prior, index);                                                   \
        break;
        PACUDA_RAMBO(3) PACUDA_RAMBO(4) PACUDA_RAMBO(5) PACUDA_RAMBO(6) PACUDA_RAMBO(7) PACUDA_RAMBO(8)
#undef PACUDA_RAMBO
    default: return PACMAN_RECURRENCE_INVALID;
    }
    return hip_failed(hipGetLastError(), "rambo") ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK;
}

extern "C" int pacuda_gather(const double *external, uint32_t stride, uint32_t rows, const uint32_t *index, uint32_t count,
                             double *out, const float *features, uint32_t width, float *features_out) {
    if (!external || !index || !out || (features && !features_out)) return PACMAN_RECURRENCE_INVALID;
    if (count == 0) return PACMAN_RECURRENCE_OK;
    gather_kernel<<<(count + 255) / 256, 256, 0, runner_stream()>>>(external, stride, rows, index, count, out);
    if (features) gather_features<<<(count + 255) / 256, 256, 0, runner_stream()>>>(features, width, index, count, features_out);
    return hip_failed(hipGetLastError(), "gather") ? PACMAN_RECURRENCE_DEVICE_ERROR : PACMAN_RECURRENCE_OK;
}

// ---- the reference's own stream, on the host -------------------------------------------------
//
// convert/src/record.rs:17-52: one xorshift64 (13, 7, 17) stream, uniform u = ((s >> 11) + 0.5) /
// 2^53; per outgoing parton c = 2u - 1, s = sqrt(1 - c c), f = 2 pi u, e = -ln(u u) (the two
// uniforms in that order), then the boost and scale to sqrt(s) = ecm, beams (ecm/2, 0, 0, +-ecm/2).
// Points that fail the cuts (pT, |eta|, Delta R of the outgoing partons, when pt_min > 0) are
// drawn and dropped; `drawn` counts every point drawn up to the last one kept. The generation is
// sequential, so it runs on the host; the points are then the same bits everywhere.
namespace {

bool passes(const double *p, uint32_t legs, double pt_min, double eta_max, double dr_min) {
    double eta[MAX_LEGS], phi[MAX_LEGS];
    for (uint32_t l = 2; l < legs; l++) {
        const double *q = p + 4 * l;
        const double pt = hypot(q[1], q[2]), mag = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        eta[l] = 0.5 * log((mag + q[3]) / (mag - q[3]));
        phi[l] = atan2(q[2], q[1]);
        if (!(pt > pt_min) || !(fabs(eta[l]) < eta_max)) return false;
    }
    for (uint32_t a = 2; a < legs; a++)
        for (uint32_t b = a + 1; b < legs; b++) {
            double dphi = fabs(phi[a] - phi[b]);
            if (dphi > M_PI) dphi = 2 * M_PI - dphi;
            const double deta = eta[a] - eta[b];
            if (!(deta * deta + dphi * dphi > dr_min * dr_min)) return false;
        }
    return true;
}

}  // namespace

extern "C" int pacuda_rambo_record(uint64_t state, uint32_t legs, uint32_t points, double ecm, double pt_min,
                                   double eta_max, double dr_min, double *momenta, uint64_t *drawn) {
    if (legs < 3 || legs > MAX_LEGS || state == 0 || (points && !momenta)) return PACMAN_RECURRENCE_INVALID;
    auto uniform = [&state]() {
        state ^= state << 13;
        state ^= state >> 7;
        state ^= state << 17;
        return ((double)(state >> 11) + 0.5) / 9007199254740992.0;
    };
    uint64_t count = 0;
    std::vector<double> point(4 * legs);
    for (uint32_t kept = 0; kept < points;) {
        double q[MAX_LEGS][4];
        const uint32_t m = legs - 2;
        for (uint32_t k = 0; k < m; k++) {
            const double c = 2.0 * uniform() - 1.0;
            const double s = sqrt(1.0 - c * c);
            const double f = 2.0 * M_PI * uniform();
            const double u1 = uniform();
            const double u2 = uniform();
            const double e = -log(u1 * u2);
            q[k][0] = e;
            q[k][1] = e * s * cos(f);
            q[k][2] = e * s * sin(f);
            q[k][3] = e * c;
        }
        double total[4];
        for (int k = 0; k < 4; k++) {
            total[k] = 0.0;
            for (uint32_t j = 0; j < m; j++) total[k] = total[k] + q[j][k];
        }
        const double mass = sqrt(total[0] * total[0] - total[1] * total[1] - total[2] * total[2] - total[3] * total[3]);
        const double b[3] = {-total[1] / mass, -total[2] / mass, -total[3] / mass};
        const double x = ecm / mass, g = total[0] / mass, a = 1.0 / (1.0 + g);
        double *p = point.data();
        p[0] = ecm / 2; p[1] = 0.0; p[2] = 0.0; p[3] = ecm / 2;
        p[4] = ecm / 2; p[5] = 0.0; p[6] = 0.0; p[7] = -ecm / 2;
        for (uint32_t j = 0; j < m; j++) {
            const double bq = b[0] * q[j][1] + b[1] * q[j][2] + b[2] * q[j][3];
            double *o = p + 4 * (j + 2);
            o[0] = x * (g * q[j][0] + bq);
            for (int k = 0; k < 3; k++) o[k + 1] = x * (q[j][k + 1] + b[k] * (q[j][0] + a * bq));
        }
        count++;
        if (pt_min > 0 && !passes(p, legs, pt_min, eta_max, dr_min)) continue;
        std::copy(point.begin(), point.end(), momenta + (size_t)kept * 4 * legs);
        kept++;
    }
    if (drawn) *drawn = count;
    return PACMAN_RECURRENCE_OK;
}
