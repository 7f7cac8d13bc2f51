/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the reference KPM moment engine (pybinding, cppcore/), used as the
 * checker for the HIP product path and as the CPU baseline.  Only tests/, the smoke check and
 * the benchmark's CPU-baseline leg may load it.  The product never links or calls it.
 *
 * What is restated here (reference file:line):
 *   - std::mt19937 default seed 5489 and libstdc++ generate_canonical
 *       (cppcore/include/numeric/random.hpp:24-50, cppcore/src/kpm/Starter.cpp:48-83)
 *   - random starters: Rademacher (real) and exp(2*pi_f*i*U) phases (complex), columns drawn
 *       one after another (cppcore/include/kpm/Starter.hpp:42-50, Starter.cpp:59-82)
 *   - make_r1 = 0.5 * H2 * r0 (Starter.hpp:52-70)
 *   - kpm_spmv / kpm_spmv_diagonal CSR semantics (cppcore/include/compute/kernel_polynomial.hpp:18-81)
 *   - calc_moments::basic diagonal and off-diagonal (cppcore/include/kpm/calc_moments.hpp:35-51,102-114)
 *   - Diagonal / Generic / MultiUnit / DenseMatrix collectors
 *       (cppcore/src/kpm/default/collectors.cpp:5-106)
 *
 * The file is a template instantiated four times (f32, f64, c64, c128) via kpm_oracle_impl.h.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>
#include <complex.h>

/* ---------------------------------------------------------------- mt19937 (MT19937, 32 bit) */
/* Matsumoto & Nishimura 1998; parameters of std::mt19937 ([rand.predef]). */
typedef struct { uint32_t mt[624]; int mti; } orc_mt;

static void orc_mt_seed(orc_mt* g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i) {
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    }
    g->mti = 624;
}

static uint32_t orc_mt_next(orc_mt* g) {
    if (g->mti >= 624) {
        for (int k = 0; k < 624; ++k) {
            uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(k + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[k] = v;
        }
        g->mti = 0;
    }
    uint32_t y = g->mt[g->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* Raw 32-bit stream of a generator seeded with `seed` (5489 = default construction). */
void orc_mt19937_stream(uint32_t seed, int64_t n, uint32_t* out) {
    orc_mt g;
    orc_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = orc_mt_next(&g);
}

/* libstdc++ generate_canonical<double, 53>: two draws, low word first, summed in double. */
static double orc_canonical_f64(orc_mt* g) {
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
        sum += (double)orc_mt_next(g) * tmp;
        tmp *= 4294967296.0;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

/* libstdc++ generate_canonical<float, 24>: a single draw rounded to float. */
static float orc_canonical_f32(orc_mt* g) {
    float sum = 0.0f, tmp = 1.0f;
    sum += (float)orc_mt_next(g) * tmp;
    tmp *= 4294967296.0f;
    float ret = sum / tmp;
    if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
    return ret;
}

/* Uniform [0,1) draws exactly as std::uniform_real_distribution<real_t>{} yields them. */
void orc_uniform_f64(int64_t n, double* out) {
    orc_mt g; orc_mt_seed(&g, 5489u);
    for (int64_t i = 0; i < n; ++i) out[i] = orc_canonical_f64(&g);
}
void orc_uniform_f32(int64_t n, float* out) {
    orc_mt g; orc_mt_seed(&g, 5489u);
    for (int64_t i = 0; i < n; ++i) out[i] = orc_canonical_f32(&g);
}

/* 2 * constant::pi evaluated in float (cppcore/include/numeric/constant.hpp:8, Starter.cpp:75) */
static const float ORC_TWO_PI_F = 2.0f * 3.14159265358979323846f;

#define ORC_T float
#define ORC_R float
#define ORC_SFX f32
#define ORC_CPLX 0
#include "kpm_oracle_impl.h"

#define ORC_T double
#define ORC_R double
#define ORC_SFX f64
#define ORC_CPLX 0
#include "kpm_oracle_impl.h"

#define ORC_T float _Complex
#define ORC_R float
#define ORC_SFX c64
#define ORC_CPLX 1
#include "kpm_oracle_impl.h"

#define ORC_T double _Complex
#define ORC_R double
#define ORC_SFX c128
#define ORC_CPLX 1
#include "kpm_oracle_impl.h"
