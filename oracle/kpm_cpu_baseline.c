/*
 * ORACLE -- test infrastructure only (CPU baseline leg of the benchmark).
 *
 * A C restatement of the reference's default CPU compute for stochastic DOS moments
 * (`DefaultCompute` handling `BatchDiagonalMoments`):
 *   - ELLPACK storage, column-major with a 32-byte aligned pitch, padding value 0 and the index
 *     of the previous row (cppcore/include/numeric/ellmatrix.hpp:10-82)
 *   - vectors batched W at a time as an N x W row-major block, W = 32 bytes / sizeof(scalar)
 *     (cppcore/include/support/simd.hpp:35-44), leftover singles promoted to one more batch when
 *     num_singles > num_threads*W/2 (cppcore/src/kpm/default/Compute.cpp:52-88)
 *   - a pool of worker threads over batches; the starter is drawn under a lock, columns in order
 *     (Compute.cpp:23-44, cppcore/include/kpm/Starter.hpp:42-50)
 *   - ELL make_r1 (Starter.hpp:86-118), kpm_spmv_diagonal block form
 *     (cppcore/include/compute/kernel_polynomial.hpp:287-323)
 *   - the `basic` or `interleaved` diagonal recurrence over BFS slices
 *     (cppcore/include/kpm/calc_moments.hpp:35-89)
 *   - BatchAccumulator: the mean over all vectors (cppcore/src/kpm/Moments.cpp:30-48)
 * Real scalars only (f64: W=4, f32: W=8).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

typedef struct { uint32_t mt[624]; int mti; } bl_mt;
static void bl_seed(bl_mt* g, uint32_t s) {
    g->mt[0] = s;
    for (int i = 1; i < 624; ++i) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->mti = 624;
}
static uint32_t bl_next(bl_mt* g) {
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
    y ^= (y >> 11); y ^= (y << 7) & 0x9d2c5680u; y ^= (y << 15) & 0xefc60000u; y ^= (y >> 18);
    return y;
}
static double bl_canon_f64(bl_mt* g) {
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) { sum += (double)bl_next(g) * tmp; tmp *= 4294967296.0; }
    double r = sum / tmp;
    return r >= 1.0 ? 0x1.fffffffffffffp-1 : r;
}
static float bl_canon_f32(bl_mt* g) {
    float r = (float)bl_next(g) / 4294967296.0f;
    return r >= 1.0f ? 0x1.fffffep-1f : r;
}

#define BL_T double
#define BL_W 4
#define BL_SFX f64
#define BL_CANON bl_canon_f64
#include "kpm_cpu_baseline_impl.h"

#define BL_T float
#define BL_W 8
#define BL_SFX f32
#define BL_CANON bl_canon_f32
#include "kpm_cpu_baseline_impl.h"
