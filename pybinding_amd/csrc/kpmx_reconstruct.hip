// Reconstruction sums: the Chebyshev series of the spectral density and the Green's function over
// every (energy, column), and the KuboBastin energy-sample sums over the M x M moments
// (cppcore/include/kpm/reconstruct.hpp:16-143).  The per-energy quantities (acos of the scaled
// energy and the prefactor) are formed by the caller in the moments' real type, as the reference
// does; these kernels evaluate the O(E M cols) and O(P M^2) sums.  Every sum runs in a fixed
// order, so results are deterministic.
#include <hip/hip_runtime.h>

#include "kpmx_device.hpp"
#include "kpmx_internal.hpp"

namespace kpmx {
namespace {

__device__ __forceinline__ float cos_of(float x) { return cosf(x); }
__device__ __forceinline__ double cos_of(double x) { return cos(x); }
__device__ __forceinline__ void sincos_of(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ void sincos_of(double x, double* s, double* c) { sincos(x, s, c); }

constexpr int WAVE = 64;
constexpr int RB_THREADS = 256;  // 4 waves

// out[e][c] = pref[e] * sum_n mu[c][n] cos(n theta[e])                       (EXP = false)
// out[e][c] = pref[e] * sum_n mu[c][n] exp(-i n theta[e]), mu and pref complex  (EXP = true)
// Workgroup (e, column group of 4): the 256 cosines (and sines) of one chunk of orders are formed
// once in LDS and read by the four waves, one column each; lanes stride over the orders.
template <class R, bool EXP>
__global__ __launch_bounds__(RB_THREADS) void k_cheb_series(const R* __restrict__ mu, int64_t M, int64_t cols,
                                                            const R* __restrict__ theta, const R* __restrict__ pref,
                                                            double* __restrict__ out) {
    __shared__ R s_cos[RB_THREADS];
    __shared__ R s_sin[RB_THREADS];
    const int64_t e = blockIdx.x;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int64_t c = (int64_t)blockIdx.y * (RB_THREADS / WAVE) + wave;
    const R th = theta[e];
    R acc_re = 0, acc_im = 0;
    for (int64_t n0 = 0; n0 < M; n0 += RB_THREADS) {
        __syncthreads();
        const int64_t n = n0 + threadIdx.x;
        if (n < M) {
            const R ph = (R)n * th;
            if constexpr (EXP) {
                R sv, cv;
                sincos_of(ph, &sv, &cv);
                s_cos[threadIdx.x] = cv;
                s_sin[threadIdx.x] = sv;
            } else {
                s_cos[threadIdx.x] = cos_of(ph);
            }
        }
        __syncthreads();
        if (c < cols) {
            const int lim = (int)(M - n0 < RB_THREADS ? M - n0 : RB_THREADS);
            for (int i = lane; i < lim; i += WAVE) {
                if constexpr (EXP) {
                    // (mr + i mi)(cos - i sin)
                    const R mr = mu[2 * (c * M + n0 + i)], mi = mu[2 * (c * M + n0 + i) + 1];
                    acc_re += mr * s_cos[i] + mi * s_sin[i];
                    acc_im += mi * s_cos[i] - mr * s_sin[i];
                } else {
                    acc_re += mu[c * M + n0 + i] * s_cos[i];
                }
            }
        }
    }
    for (int w = WAVE / 2; w >= 1; w >>= 1) {
        acc_re += __shfl_xor(acc_re, w, WAVE);
        if constexpr (EXP) acc_im += __shfl_xor(acc_im, w, WAVE);
    }
    if (c < cols && lane == 0) {
        if constexpr (EXP) {
            const R pr = pref[2 * e], pi = pref[2 * e + 1];
            out[2 * (e * cols + c)] = (double)(pr * acc_re - pi * acc_im);
            out[2 * (e * cols + c) + 1] = (double)(pr * acc_im + pi * acc_re);
        } else {
            out[e * cols + c] = (double)(pref[e] * acc_re);
        }
    }
}

// KuboBastin basis of every sample q and order m (reconstruct.hpp:90-143):
//   t[q][m] = cos(m th_q),  a[q][m] = (E_q - i m sqrt(1 - E_q^2)) e^{i m th_q},  th_q = acos E_q
__global__ void k_kubo_basis(const double* __restrict__ samples, int64_t P, int64_t M, double* __restrict__ t,
                             double* __restrict__ a) {
    const int64_t total = P * M;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = e / M, m = e % M;
        const double en = samples[q];
        const double th = acos(en), sq = sqrt(1.0 - en * en);
        double s, c;
        sincos(th * (double)m, &s, &c);
        t[e] = c;
        a[2 * e] = en * c + (double)m * sq * s;
        a[2 * e + 1] = en * s - (double)m * sq * c;
    }
}

// Partial sums over row groups: workgroup (q, g) takes rows n = 4 g + wave + 4 G k and forms
//   sum_n [ t_n sum_m mu_nm a_m + conj(a_n) sum_m mu_nm t_m ]
// (the reference's sum_nm mu_nm Gamma_nm with Gamma_nm = g(n,m) + conj(g(m,n)), g(n,m) = t_n a_m).
__global__ __launch_bounds__(RB_THREADS) void k_kubo_sums(const double* __restrict__ mu, int64_t M,
                                                          const double* __restrict__ t, const double* __restrict__ a,
                                                          int groups, double* __restrict__ partial) {
    __shared__ double red[2][RB_THREADS / WAVE];
    const int64_t q = blockIdx.x;
    const int g = blockIdx.y;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const double* tq = t + q * M;
    const double* aq = a + 2 * q * M;
    double sr = 0, si = 0;
    for (int64_t n = (int64_t)g * (RB_THREADS / WAVE) + wave; n < M; n += (int64_t)groups * (RB_THREADS / WAVE)) {
        double yr = 0, yi = 0, zr = 0, zi = 0;  // y = sum_m mu_nm a_m, z = sum_m mu_nm t_m
        const double* row = mu + 2 * n * M;
        for (int64_t m = lane; m < M; m += WAVE) {
            const double mr = row[2 * m], mi = row[2 * m + 1];
            const double ar = aq[2 * m], ai = aq[2 * m + 1], tm = tq[m];
            yr += mr * ar - mi * ai;
            yi += mr * ai + mi * ar;
            zr += mr * tm;
            zi += mi * tm;
        }
        const double tn = tq[n], anr = aq[2 * n], ani = aq[2 * n + 1];
        sr += tn * yr + (anr * zr + ani * zi);
        si += tn * yi + (anr * zi - ani * zr);
    }
    for (int w = WAVE / 2; w >= 1; w >>= 1) {
        sr += __shfl_xor(sr, w, WAVE);
        si += __shfl_xor(si, w, WAVE);
    }
    if (lane == 0) {
        red[0][wave] = sr;
        red[1][wave] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partial[2 * (q * groups + g)] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        partial[2 * (q * groups + g) + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    }
}

// out[q] = (sum over groups, in order) / (1 - E_q^2)^2
__global__ void k_kubo_finish(const double* __restrict__ partial, int groups, const double* __restrict__ samples,
                              int64_t P, double* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    double sr = 0, si = 0;
    for (int g = 0; g < groups; ++g) {
        sr += partial[2 * (q * groups + g)];
        si += partial[2 * (q * groups + g) + 1];
    }
    const double en = samples[q];
    const double k = 1.0 / ((1.0 - en * en) * (1.0 - en * en));
    out[2 * q] = k * sr;
    out[2 * q + 1] = k * si;
}

}  // namespace

int launch_cheb_series(int real, bool exp_series, const void* mu, int64_t M, int64_t cols, int64_t num_energy,
                       const void* theta, const void* pref, double* out, hipStream_t s) {
    if (num_energy <= 0 || cols <= 0) return 0;
    const int64_t col_groups = (cols + RB_THREADS / WAVE - 1) / (RB_THREADS / WAVE);
    if (col_groups > 65535 || num_energy > 0x7fffffff) return -1;
    const dim3 grid((unsigned)num_energy, (unsigned)col_groups);
    auto go = [&](auto kernel, auto tag) {
        using R = decltype(tag);
        hipLaunchKernelGGL(kernel, grid, dim3(RB_THREADS), 0, s, static_cast<const R*>(mu), M, cols,
                           static_cast<const R*>(theta), static_cast<const R*>(pref), out);
    };
    if (real == 0) {
        if (exp_series) go(k_cheb_series<float, true>, float{});
        else go(k_cheb_series<float, false>, float{});
    } else if (real == 1) {
        if (exp_series) go(k_cheb_series<double, true>, double{});
        else go(k_cheb_series<double, false>, double{});
    } else {
        return -1;
    }
    return 0;
}

int kubo_sums_groups(int64_t M, int64_t P) {
    // about 4096 workgroups in all, at least 4 rows per group
    int64_t g = (4096 + P - 1) / (P > 0 ? P : 1);
    const int64_t gmax = (M + 15) / 16;
    if (g > gmax) g = gmax;
    if (g < 1) g = 1;
    if (g > 65535) g = 65535;
    return (int)g;
}

void launch_kubo_sums(const double* mu, int64_t M, const double* samples, int64_t P, double* t, double* a,
                      double* partial, int groups, double* out, hipStream_t s) {
    if (P <= 0) return;
    const int64_t total = P * M;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_kubo_basis, dim3((unsigned)blocks), dim3(256), 0, s, samples, P, M, t, a);
    hipLaunchKernelGGL(k_kubo_sums, dim3((unsigned)P, (unsigned)groups), dim3(RB_THREADS), 0, s, mu, M, t, a, groups,
                       partial);
    hipLaunchKernelGGL(k_kubo_finish, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, partial, groups, samples, P,
                       out);
}

}  // namespace kpmx
