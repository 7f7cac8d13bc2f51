// Kubo-Bastin moment contraction  mu_nm += sum_k L[n][k] conj(R[m][k])
// (the reference's MomentMultiplication::matrix_mul_add, cppcore/src/kpm/Moments.cpp:92-101).
//
// Two implementations of the same M x M contraction over a long k (= sites x vectors):
//   mode 0 (FP64): 64x64 output tiles, 16-deep k panels staged in LDS, every product and sum in
//                  double on the vector ALUs (FP64 vector peak equals the FP64 matrix peak on
//                  gfx950), accumulation exact to double rounding.
//   mode 1 (FP32 split MFMA): each operand is split into a float "hi" part and a float "lo"
//                  remainder (x = hi + lo), and L R^H ~ Lhi Rhi^H + Lhi Rlo^H + Llo Rhi^H is
//                  formed with v_mfma_f32_32x32x2_f32; the float accumulators are flushed into
//                  double every 32-deep k panel, so the rounding of the float sums stays bounded
//                  by one panel's depth.
// Both split k into S slices written to a [S][M][M] workspace and summed in slice order, so the
// result is deterministic.
#include "kpmx_device.hpp"
#include "kpmx_internal.hpp"

namespace kpmx {

namespace {

constexpr int TM = 64;   // output tile edge
constexpr int TK = 16;   // k panel depth (FP64 path)

__device__ __forceinline__ d2 to_d2(float v) { return {double(v), 0.0}; }
__device__ __forceinline__ d2 to_d2(double v) { return {v, 0.0}; }
template <class R> __device__ __forceinline__ d2 to_d2(cplx<R> v) { return {double(v.re), double(v.im)}; }

template <class T>
__global__ __launch_bounds__(256) void k_gemm_nc_fp64(const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                                      int64_t ldb, int64_t M, int64_t kdim, int64_t kslice,
                                                      double* __restrict__ P) {
    constexpr bool CPLX = traits<T>::is_complex;
    __shared__ d2 As[TK][TM + 1];
    __shared__ d2 Bs[TK][TM + 1];
    const int t = threadIdx.x;
    const int tx = t % 16, ty = t / 16;
    const int64_t tiles = (M + TM - 1) / TM;
    const int64_t tile = blockIdx.x % (tiles * tiles);
    const int64_t s = blockIdx.x / (tiles * tiles);
    const int64_t n0 = (tile / tiles) * TM, m0 = (tile % tiles) * TM;
    const int64_t kbeg = s * kslice;
    const int64_t kend = (kbeg + kslice < kdim) ? kbeg + kslice : kdim;

    double cr[4][4], ci[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) cr[i][j] = ci[i][j] = 0.0;

    for (int64_t k0 = kbeg; k0 < kend; k0 += TK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int r = e / TK, kk = e % TK;
            const int64_t k = k0 + kk;
            const bool kin = k < kend;
            As[kk][r] = (kin && n0 + r < M) ? to_d2(A[(n0 + r) * lda + k]) : d2{0.0, 0.0};
            Bs[kk][r] = (kin && m0 + r < M) ? to_d2(B[(m0 + r) * ldb + k]) : d2{0.0, 0.0};
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            d2 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    // a * conj(b)
                    cr[i][j] += a[i].re * b[j].re;
                    if (CPLX) {
                        cr[i][j] += a[i].im * b[j].im;
                        ci[i][j] += a[i].im * b[j].re - a[i].re * b[j].im;
                    }
                }
        }
        __syncthreads();
    }
    const int width = CPLX ? 2 : 1;
    double* out = P + s * M * M * width;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t n = n0 + ty + 16 * i, m = m0 + tx + 16 * j;
            if (n < M && m < M) {
                if (CPLX) {
                    out[(n * M + m) * 2] = cr[i][j];
                    out[(n * M + m) * 2 + 1] = ci[i][j];
                } else {
                    out[n * M + m] = cr[i][j];
                }
            }
        }
}

// ---------------------------------------------------------------------------- FP32 split MFMA
// One 64-lane wave computes a 32x32 output block with v_mfma_f32_32x32x2_f32:
//   A operand: lane l holds A[row = l % 32][k = l / 32]  (2 k per instruction)
//   B operand: lane l holds B[k = l / 32][col = l % 32]
//   C/D: 16 floats per lane, element i at row 8*(i/4) + 4*(l/32) + (i%4), column l % 32.
// A block of 4 waves computes a 64x64 tile (2x2 waves).  K panels of KP = 32 are staged in LDS
// as hi/lo float pairs; the complex product a conj(b) = (ar br + ai bi) + i (ai br - ar bi) is
// formed from real MFMAs on the stacked operands.
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int KP = 32;

struct hilo { float hi, lo; };
__device__ __forceinline__ hilo split(double v) {
    const float h = (float)v;
    return {h, (float)(v - (double)h)};
}

template <class T>
__global__ __launch_bounds__(256) void k_gemm_nc_split(const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                                       int64_t ldb, int64_t M, int64_t kdim, int64_t kslice,
                                                       double* __restrict__ P) {
    constexpr bool CPLX = traits<T>::is_complex;
    // [k][row] of hi/lo for re and im parts
    __shared__ float Ar_h[KP][TM], Ar_l[KP][TM], Ai_h[KP][TM], Ai_l[KP][TM];
    __shared__ float Br_h[KP][TM], Br_l[KP][TM], Bi_h[KP][TM], Bi_l[KP][TM];
    const int t = threadIdx.x;
    const int lane = t % 64, wave = t / 64;
    const int wr = wave / 2, wc = wave % 2;  // 32x32 quadrant of this wave
    const int64_t tiles = (M + TM - 1) / TM;
    const int64_t tile = blockIdx.x % (tiles * tiles);
    const int64_t s = blockIdx.x / (tiles * tiles);
    const int64_t n0 = (tile / tiles) * TM, m0 = (tile % tiles) * TM;
    const int64_t kbeg = s * kslice;
    const int64_t kend = (kbeg + kslice < kdim) ? kbeg + kslice : kdim;

    double dre[16], dim_[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) dre[i] = dim_[i] = 0.0;

    for (int64_t k0 = kbeg; k0 < kend; k0 += KP) {
        // stage: 64 rows x 32 k per operand = 2048 elements, 8 per thread
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = t + 256 * q;
            const int r = e / KP, kk = e % KP;
            const int64_t k = k0 + kk;
            const bool kin = k < kend;
            const d2 av = (kin && n0 + r < M) ? to_d2(A[(n0 + r) * lda + k]) : d2{0.0, 0.0};
            const d2 bv = (kin && m0 + r < M) ? to_d2(B[(m0 + r) * ldb + k]) : d2{0.0, 0.0};
            const hilo ar = split(av.re), ai = split(av.im), br = split(bv.re), bi = split(bv.im);
            Ar_h[kk][r] = ar.hi; Ar_l[kk][r] = ar.lo; Ai_h[kk][r] = ai.hi; Ai_l[kk][r] = ai.lo;
            Br_h[kk][r] = br.hi; Br_l[kk][r] = br.lo; Bi_h[kk][r] = bi.hi; Bi_l[kk][r] = bi.lo;
        }
        __syncthreads();
        f32x16 cre, cim;
#pragma unroll
        for (int i = 0; i < 16; ++i) cre[i] = cim[i] = 0.f;
        const int arow = wr * 32 + (lane % 32);
        const int bcol = wc * 32 + (lane % 32);
        const int kl = lane / 32;
#pragma unroll 4
        for (int kk = 0; kk < KP; kk += 2) {
            const int k = kk + kl;
            const float arh = Ar_h[k][arow], arl = Ar_l[k][arow];
            const float brh = Br_h[k][bcol], brl = Br_l[k][bcol];
            // re += ar br
            cre = __builtin_amdgcn_mfma_f32_32x32x2f32(arh, brh, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x2f32(arh, brl, cre, 0, 0, 0);
            cre = __builtin_amdgcn_mfma_f32_32x32x2f32(arl, brh, cre, 0, 0, 0);
            if (CPLX) {
                const float aih = Ai_h[k][arow], ail = Ai_l[k][arow];
                const float bih = Bi_h[k][bcol], bil = Bi_l[k][bcol];
                // re += ai bi
                cre = __builtin_amdgcn_mfma_f32_32x32x2f32(aih, bih, cre, 0, 0, 0);
                cre = __builtin_amdgcn_mfma_f32_32x32x2f32(aih, bil, cre, 0, 0, 0);
                cre = __builtin_amdgcn_mfma_f32_32x32x2f32(ail, bih, cre, 0, 0, 0);
                // im += ai br - ar bi
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(aih, brh, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(aih, brl, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(ail, brh, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(-arh, bih, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(-arh, bil, cim, 0, 0, 0);
                cim = __builtin_amdgcn_mfma_f32_32x32x2f32(-arl, bih, cim, 0, 0, 0);
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            dre[i] += (double)cre[i];
            dim_[i] += (double)cim[i];
        }
        __syncthreads();
    }
    const int width = CPLX ? 2 : 1;
    double* out = P + s * M * M * width;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int64_t n = n0 + wr * 32 + 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
        const int64_t m = m0 + wc * 32 + (lane % 32);
        if (n < M && m < M) {
            if (CPLX) {
                out[(n * M + m) * 2] = dre[i];
                out[(n * M + m) * 2 + 1] = dim_[i];
            } else {
                out[n * M + m] = dre[i];
            }
        }
    }
}

__global__ void k_gemm_reduce(const double* P, int64_t slices, int64_t count, double* C) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x) {
        double s = C[e];
        for (int64_t q = 0; q < slices; ++q) s += P[q * count + e];
        C[e] = s;
    }
}

int64_t num_slices(int64_t M, int64_t kdim) {
    const int64_t tiles = (M + TM - 1) / TM;
    int64_t s = 2048 / (tiles * tiles);
    if (s < 1) s = 1;
    const int64_t max_s = (kdim + 255) / 256;
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    return s;
}

}  // namespace

int64_t gemm_workspace_doubles(int64_t M, int64_t kdim, int scalar) {
    return num_slices(M, kdim) * M * M * (scalar >= 2 ? 2 : 1);
}

template <class T>
static void gemm_t(int mode, const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M, int64_t kdim, double* C,
                   double* work, hipStream_t s) {
    const int64_t S = num_slices(M, kdim);
    int64_t kslice = (kdim + S - 1) / S;
    const int64_t align = mode == 1 ? KP : TK;
    kslice = (kslice + align - 1) / align * align;
    const int64_t tiles = (M + TM - 1) / TM;
    const dim3 grid((unsigned)(tiles * tiles * S)), block(256);
    if (mode == 1)
        hipLaunchKernelGGL(k_gemm_nc_split<T>, grid, block, 0, s, static_cast<const T*>(A), lda, static_cast<const T*>(B),
                           ldb, M, kdim, kslice, work);
    else
        hipLaunchKernelGGL(k_gemm_nc_fp64<T>, grid, block, 0, s, static_cast<const T*>(A), lda, static_cast<const T*>(B),
                           ldb, M, kdim, kslice, work);
    const int64_t count = M * M * (traits<T>::is_complex ? 2 : 1);
    int64_t g = (count + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)g), dim3(256), 0, s, work, S, count, C);
}

int launch_gemm_nc(int scalar, int mode, const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M, int64_t kdim,
                   double* C, double* work, hipStream_t s) {
    if (mode != 0 && mode != 1) return -1;
    switch (scalar) {
        case 0: gemm_t<float>(mode, A, lda, B, ldb, M, kdim, C, work, s); break;
        case 1: gemm_t<double>(mode, A, lda, B, ldb, M, kdim, C, work, s); break;
        case 2: gemm_t<cf32>(mode, A, lda, B, ldb, M, kdim, C, work, s); break;
        case 3: gemm_t<cf64>(mode, A, lda, B, ldb, M, kdim, C, work, s); break;
        default: return -1;
    }
    return 0;
}

}  // namespace kpmx
