// Kubo-Bastin moment contraction  mu_nm += sum_k L[n][k] conj(R[m][k])
// (the reference's MomentMultiplication::matrix_mul_add, cppcore/src/kpm/Moments.cpp:92-101).
//
// Three implementations of the same M x M contraction over a long k (= sites x vectors):
//   mode 0 (FP64 MFMA): 128x64 output tiles on v_mfma_f64_16x16x4f64, 16-deep k panels staged
//                  in LDS with the next panel prefetched into registers, every product and sum in
//                  double (accumulation exact to double rounding); persistent XCD-aware grid
//                  (x3_plan) like mode 2.
//   mode 1 (FP32 split MFMA): each operand is split into a float "hi" part and a float "lo"
//                  remainder (x = hi + lo), and L R^H ~ Lhi Rhi^H + Lhi Rlo^H + Llo Rhi^H is
//                  formed with v_mfma_f32_32x32x2_f32; the float accumulators are flushed into
//                  double every 32-deep k panel, so the rounding of the float sums stays bounded
//                  by one panel's depth.
//   mode 2 (bf16x3 split MFMA): three bf16 pieces per operand value (written by the sweeps that
//                  produce L and R), six piece products per real product on
//                  v_mfma_f32_32x32x16_bf16 (fp32-accurate), see below.
// All split k into S slices written to a [S][M][M] workspace and summed in slice order, so the
// result is deterministic.
#include <algorithm>

#include "kpmx_device.hpp"
#include "kpmx_internal.hpp"

namespace kpmx {

namespace {

constexpr int TM = 64;   // output tile edge

__device__ __forceinline__ d2 to_d2(float v) { return {double(v), 0.0}; }
__device__ __forceinline__ d2 to_d2(double v) { return {v, 0.0}; }
template <class R> __device__ __forceinline__ d2 to_d2(cplx<R> v) { return {double(v.re), double(v.im)}; }

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

// ------------------------------------------------------------------- bf16x3 split-precision MFMA
// Every operand value x (fp64 or fp32) is split into three bf16 pieces x = x0 + x1 + x2 + O(2^-24 x)
// (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1), the remainders exact in the input
// precision), and a product x y is formed from the six piece products with i + j <= 2,
// accumulated by v_mfma_f32_32x32x16_bf16 in fp32 and flushed into fp64 every 64 k, giving an
// fp32-accurate contraction on the bf16 matrix cores.  Complex products use four real products
// (Cr += Lr Rr + Li Ri, Ci += Li Rr - Lr Ri), i.e. 24 MFMAs per 32x32x16 complex block.
//
// The pieces are produced once, by the sweeps that write the L and R panels (store_split), so
// the contraction only copies 16-byte chunks into LDS.
// Work decomposition (persistent, XCD-aware): a workgroup of 8 waves computes one 128x64 output
// tile over one K range; each wave owns a 32x32 block with two fp32 accumulator chains (real,
// imaginary) plus the fp64 sums.  The T output tiles of one K range run on the same XCD
// (workgroup b sits on XCD b % 8), so the K range's L and R rows are read from HBM once and
// served to the T workgroups from that XCD's L2.  Partial tiles go to [KG][M][M] and are summed
// in K-range order (deterministic).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int XTN = 128;         // output tile rows (n)
constexpr int XTM = 64;          // output tile columns (m)
constexpr int XK = 32;           // k per LDS stage (two MFMA k16 steps)
constexpr int XS = XK + 8;       // LDS row stride in bf16 (pads the 16-byte lane reads off bank conflicts)
constexpr int XTHREADS = 512;    // 8 waves: 4 x 2 blocks of 32 x 32

__device__ __forceinline__ bf16x8 negate(bf16x8 v) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

// The bf16x3 contraction on operands already split into bf16 planes by the producing sweeps
// (store_split): A planes pa(p, j) = A + (p * 3 + j) * plane, rows of `ld` elements (ld a
// multiple of 8, the row tail [kdim, ld) zero).  Staging is a plain 16-byte copy, software
// pipelined: the next k panel is loaded into registers while the current one is multiplied.
template <int NP>
__global__ __launch_bounds__(XTHREADS) void k_gemm_nc_planes(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                             int64_t ld, int64_t plane, int64_t M, int64_t kdim,
                                                             int64_t kslice, int groups_per_xcd, int64_t tiles_per_wg,
                                                             double* __restrict__ P) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    __shared__ __align__(16) uint16_t As[NP][3][XTN][XS];
    __shared__ __align__(16) uint16_t Bs[NP][3][XTM][XS];
    constexpr int CA = NP * 3 * XTN * (XK / 8);   // 16-byte chunks per stage
    constexpr int CB = NP * 3 * XTM * (XK / 8);
    constexpr int QA = (CA + XTHREADS - 1) / XTHREADS;
    constexpr int QB = (CB + XTHREADS - 1) / XTHREADS;
    const int t = threadIdx.x;
    const int lane = t % 64, wave = t / 64;
    const int br = wave / 2, bc = wave % 2;
    const int64_t tiles_n = (M + XTN - 1) / XTN, tiles_m = (M + XTM - 1) / XTM;
    const int64_t T2 = tiles_n * tiles_m;
    const int xcd = blockIdx.x % 8;
    const int slot = blockIdx.x / 8;
    const int slots = gridDim.x / 8;
    int64_t kg, tile0, tile_step;
    if (T2 <= slots) {
        if (slot >= groups_per_xcd * T2) return;   // idle workgroup (the whole block exits before any barrier)
        kg = (int64_t)xcd * groups_per_xcd + slot / T2;
        tile0 = slot % T2;
        tile_step = T2;
    } else {
        kg = xcd;
        tile0 = slot;
        tile_step = slots;
    }
    const int64_t kbeg = kg * kslice;
    const int64_t kend = (kbeg + kslice < kdim) ? kbeg + kslice : kdim;
    u32x4 ra[QA], rb[QB];
    // chunk e of a stage: plane pj, row r, 8-element chunk q (rows beyond M, chunks at or beyond
    // kend: zero)
    auto load = [&](const uint16_t* __restrict__ src, int64_t r0, int rows, int64_t k0, int e) -> u32x4 {
        const int q = e % (XK / 8), r = (e / (XK / 8)) % rows, pj = e / (rows * (XK / 8));
        const int64_t k = k0 + q * 8;
        if (r0 + r >= M || k >= kend) return u32x4{0, 0, 0, 0};
        return *reinterpret_cast<const u32x4*>(src + pj * plane + (r0 + r) * ld + k);
    };
    for (int64_t tile = tile0, it = 0; tile < T2 && it < tiles_per_wg; tile += tile_step, ++it) {
        const int64_t n0 = (tile / tiles_m) * XTN, m0 = (tile % tiles_m) * XTM;
        double dre[16], dim_[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) dre[i] = dim_[i] = 0.0;
        f32x16 cre, cim;   // two accumulator chains (registers: the prefetched panel shares the budget)
#pragma unroll
        for (int i = 0; i < 16; ++i) cre[i] = cim[i] = 0.f;
        if (kbeg < kend) {
#pragma unroll
            for (int q = 0; q < QA; ++q)
                if (t + q * XTHREADS < CA) ra[q] = load(A, n0, XTN, kbeg, t + q * XTHREADS);
#pragma unroll
            for (int q = 0; q < QB; ++q)
                if (t + q * XTHREADS < CB) rb[q] = load(B, m0, XTM, kbeg, t + q * XTHREADS);
        }
        int stage = 0;
        for (int64_t k0 = kbeg; k0 < kend; k0 += XK, ++stage) {
            __syncthreads();   // the previous stage's LDS operands are no longer read
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                const int e = t + q * XTHREADS;
                if (e < CA) {
                    const int qq = e % (XK / 8), r = (e / (XK / 8)) % XTN, pj = e / (XTN * (XK / 8));
                    *reinterpret_cast<u32x4*>(&As[pj / 3][pj % 3][r][qq * 8]) = ra[q];
                }
            }
#pragma unroll
            for (int q = 0; q < QB; ++q) {
                const int e = t + q * XTHREADS;
                if (e < CB) {
                    const int qq = e % (XK / 8), r = (e / (XK / 8)) % XTM, pj = e / (XTM * (XK / 8));
                    *reinterpret_cast<u32x4*>(&Bs[pj / 3][pj % 3][r][qq * 8]) = rb[q];
                }
            }
            __syncthreads();
            if (k0 + XK < kend) {   // the next panel is in flight during the multiplications
#pragma unroll
                for (int q = 0; q < QA; ++q)
                    if (t + q * XTHREADS < CA) ra[q] = load(A, n0, XTN, k0 + XK, t + q * XTHREADS);
#pragma unroll
                for (int q = 0; q < QB; ++q)
                    if (t + q * XTHREADS < CB) rb[q] = load(B, m0, XTM, k0 + XK, t + q * XTHREADS);
            }
            const int ar = br * 32 + (lane % 32);
            const int bcl = bc * 32 + (lane % 32);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int ko = s * 16 + (lane / 32) * 8;
                bf16x8 a[NP][3], b[NP][3];
#pragma unroll
                for (int p = 0; p < NP; ++p)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        a[p][j] = *reinterpret_cast<const bf16x8*>(&As[p][j][ar][ko]);
                        b[p][j] = *reinterpret_cast<const bf16x8*>(&Bs[p][j][bcl][ko]);
                    }
#define KPMX_PAIRS(acc, X, Y)                                                        \
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[2], Y[0], acc, 0, 0, 0);         \
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[1], Y[1], acc, 0, 0, 0);         \
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[0], Y[2], acc, 0, 0, 0);         \
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[1], Y[0], acc, 0, 0, 0);         \
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[0], Y[1], acc, 0, 0, 0);         \
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(X[0], Y[0], acc, 0, 0, 0);
                KPMX_PAIRS(cre, a[0], b[0]);
                if constexpr (NP == 2) {
                    KPMX_PAIRS(cim, a[1], b[0]);
                    KPMX_PAIRS(cre, a[1], b[1]);
                    b[1][0] = negate(b[1][0]);
                    b[1][1] = negate(b[1][1]);
                    b[1][2] = negate(b[1][2]);
                    KPMX_PAIRS(cim, a[0], b[1]);
                }
#undef KPMX_PAIRS
            }
            if ((stage & 1) == 1 || k0 + XK >= kend) {   // flush the fp32 sums every 64 k
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    dre[i] += (double)cre[i];
                    dim_[i] += (double)cim[i];
                    cre[i] = cim[i] = 0.f;
                }
            }
        }
        double* out = P + kg * M * M * NP;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int64_t n = n0 + br * 32 + 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
            const int64_t m = m0 + bc * 32 + (lane % 32);
            if (n < M && m < M) {
                if (NP == 2) {
                    out[(n * M + m) * 2] = dre[i];
                    out[(n * M + m) * 2 + 1] = dim_[i];
                } else {
                    out[n * M + m] = dre[i];
                }
            }
        }
    }
}

struct X3Plan {
    int64_t kg_total, kslice, tiles_per_wg;
    int groups_per_xcd, grid;
};

X3Plan x3_plan(int64_t M, int64_t kdim, int num_cu) {
    X3Plan p;
    const int64_t T2 = ((M + XTN - 1) / XTN) * ((M + XTM - 1) / XTM);
    p.grid = std::max(8, (num_cu / 8) * 8);
    const int slots = p.grid / 8;
    if (T2 <= slots) {
        p.groups_per_xcd = (int)(slots / T2);
        p.kg_total = 8 * (int64_t)p.groups_per_xcd;
        p.tiles_per_wg = 1;
    } else {
        p.groups_per_xcd = 1;
        p.kg_total = 8;
        p.tiles_per_wg = (T2 + slots - 1) / slots;
    }
    int64_t ks = (kdim + p.kg_total - 1) / p.kg_total;
    p.kslice = std::max<int64_t>(XK, (ks + XK - 1) / XK * XK);
    return p;
}

// ------------------------------------------------------------------- FP64 on the matrix cores
// The FP64 contraction with v_mfma_f64_16x16x4_f64: the work decomposition of the bf16x3 kernel
// (persistent, XCD-aware: 128x64 output tiles of one K range on one XCD, partial tiles per K range
// summed in order), every product and sum in double.  One 16x16x4 MFMA takes lane l's A value at
// (row l % 16, k l / 16) and B value at (k l / 16, column l % 16) and accumulates a 16x16 block
// whose lane l holds rows 4 j + l / 16, j < 4, of column l % 16 (result j in the j-th register
// pair: the f64 form interleaves the lane groups row by row).  A wave owns a 32x32 block (2x2
// MFMA blocks, real and imaginary accumulators); complex a conj(b) = (ar br + ai bi) + i (ai br - ar bi)
// takes four real MFMAs.  K panels of D64K = 16 are staged in LDS as [k][row] pairs; the next
// panel is loaded into registers while the current one is multiplied.
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int D64K = 16;

template <class T>
__global__ __launch_bounds__(XTHREADS) void k_gemm_nc_mfma64(const T* __restrict__ A, int64_t lda,
                                                             const T* __restrict__ B, int64_t ldb, int64_t M,
                                                             int64_t kdim, int64_t kslice, int groups_per_xcd,
                                                             int64_t tiles_per_wg, double* __restrict__ P) {
    constexpr bool CPLX = traits<T>::is_complex;
    constexpr int NP = CPLX ? 2 : 1;
    __shared__ d2 As[D64K][XTN + 1];
    __shared__ d2 Bs[D64K][XTM + 1];
    constexpr int EA = XTN * D64K / XTHREADS;   // elements staged per thread (A: 4, B: 2)
    constexpr int EB = XTM * D64K / XTHREADS;
    const int t = threadIdx.x;
    const int lane = t % 64, wave = t / 64;
    const int br = wave / 2, bc = wave % 2;
    const int64_t tiles_n = (M + XTN - 1) / XTN, tiles_m = (M + XTM - 1) / XTM;
    const int64_t T2 = tiles_n * tiles_m;
    const int xcd = blockIdx.x % 8;
    const int slot = blockIdx.x / 8;
    const int slots = gridDim.x / 8;
    int64_t kg, tile0, tile_step;
    if (T2 <= slots) {
        if (slot >= groups_per_xcd * T2) return;   // idle workgroup (the whole block exits before any barrier)
        kg = (int64_t)xcd * groups_per_xcd + slot / T2;
        tile0 = slot % T2;
        tile_step = T2;
    } else {
        kg = xcd;
        tile0 = slot;
        tile_step = slots;
    }
    const int64_t kbeg = kg * kslice;
    const int64_t kend = (kbeg + kslice < kdim) ? kbeg + kslice : kdim;
    d2 ra[EA], rb[EB];
    // element e of a stage: row e / D64K, k e % D64K (16 consecutive k of one row per 16 threads)
    auto load = [&](const T* __restrict__ src, int64_t ld, int64_t r0, int64_t k0, int e) -> d2 {
        const int r = e / D64K, kk = e % D64K;
        const int64_t k = k0 + kk;
        if (r0 + r >= M || k >= kend) return d2{0.0, 0.0};
        return to_d2(src[(r0 + r) * ld + k]);
    };
    for (int64_t tile = tile0, it = 0; tile < T2 && it < tiles_per_wg; tile += tile_step, ++it) {
        const int64_t n0 = (tile / tiles_m) * XTN, m0 = (tile % tiles_m) * XTM;
        f64x4 cr[2][2], ci[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) cr[i][j] = ci[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
        if (kbeg < kend) {
#pragma unroll
            for (int q = 0; q < EA; ++q) ra[q] = load(A, lda, n0, kbeg, t + q * XTHREADS);
#pragma unroll
            for (int q = 0; q < EB; ++q) rb[q] = load(B, ldb, m0, kbeg, t + q * XTHREADS);
        }
        for (int64_t k0 = kbeg; k0 < kend; k0 += D64K) {
            __syncthreads();   // the previous stage's LDS operands are no longer read
#pragma unroll
            for (int q = 0; q < EA; ++q) {
                const int e = t + q * XTHREADS;
                As[e % D64K][e / D64K] = ra[q];
            }
#pragma unroll
            for (int q = 0; q < EB; ++q) {
                const int e = t + q * XTHREADS;
                Bs[e % D64K][e / D64K] = rb[q];
            }
            __syncthreads();
            if (k0 + D64K < kend) {   // the next panel is in flight during the multiplications
#pragma unroll
                for (int q = 0; q < EA; ++q) ra[q] = load(A, lda, n0, k0 + D64K, t + q * XTHREADS);
#pragma unroll
                for (int q = 0; q < EB; ++q) rb[q] = load(B, ldb, m0, k0 + D64K, t + q * XTHREADS);
            }
#pragma unroll
            for (int ks = 0; ks < D64K; ks += 4) {
                const int k = ks + lane / 16;
                d2 a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) a[i] = As[k][br * 32 + i * 16 + lane % 16];
#pragma unroll
                for (int j = 0; j < 2; ++j) b[j] = Bs[k][bc * 32 + j * 16 + lane % 16];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].re, b[j].re, cr[i][j], 0, 0, 0);
                        if constexpr (CPLX) {
                            cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].im, b[j].im, cr[i][j], 0, 0, 0);
                            ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].im, b[j].re, ci[i][j], 0, 0, 0);
                            ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[i].re, b[j].im, ci[i][j], 0, 0, 0);
                        }
                    }
            }
        }
        double* out = P + kg * M * M * NP;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int64_t n = n0 + br * 32 + i * 16 + 4 * q + (lane / 16);
                    const int64_t m = m0 + bc * 32 + j * 16 + (lane % 16);
                    if (n < M && m < M) {
                        if constexpr (CPLX) {
                            out[(n * M + m) * 2] = cr[i][j][q];
                            out[(n * M + m) * 2 + 1] = ci[i][j][q];
                        } else {
                            out[n * M + m] = cr[i][j][q];
                        }
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

int64_t gemm_workspace_doubles(int64_t M, int64_t kdim, int scalar, int mode, int num_cu) {
    const int64_t S = (mode == 2 || mode == 0) ? x3_plan(M, kdim, num_cu).kg_total : num_slices(M, kdim);
    return S * M * M * (scalar >= 2 ? 2 : 1);
}

template <class T>
static void gemm_t(int mode, const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M, int64_t kdim, double* C,
                   double* work, int num_cu, hipStream_t s) {
    const int64_t count = M * M * (traits<T>::is_complex ? 2 : 1);
    if (mode == 0) {   // FP64 on the matrix cores, XCD-aware persistent grid
        const X3Plan p = x3_plan(M, kdim, num_cu);
        hipLaunchKernelGGL(k_gemm_nc_mfma64<T>, dim3((unsigned)p.grid), dim3(XTHREADS), 0, s, static_cast<const T*>(A),
                           lda, static_cast<const T*>(B), ldb, M, kdim, p.kslice, p.groups_per_xcd, p.tiles_per_wg, work);
        int64_t g = (count + 255) / 256;
        if (g > 65536) g = 65536;
        hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)g), dim3(256), 0, s, work, p.kg_total, count, C);
        return;
    }
    const int64_t S = num_slices(M, kdim);
    {
        int64_t kslice = (kdim + S - 1) / S;
        kslice = (kslice + KP - 1) / KP * KP;
        const int64_t tiles = (M + TM - 1) / TM;
        const dim3 grid((unsigned)(tiles * tiles * S)), block(256);
        hipLaunchKernelGGL(k_gemm_nc_split<T>, grid, block, 0, s, static_cast<const T*>(A), lda,
                           static_cast<const T*>(B), ldb, M, kdim, kslice, work);
    }
    int64_t g = (count + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)g), dim3(256), 0, s, work, S, count, C);
}

int launch_gemm_nc_planes(bool cplx, const uint16_t* A, const uint16_t* B, int64_t ld, int64_t plane, int64_t M,
                          int64_t kdim, double* C, double* work, int num_cu, hipStream_t s) {
    if (ld % 8 != 0) return -1;
    const X3Plan p = x3_plan(M, kdim, num_cu);
    if (cplx)
        hipLaunchKernelGGL(k_gemm_nc_planes<2>, dim3((unsigned)p.grid), dim3(XTHREADS), 0, s, A, B, ld, plane, M, kdim,
                           p.kslice, p.groups_per_xcd, p.tiles_per_wg, work);
    else
        hipLaunchKernelGGL(k_gemm_nc_planes<1>, dim3((unsigned)p.grid), dim3(XTHREADS), 0, s, A, B, ld, plane, M, kdim,
                           p.kslice, p.groups_per_xcd, p.tiles_per_wg, work);
    const int64_t count = M * M * (cplx ? 2 : 1);
    int64_t g = (count + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)g), dim3(256), 0, s, work, p.kg_total, count, C);
    return 0;
}

int launch_gemm_nc(int scalar, int mode, const void* A, int64_t lda, const void* B, int64_t ldb, int64_t M, int64_t kdim,
                   double* C, double* work, int num_cu, hipStream_t s) {
    if (mode != 0 && mode != 1) return -1;   // mode 2 runs on pre-split planes (launch_gemm_nc_planes)
    switch (scalar) {
        case 0: gemm_t<float>(mode, A, lda, B, ldb, M, kdim, C, work, num_cu, s); break;
        case 1: gemm_t<double>(mode, A, lda, B, ldb, M, kdim, C, work, num_cu, s); break;
        case 2: gemm_t<cf32>(mode, A, lda, B, ldb, M, kdim, C, work, num_cu, s); break;
        case 3: gemm_t<cf64>(mode, A, lda, B, ldb, M, kdim, C, work, num_cu, s); break;
        default: return -1;
    }
    return 0;
}

}  // namespace kpmx
