// Device-side scalar helpers for the KPM kernels (gfx950).
//
// Scalars: float, double, and the complex pairs cf32 / cf64 laid out as {re, im} exactly like
// std::complex<T>, so host arrays of std::complex can be copied byte for byte.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kpmx {

template <class R>
struct cplx {
    R re, im;
};
using cf32 = cplx<float>;
using cf64 = cplx<double>;

// Two adjacent real fp64 vector columns handled by one lane (16-byte accesses; two rows of a
// 64-column block per wave instruction).  Arithmetic is per column; the matrix stays real double.
struct dpair {
    double a, b;
};

// real: the type of the scale factor and sums; value: the matrix entry type (a dpair block is
// multiplied by a real double matrix)
template <class T> struct traits;
template <> struct traits<float>  { using real = float;  using value = float;  static constexpr bool is_complex = false; static constexpr int code = 0; };
template <> struct traits<double> { using real = double; using value = double; static constexpr bool is_complex = false; static constexpr int code = 1; };
template <> struct traits<cf32>   { using real = float;  using value = cf32;   static constexpr bool is_complex = true;  static constexpr int code = 2; };
template <> struct traits<cf64>   { using real = double; using value = cf64;   static constexpr bool is_complex = true;  static constexpr int code = 3; };
template <> struct traits<dpair>  { using real = double; using value = double; static constexpr bool is_complex = false; static constexpr int code = 1; };

// ----------------------------------------------------------------------------- arithmetic
__host__ __device__ __forceinline__ float  zero(float)  { return 0.f; }
__host__ __device__ __forceinline__ double zero(double) { return 0.0; }
template <class R> __host__ __device__ __forceinline__ cplx<R> zero(cplx<R>) { return {R(0), R(0)}; }

// a*b + c (plain product, no complex special-value handling, like the reference's detail::mul)
__device__ __forceinline__ float  fmac(float a, float b, float c)    { return a * b + c; }
__device__ __forceinline__ double fmac(double a, double b, double c) { return a * b + c; }
template <class R>
__device__ __forceinline__ cplx<R> fmac(cplx<R> a, cplx<R> b, cplx<R> c) {
    return {a.re * b.re - a.im * b.im + c.re, a.re * b.im + a.im * b.re + c.im};
}

__device__ __forceinline__ float  scal(float a, float s)   { return a * s; }
__device__ __forceinline__ double scal(double a, double s) { return a * s; }
template <class R> __device__ __forceinline__ cplx<R> scal(cplx<R> a, R s) { return {a.re * s, a.im * s}; }

__device__ __forceinline__ float  sub(float a, float b)   { return a - b; }
__device__ __forceinline__ double sub(double a, double b) { return a - b; }
template <class R> __device__ __forceinline__ cplx<R> sub(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }

__host__ __device__ __forceinline__ dpair zero(dpair) { return {0.0, 0.0}; }
__device__ __forceinline__ dpair fmac(double a, dpair b, dpair c) { return {a * b.a + c.a, a * b.b + c.b}; }
__device__ __forceinline__ dpair scal(dpair a, double s) { return {a.a * s, a.b * s}; }
__device__ __forceinline__ dpair sub(dpair a, dpair b) { return {a.a - b.a, a.b - b.b}; }

// |a|^2 in double
__device__ __forceinline__ double sqabs(float a)  { return double(a) * double(a); }
__device__ __forceinline__ double sqabs(double a) { return a * a; }
template <class R> __device__ __forceinline__ double sqabs(cplx<R> a) {
    return double(a.re) * double(a.re) + double(a.im) * double(a.im);
}

// conj(a)*b accumulated in double: returns {re, im}
struct d2 { double re, im; };
__device__ __forceinline__ d2 cdot(float a, float b)   { return {double(a) * double(b), 0.0}; }
__device__ __forceinline__ d2 cdot(double a, double b) { return {a * b, 0.0}; }
template <class R> __device__ __forceinline__ d2 cdot(cplx<R> a, cplx<R> b) {
    return {double(a.re) * double(b.re) + double(a.im) * double(b.im),
            double(a.re) * double(b.im) - double(a.im) * double(b.re)};
}

// The two diagonal-moment sums of one element pair (x, y_new) added to the running sums
// (A = sum |x|^2, B = sum conj(y_new) x): a real or complex column fills (a_re, b_re, b_im); a
// column pair puts column 2p's sums in (a_re, b_re) and column 2p+1's in (a_im, b_im).
template <class T>
__device__ __forceinline__ void add_diag_sums(T ynew, T x, double& a_re, double& a_im, double& b_re, double& b_im) {
    a_re += sqabs(x);
    const d2 p = cdot(ynew, x);
    b_re += p.re;
    b_im += p.im;
}
__device__ __forceinline__ void add_diag_sums(dpair ynew, dpair x, double& a_re, double& a_im, double& b_re,
                                              double& b_im) {
    a_re += x.a * x.a;
    a_im += x.b * x.b;
    b_re += ynew.a * x.a;
    b_im += ynew.b * x.b;
}

// Non-temporal (streaming) loads and stores: data touched once per step, kept from evicting
// reused data in L2.
__device__ __forceinline__ float nt_load(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double nt_load(const double* p) { return __builtin_nontemporal_load(p); }
template <class R> __device__ __forceinline__ cplx<R> nt_load(const cplx<R>* p) {
    const R* q = reinterpret_cast<const R*>(p);
    return {__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1)};
}
__device__ __forceinline__ dpair nt_load(const dpair* p) {
    const double* q = reinterpret_cast<const double*>(p);
    return {__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1)};
}
__device__ __forceinline__ void nt_store(dpair v, dpair* p) {
    double* q = reinterpret_cast<double*>(p);
    __builtin_nontemporal_store(v.a, q);
    __builtin_nontemporal_store(v.b, q + 1);
}
__device__ __forceinline__ void nt_store(float v, float* p) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void nt_store(double v, double* p) { __builtin_nontemporal_store(v, p); }
template <class R> __device__ __forceinline__ void nt_store(cplx<R> v, cplx<R>* p) {
    R* q = reinterpret_cast<R*>(p);
    __builtin_nontemporal_store(v.re, q);
    __builtin_nontemporal_store(v.im, q + 1);
}


// ---- three-piece bf16 split (fp32-accurate operands for bf16 MFMA) ----
// x = x0 + x1 + x2 + O(2^-24 x): x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1), the
// remainders exact in the input precision; pieces returned as 16-bit patterns.
__device__ __forceinline__ uint32_t bf16_bits_rn(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;   // round to nearest even (finite inputs)
}
__device__ __forceinline__ float bf16_value(uint32_t h) { return __uint_as_float(h << 16); }
template <class R>
__device__ __forceinline__ void split3(R v, uint16_t& p0, uint16_t& p1, uint16_t& p2) {
    const uint32_t h0 = bf16_bits_rn((float)v);
    const R r1 = v - (R)bf16_value(h0);
    const uint32_t h1 = bf16_bits_rn((float)r1);
    const R r2 = r1 - (R)bf16_value(h1);
    const uint32_t h2 = bf16_bits_rn((float)r2);
    p0 = (uint16_t)h0;
    p1 = (uint16_t)h1;
    p2 = (uint16_t)h2;
}
// Write the pieces of v (column col, row row) into split planes: plane ((col * NP + p) * 3 + j)
// of `base` (stride `plane` elements) holds piece j of part p (NP = 1 real, 2 complex: re, im).
template <class R>
__device__ __forceinline__ void store_split(uint16_t* base, int64_t plane, int col, int64_t row, R v) {
    uint16_t a, b, c;
    split3<R>(v, a, b, c);
    uint16_t* q = base + (int64_t)col * 3 * plane + row;
    q[0] = a;
    q[plane] = b;
    q[2 * plane] = c;
}
template <class R>
__device__ __forceinline__ void store_split(uint16_t* base, int64_t plane, int col, int64_t row, cplx<R> v) {
    uint16_t a, b, c;
    uint16_t* q = base + (int64_t)col * 6 * plane + row;
    split3<R>(v.re, a, b, c);
    q[0] = a;
    q[plane] = b;
    q[2 * plane] = c;
    split3<R>(v.im, a, b, c);
    q[3 * plane] = a;
    q[4 * plane] = b;
    q[5 * plane] = c;
}

}  // namespace kpmx
