/*
 * ORACLE -- test infrastructure only.  Template body of kpm_oracle.c, included once per scalar
 * type with ORC_T (scalar), ORC_R (real type), ORC_SFX (name suffix), ORC_CPLX (0/1) defined.
 */
#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define FN(name) ORC_CAT(name, ORC_SFX)

#if ORC_CPLX
static inline ORC_T FN(orc_mul)(ORC_T a, ORC_T b) {
    /* compute::detail::mul (cppcore/include/compute/detail.hpp:16-20): raw complex product */
    ORC_R ar = __real__ a, ai = __imag__ a, br = __real__ b, bi = __imag__ b;
    ORC_T r;
    __real__ r = ar * br - ai * bi;
    __imag__ r = ar * bi + ai * br;
    return r;
}
static inline ORC_T FN(orc_conj)(ORC_T a) { ORC_T r; __real__ r = __real__ a; __imag__ r = -__imag__ a; return r; }
static inline ORC_R FN(orc_sq)(ORC_T a) { return __real__ a * __real__ a + __imag__ a * __imag__ a; }
#else
static inline ORC_T FN(orc_mul)(ORC_T a, ORC_T b) { return a * b; }
static inline ORC_T FN(orc_conj)(ORC_T a) { return a; }
static inline ORC_R FN(orc_sq)(ORC_T a) { return a * a; }
#endif

/* Random starter columns, drawn one after another from one generator seeded 5489
 * (Starter.cpp:59-82; Starter.hpp:42-50).  Output is vector-contiguous: out[v*n + i]. */
void FN(orc_random_starter)(int64_t n, int64_t nvec, ORC_T* out) {
    orc_mt g;
    orc_mt_seed(&g, 5489u);
    for (int64_t v = 0; v < nvec; ++v) {
        for (int64_t i = 0; i < n; ++i) {
#if ORC_CPLX
            ORC_R phase = (sizeof(ORC_R) == 8) ? (ORC_R)orc_canonical_f64(&g) : (ORC_R)orc_canonical_f32(&g);
            ORC_R theta = (ORC_R)ORC_TWO_PI_F * phase;
            ORC_T z;
            __real__ z = (ORC_R)0; __imag__ z = theta;
            out[v * n + i] = (sizeof(ORC_R) == 8) ? (ORC_T)cexp((double _Complex)z) : (ORC_T)cexpf((float _Complex)z);
#else
            ORC_R x = (sizeof(ORC_R) == 8) ? (ORC_R)orc_canonical_f64(&g) : (ORC_R)orc_canonical_f32(&g);
            out[v * n + i] = (x < 0.5f) ? (ORC_R)-1 : (ORC_R)1;
#endif
        }
    }
}

/* y = A*x (plain CSR product, Eigen row-major semantics: in-row order accumulation) */
void FN(orc_csr_matvec)(int64_t n, const int32_t* indptr, const int32_t* indices, const ORC_T* data,
                        const ORC_T* x, ORC_T* y) {
    for (int64_t row = 0; row < n; ++row) {
        ORC_T r = 0;
        for (int32_t k = indptr[row]; k < indptr[row + 1]; ++k) r += FN(orc_mul)(data[k], x[indices[k]]);
        y[row] = r;
    }
}

/* make_r1: r1 = 0.5 * H2 * r0 (Starter.hpp:52-70) */
static void FN(orc_make_r1)(int64_t n, const int32_t* indptr, const int32_t* indices, const ORC_T* data,
                            const ORC_T* r0, ORC_T* r1) {
    for (int64_t row = 0; row < n; ++row) {
        ORC_T tmp = 0;
        for (int32_t k = indptr[row]; k < indptr[row + 1]; ++k) tmp += FN(orc_mul)(data[k], r0[indices[k]]);
        r1[row] = tmp * (ORC_R)0.5;
    }
}

/* kpm_spmv, CSR: y = H2 * x - y (kernel_polynomial.hpp:18-32) */
void FN(orc_kpm_spmv)(int64_t n, const int32_t* indptr, const int32_t* indices, const ORC_T* data,
                      const ORC_T* x, ORC_T* y) {
    for (int64_t row = 0; row < n; ++row) {
        ORC_T r = 0;
        for (int32_t k = indptr[row]; k < indptr[row + 1]; ++k) r += FN(orc_mul)(data[k], x[indices[k]]);
        y[row] = r - y[row];
    }
}

/* sum conj(a_i) b_i -- Eigen's a.dot(b) */
static ORC_T FN(orc_dotc)(int64_t n, const ORC_T* a, const ORC_T* b) {
    ORC_T s = 0;
    for (int64_t i = 0; i < n; ++i) s += FN(orc_mul)(FN(orc_conj)(a[i]), b[i]);
    return s;
}

static ORC_R FN(orc_sqnorm)(int64_t n, const ORC_T* a) {
    ORC_R s = 0;
    for (int64_t i = 0; i < n; ++i) s += FN(orc_sq)(a[i]);
    return s;
}

/* Diagonal moments mu_n = <r|T_n(H~)|r> for `nvec` starters (vector-contiguous r0s, nvec x n),
 * calc_moments::basic diagonal (calc_moments.hpp:35-51) with the DiagonalCollector
 * (collectors.cpp:5-15).  mu is nvec x M; M must be even. */
int FN(orc_diag_moments)(int64_t n, const int32_t* indptr, const int32_t* indices, const ORC_T* data,
                         int64_t nvec, const ORC_T* r0s, int64_t M, ORC_T* mu) {
    if (M < 2 || (M % 2) != 0) return -1;
    ORC_T* r0 = (ORC_T*)malloc(sizeof(ORC_T) * (size_t)(n > 0 ? n : 1));
    ORC_T* r1 = (ORC_T*)malloc(sizeof(ORC_T) * (size_t)(n > 0 ? n : 1));
    if (!r0 || !r1) { free(r0); free(r1); return -2; }
    for (int64_t v = 0; v < nvec; ++v) {
        ORC_T* out = mu + v * M;
        memcpy(r0, r0s + v * n, sizeof(ORC_T) * (size_t)n);
        FN(orc_make_r1)(n, indptr, indices, data, r0, r1);
        ORC_T m0 = (ORC_T)(FN(orc_sqnorm)(n, r0) * (ORC_R)0.5);
        ORC_T m1 = FN(orc_dotc)(n, r1, r0);
        out[0] = m0;
        out[1] = m1;
        for (int64_t k = 2; k <= M / 2; ++k) {
            /* x = r1, y = r0: r0 <- H2 r1 - r0; m2 = |r1|^2; m3 = <r0_new|r1> */
            FN(orc_kpm_spmv)(n, indptr, indices, data, r1, r0);
            ORC_T m2 = (ORC_T)FN(orc_sqnorm)(n, r1);
            ORC_T m3 = FN(orc_dotc)(n, r0, r1);
            out[2 * (k - 1)] = (ORC_R)2 * (m2 - m0);
            out[2 * (k - 1) + 1] = (ORC_R)2 * m3 - m1;
            ORC_T* t = r0; r0 = r1; r1 = t;
        }
    }
    free(r0); free(r1);
    return 0;
}

/* Off-diagonal recurrence (calc_moments.hpp:102-114) feeding one of three collectors:
 *   kind 0 Generic   : mu[n] = beta . (op r_n)  (op optional), mu[0] halved     (collectors.cpp:36-58)
 *   kind 1 MultiUnit : mu[i*M + n] = r_n[dest_i], mu[i*M + 0] halved           (collectors.cpp:60-75)
 *   kind 2 Dense     : mu[n*n_rows + :] = op r_n (or r_n), row 0 halved         (collectors.cpp:77-106)
 */
int FN(orc_offdiag_moments)(int64_t n, const int32_t* indptr, const int32_t* indices, const ORC_T* data,
                            const ORC_T* r0_in, int64_t M, int kind,
                            const ORC_T* beta,
                            const int32_t* op_indptr, const int32_t* op_indices, const ORC_T* op_data,
                            int64_t ndest, const int32_t* dest,
                            ORC_T* mu) {
    if (M < 2) return -1;
    size_t bytes = sizeof(ORC_T) * (size_t)(n > 0 ? n : 1);
    ORC_T* r0 = (ORC_T*)malloc(bytes);
    ORC_T* r1 = (ORC_T*)malloc(bytes);
    ORC_T* tmp = (ORC_T*)malloc(bytes);
    if (!r0 || !r1 || !tmp) { free(r0); free(r1); free(tmp); return -2; }
    memcpy(r0, r0_in, sizeof(ORC_T) * (size_t)n);
    FN(orc_make_r1)(n, indptr, indices, data, r0, r1);

    for (int64_t m = 0; m < M; ++m) {
        const ORC_T* r;
        if (m == 0) {
            r = r0;
        } else if (m == 1) {
            r = r1;
        } else {
            FN(orc_kpm_spmv)(n, indptr, indices, data, r1, r0);  /* r0 = H2 r1 - r0 */
            ORC_T* t = r0; r0 = r1; r1 = t;                      /* swap, collect r1 */
            r = r1;
        }
        ORC_R half = (m == 0) ? (ORC_R)0.5 : (ORC_R)1;
        const ORC_T* v = r;
        if (op_data && kind != 1) {
            FN(orc_csr_matvec)(n, op_indptr, op_indices, op_data, r, tmp);
            v = tmp;
        }
        if (kind == 0) {
            ORC_T s = FN(orc_dotc)(n, beta, v);
            mu[m] = (m == 0) ? s * 0.5f : s;
        } else if (kind == 1) {
            for (int64_t i = 0; i < ndest; ++i) mu[i * M + m] = r[dest[i]] * half;
        } else {
            for (int64_t i = 0; i < n; ++i) mu[m * n + i] = v[i] * half;
        }
    }
    free(r0); free(r1); free(tmp);
    return 0;
}

#undef FN
#undef ORC_CAT
#undef ORC_CAT2
#undef ORC_T
#undef ORC_R
#undef ORC_SFX
#undef ORC_CPLX
