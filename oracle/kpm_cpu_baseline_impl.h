/* ORACLE -- test infrastructure only.  Template body of kpm_cpu_baseline.c (BL_T, BL_W, BL_SFX, BL_CANON). */
#define BL_CAT2(a, b) a##_##b
#define BL_CAT(a, b) BL_CAT2(a, b)
#define BFN(name) BL_CAT(name, BL_SFX)

typedef struct {
    int64_t n, pitch, width;
    const BL_T* data;        /* pitch x width, column-major */
    const int32_t* idx;      /* pitch x width, column-major */
    const int32_t* perm;     /* original index -> optimized index (NULL = identity) */
    const int32_t* borders;  /* BFS slice borders; last == n */
    int64_t nborders;
    int64_t M, R, nthreads;
    int interleaved;
    /* shared state */
    pthread_mutex_t lock;
    bl_mt gen;
    int64_t next_job, num_batches, num_singles;
    BL_T* sum;               /* M accumulated moments */
} BFN(bl_job);

/* kpm_spmv_diagonal, ELL block form, W columns (kernel_polynomial.hpp:287-323 semantics). */
static void BFN(bl_step)(const BFN(bl_job)* J, int64_t w, int64_t start, int64_t end,
                         const BL_T* x, BL_T* y, BL_T* m2, BL_T* m3) {
    const int64_t last = J->width - 1;
    for (int64_t k = 0; k < last; ++k) {
        const BL_T* a = J->data + k * J->pitch;
        const int32_t* ix = J->idx + k * J->pitch;
        for (int64_t row = start; row < end; ++row) {
            const BL_T av = a[row];
            const BL_T* xr = x + (int64_t)ix[row] * w;
            BL_T* yr = y + row * w;
            if (k == 0) { for (int64_t j = 0; j < w; ++j) yr[j] = av * xr[j] - yr[j]; }
            else        { for (int64_t j = 0; j < w; ++j) yr[j] = av * xr[j] + yr[j]; }
        }
    }
    const BL_T* a = J->data + last * J->pitch;
    const int32_t* ix = J->idx + last * J->pitch;
    BL_T s2[BL_W], s3[BL_W];
    for (int64_t j = 0; j < w; ++j) { s2[j] = 0; s3[j] = 0; }
    for (int64_t row = start; row < end; ++row) {
        const BL_T av = a[row];
        const BL_T* xr = x + (int64_t)ix[row] * w;
        const BL_T* xo = x + row * w;
        BL_T* yr = y + row * w;
        for (int64_t j = 0; j < w; ++j) {
            BL_T c = (last == 0) ? -yr[j] : yr[j];
            BL_T r2 = av * xr[j] + c;
            s2[j] += xo[j] * xo[j];
            s3[j] += r2 * xo[j];
            yr[j] = r2;
        }
    }
    for (int64_t j = 0; j < w; ++j) { m2[j] += s2[j]; m3[j] += s3[j]; }
}

static void BFN(bl_batch)(BFN(bl_job)* J, int64_t w, BL_T* r0, BL_T* r1, BL_T* mom /* M x w */) {
    const int64_t n = J->n, M = J->M;
    /* draw w columns under the starter lock, then permute into optimized order */
    pthread_mutex_lock(&J->lock);
    for (int64_t j = 0; j < w; ++j) {
        for (int64_t i = 0; i < n; ++i) {
            BL_T u = (BL_T)BL_CANON(&J->gen);
            int64_t dst = J->perm ? J->perm[i] : i;
            r0[dst * w + j] = (u < 0.5f) ? (BL_T)-1 : (BL_T)1;
        }
    }
    pthread_mutex_unlock(&J->lock);

    /* make_r1, ELL form: r1 = sum_k a*r0[idx]*0.5, column by column */
    memset(r1, 0, sizeof(BL_T) * (size_t)(n * w));
    for (int64_t k = 0; k < J->width; ++k) {
        const BL_T* a = J->data + k * J->pitch;
        const int32_t* ix = J->idx + k * J->pitch;
        for (int64_t row = 0; row < n; ++row)
            for (int64_t j = 0; j < w; ++j) r1[row * w + j] += a[row] * r0[(int64_t)ix[row] * w + j] * (BL_T)0.5;
    }
    BL_T m0[BL_W], m1[BL_W];
    for (int64_t j = 0; j < w; ++j) {
        BL_T s = 0, d = 0;
        for (int64_t i = 0; i < n; ++i) { s += r0[i * w + j] * r0[i * w + j]; d += r1[i * w + j] * r0[i * w + j]; }
        m0[j] = s * (BL_T)0.5; m1[j] = d;
        mom[0 * w + j] = m0[j]; mom[1 * w + j] = m1[j];
    }

    if (!J->interleaved) {
        for (int64_t s = 2; s <= M / 2; ++s) {
            BL_T m2[BL_W] = {0}, m3[BL_W] = {0};
            BFN(bl_step)(J, w, 0, n, r1, r0, m2, m3);
            for (int64_t j = 0; j < w; ++j) {
                mom[(2 * (s - 1)) * w + j] = (BL_T)2 * (m2[j] - m0[j]);
                mom[(2 * (s - 1) + 1) * w + j] = (BL_T)2 * m3[j] - m1[j];
            }
            BL_T* t = r0; r0 = r1; r1 = t;
        }
    } else {
        const int64_t last = J->nborders - 1;
        for (int64_t s = 2; s <= M / 2; s += 2) {
            BL_T m2[BL_W] = {0}, m3[BL_W] = {0}, m4[BL_W] = {0}, m5[BL_W] = {0};
            int64_t start0 = 0, start1 = 0;
            for (int64_t k = 0; k <= last; ++k) {
                int64_t end0 = J->borders[k];
                int64_t end1 = (k == last) ? J->borders[last] : start0;
                BFN(bl_step)(J, w, start0, end0, r1, r0, m2, m3);
                BFN(bl_step)(J, w, start1, end1, r0, r1, m4, m5);
                start1 = end1; start0 = end0;
            }
            for (int64_t j = 0; j < w; ++j) {
                mom[(2 * (s - 1)) * w + j] = (BL_T)2 * (m2[j] - m0[j]);
                mom[(2 * (s - 1) + 1) * w + j] = (BL_T)2 * m3[j] - m1[j];
                mom[(2 * s) * w + j] = (BL_T)2 * (m4[j] - m0[j]);
                mom[(2 * s + 1) * w + j] = (BL_T)2 * m5[j] - m1[j];
            }
        }
    }
}

static void* BFN(bl_worker)(void* arg) {
    BFN(bl_job)* J = (BFN(bl_job)*)arg;
    const int64_t n = J->n, M = J->M;
    BL_T* r0 = (BL_T*)malloc(sizeof(BL_T) * (size_t)(n * BL_W));
    BL_T* r1 = (BL_T*)malloc(sizeof(BL_T) * (size_t)(n * BL_W));
    BL_T* mom = (BL_T*)malloc(sizeof(BL_T) * (size_t)(M * BL_W));
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int64_t job = J->next_job++;
        pthread_mutex_unlock(&J->lock);
        if (job >= J->num_batches + J->num_singles) break;
        int64_t w = (job < J->num_batches) ? BL_W : 1;
        BFN(bl_batch)(J, w, r0, r1, mom);
        /* BatchAccumulator: sum the valid columns (columns beyond R in the last batch dropped) */
        pthread_mutex_lock(&J->lock);
        int64_t first = (job < J->num_batches) ? job * BL_W : J->num_batches * BL_W + (job - J->num_batches);
        int64_t cols = (J->R - first) < w ? (J->R - first) : w;
        for (int64_t m = 0; m < M; ++m)
            for (int64_t j = 0; j < cols; ++j) J->sum[m] += mom[m * w + j];
        pthread_mutex_unlock(&J->lock);
    }
    free(r0); free(r1); free(mom);
    return NULL;
}

/* Mean DOS moments over R random vectors.  Returns the moments_timer-scope seconds (starter
 * draws + recurrence), or a negative value on error. */
double BFN(bl_dos_moments)(int64_t n, int64_t pitch, int64_t width, const BL_T* data, const int32_t* idx,
                           const int32_t* perm, const int32_t* borders, int64_t nborders,
                           int64_t M, int64_t R, int64_t nthreads, int interleaved, BL_T* out) {
    if (M < 2 || M % 2 || (interleaved && (M - 2) % 4)) return -1.0;
    BFN(bl_job) J;
    memset(&J, 0, sizeof(J));
    J.n = n; J.pitch = pitch; J.width = width; J.data = data; J.idx = idx; J.perm = perm;
    J.borders = borders; J.nborders = nborders; J.M = M; J.R = R; J.nthreads = nthreads;
    J.interleaved = interleaved;
    pthread_mutex_init(&J.lock, NULL);
    bl_seed(&J.gen, 5489u);
    J.num_batches = R / BL_W;
    J.num_singles = R % BL_W;
    if (J.num_singles > nthreads * BL_W / 2) { J.num_batches += 1; J.num_singles = 0; }
    J.sum = (BL_T*)calloc((size_t)M, sizeof(BL_T));

    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int64_t t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, BFN(bl_worker), &J);
    for (int64_t t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(th);

    for (int64_t m = 0; m < M; ++m) out[m] = (R != 1) ? J.sum[m] / (BL_T)R : J.sum[m];
    free(J.sum);
    pthread_mutex_destroy(&J.lock);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

#undef BFN
#undef BL_CAT
#undef BL_CAT2
#undef BL_T
#undef BL_W
#undef BL_SFX
#undef BL_CANON
