/*
 * tci_oracle.c — CPU restatement of the reference TCI2 path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, never part of
 * it.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off).  Each function
 * cites the reference lines it follows (paths relative to the reference
 * repository root, TensorCrossInterpolation.jl v0.9.19).
 */
#include "tci_oracle.h"
#include "../include/tci_functors.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char g_err[256];
static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
const char *orc_last_error(void) { return g_err; }

/* Julia's max for Float64: NaN-propagating (Base.max). */
static double jl_max(double a, double b)
{
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

/* ===================================================================== */
/* rrLU — matrixlu.jl                                                      */
/* ===================================================================== */

/* submatrixargmax(abs2, A, k) — matrixlu.jl:1-32, 63-65.  Column-major scan
 * over rows k:m, cols k:n; strict '>' so the first maximum wins; the running
 * maximum starts at typemin(Float64) = -Inf, so NaN is never selected and an
 * all-NaN block returns (k,k).  Indices are 0-based here. */
void orc_submatrixargmax(const double *A, int64_t lda, int64_t m, int64_t n, int64_t k,
                         int64_t *r, int64_t *c)
{
    double best = -INFINITY;
    int64_t mr = k, mc = k;
    for (int64_t cc = k; cc < n; ++cc) {
        for (int64_t rr = k; rr < m; ++rr) {
            double a = A[rr + lda * cc];
            double v = a * a;
            if (v > best) { best = v; mr = rr; mc = cc; }
        }
    }
    *r = mr;
    *c = mc;
}

/* swaprow!/swapcol! (matrixlu.jl:98-112): full rows/columns. */
static void swaprow(double *A, int64_t lda, int64_t n, int64_t *perm, int64_t a, int64_t b)
{
    int64_t t = perm[a]; perm[a] = perm[b]; perm[b] = t;
    for (int64_t j = 0; j < n; ++j) {
        double x = A[a + lda * j]; A[a + lda * j] = A[b + lda * j]; A[b + lda * j] = x;
    }
}
static void swapcol(double *A, int64_t lda, int64_t m, int64_t *perm, int64_t a, int64_t b)
{
    int64_t t = perm[a]; perm[a] = perm[b]; perm[b] = t;
    for (int64_t i = 0; i < m; ++i) {
        double x = A[i + lda * a]; A[i + lda * a] = A[i + lda * b]; A[i + lda * b] = x;
    }
}

/* addpivot! (matrixlu.jl:114-136): swap, scale the pivot column (leftorth)
 * or row by true division, then the unfused rank-1 update
 * A[i,j] -= x[i]*y[j] over the trailing block, j outer / i inner. */
static void addpivot(double *A, int64_t m, int64_t n, int64_t *rowperm, int64_t *colperm,
                     int leftorth, int64_t k, int64_t r, int64_t c)
{
    const int64_t lda = m;
    swaprow(A, lda, n, rowperm, k, r);
    swapcol(A, lda, m, colperm, k, c);
    double piv = A[k + lda * k];
    if (leftorth) {
        for (int64_t i = k + 1; i < m; ++i) A[i + lda * k] = A[i + lda * k] / piv;
    } else {
        for (int64_t j = k + 1; j < n; ++j) A[k + lda * j] = A[k + lda * j] / piv;
    }
    for (int64_t j = k + 1; j < n; ++j) {
        double y = A[k + lda * j];
        for (int64_t i = k + 1; i < m; ++i) {
            double t = A[i + lda * k] * y;
            A[i + lda * j] = A[i + lda * j] - t;
        }
    }
}

/* work accounting for the roofline: Σ_{k<s}(m-k)(n-k) and Σ_{k=1..p}(m-k)(n-k) */
static _Thread_local double g_stat_arg, g_stat_schur;

/* rrlu (matrixlu.jl:217-225) → rrlu! (194-204) → _optimizerrlu! (141-181). */
int orc_rrlu(const double *A0, int64_t m, int64_t n, int64_t maxrank, double reltol,
             double abstol, int32_t leftorth, int64_t *rowperm, int64_t *colperm,
             double *Lout, double *Uout, int64_t ldu, int64_t *npivot_out, double *error_out)
{
    double *A = (double *)malloc(sizeof(double) * (size_t)(m > 0 && n > 0 ? m * n : 1));
    if (m > 0 && n > 0) memcpy(A, A0, sizeof(double) * (size_t)(m * n));
    for (int64_t i = 0; i < m; ++i) rowperm[i] = i + 1;
    for (int64_t j = 0; j < n; ++j) colperm[j] = j + 1;
    int64_t mr = maxrank;
    if (m < mr) mr = m;
    if (n < mr) mr = n;
    double maxerror = 0.0, error = NAN;
    int64_t npivot = 0;
    g_stat_arg = 0.0;
    g_stat_schur = 0.0;
    while (npivot < mr) {
        int64_t k = npivot, r, c;
        orc_submatrixargmax(A, m, m, n, k, &r, &c);
        g_stat_arg += (double)(m - k) * (double)(n - k);
        error = fabs(A[r + m * c]);
        if ((error < reltol * maxerror || error < abstol) && npivot > 0) break;
        maxerror = jl_max(maxerror, error);
        addpivot(A, m, n, rowperm, colperm, leftorth, k, r, c);
        npivot += 1;
        g_stat_schur += (double)(m - npivot) * (double)(n - npivot);
    }
    const int64_t p = npivot;
    /* L = tril(A[:,1:p]), U = triu(A[1:p,:]) (matrixlu.jl:162-163) */
    int nanL = 0, nanU = 0;
    for (int64_t j = 0; j < p; ++j)
        for (int64_t i = 0; i < m; ++i) {
            double v = i >= j ? A[i + m * j] : 0.0;
            if (v != v) nanL = 1;
            Lout[i + m * j] = v;
        }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < p; ++i) {
            double v = i <= j ? A[i + m * j] : 0.0;
            if (v != v) nanU = 1;
            Uout[i + ldu * j] = v;
        }
    free(A);
    *npivot_out = p;
    /* unit diagonal (matrixlu.jl:170-174) */
    for (int64_t k = 0; k < p; ++k) {
        if (leftorth) Lout[k + m * k] = 1.0;
        else Uout[k + ldu * k] = 1.0;
    }
    int64_t mn = m < n ? m : n;
    if (npivot >= mn) error = 0.0;   /* matrixlu.jl:176-178 */
    *error_out = error;
    if (nanL) return fail(ORC_ERR_NAN_L, "lu.L contains NaNs");
    if (nanU) return fail(ORC_ERR_NAN_U, "lu.U contains NaNs");
    return ORC_OK;
}

/* ===================================================================== */
/* MatrixLUCI factors — matrixluci.jl:40-84                               */
/* ===================================================================== */

void orc_luci_left(int64_t m, int64_t n, int64_t p, int32_t leftorth, const int64_t *rowperm,
                   const int64_t *colperm, const double *L, const double *U, int64_t ldu,
                   double *out)
{
    (void)n; (void)colperm;
    double *R = (double *)calloc((size_t)(m * p > 0 ? m * p : 1), sizeof(double));
    if (leftorth) {
        /* colstimespivotinv (48-57): [I_p; L21·L11⁻¹], rows permuted back */
        for (int64_t k = 0; k < p; ++k) R[k + m * k] = 1.0;
        double *x = (double *)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
        for (int64_t i = p; i < m; ++i) {
            /* solve x·L11 = L[i,0:p] for x, back to front */
            for (int64_t j = p - 1; j >= 0; --j) {
                double s = L[i + m * j];
                for (int64_t t = j + 1; t < p; ++t) s = s - x[t] * L[t + m * j];
                x[j] = s / L[j + m * j];
            }
            for (int64_t j = 0; j < p; ++j) R[i + m * j] = x[j];
        }
        free(x);
    } else {
        /* colmatrix (40-42): left(lu) · U[:,1:p], computed row by row in
         * position order (rows are permuted back below) */
        for (int64_t i = 0; i < m; ++i)
            for (int64_t j = 0; j < p; ++j) {
                double s = 0.0;
                int64_t tmax = i < j ? i : j;
                for (int64_t t = 0; t <= tmax && t < p; ++t) s = s + L[i + m * t] * U[t + ldu * j];
                R[i + m * j] = s;
            }
    }
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < p; ++j) out[(rowperm[i] - 1) + m * j] = R[i + m * j];
    free(R);
}

void orc_luci_right(int64_t m, int64_t n, int64_t p, int32_t leftorth, const int64_t *rowperm,
                    const int64_t *colperm, const double *L, const double *U, int64_t ldu,
                    double *out)
{
    (void)rowperm;
    double *R = (double *)calloc((size_t)(p * n > 0 ? p * n : 1), sizeof(double));
    if (leftorth) {
        /* rowmatrix (44-46): L[1:p,:] · U */
        for (int64_t i = 0; i < p; ++i)
            for (int64_t j = 0; j < n; ++j) {
                double s = 0.0;
                int64_t tmax = i < j ? i : j;
                for (int64_t t = 0; t <= tmax; ++t) s = s + L[i + m * t] * U[t + ldu * j];
                R[i + p * j] = s;
            }
    } else {
        /* pivotinvtimesrows (59-68): [I_p  U11⁻¹·U12], columns permuted back */
        for (int64_t k = 0; k < p; ++k) R[k + p * k] = 1.0;
        double *x = (double *)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
        for (int64_t j = p; j < n; ++j) {
            for (int64_t i = p - 1; i >= 0; --i) {
                double s = U[i + ldu * j];
                for (int64_t t = i + 1; t < p; ++t) s = s - U[i + ldu * t] * x[t];
                x[i] = s / U[i + ldu * i];
            }
            for (int64_t i = 0; i < p; ++i) R[i + p * j] = x[i];
        }
        free(x);
    }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < p; ++i) out[i + p * (colperm[j] - 1)] = R[i + p * j];
    free(R);
}

/* ===================================================================== */
/* Function evaluation                                                    */
/* ===================================================================== */

static double feval(const orc_functor *f, const int32_t *x, int32_t L)
{
    if (f->kind == 0) return f->callback(x, L, f->user);
    uint8_t buf[256];
    for (int t = 0; t < L; ++t) buf[t] = (uint8_t)x[t];
    tci_point pt;
    pt.a = buf; pt.b = buf; pt.na = L; pt.L = L;
    return tci_functor_eval(f->kind, f->params, f->iparams, &pt);
}

/* _batchevaluate_dispatch (batcheval.jl:32-61): result[i, c, j] with the
 * left index fastest, then the central indices (first fastest), then right. */
void orc_batcheval(const orc_functor *f, int32_t L, const int32_t *localdims,
                   const int32_t *left, int64_t nI, int32_t nl, const int32_t *right,
                   int64_t nJ, int32_t nr, double *out)
{
    int32_t M = L - nl - nr;
    int64_t nc = 1;
    for (int t = 0; t < M; ++t) nc *= localdims[nl + t];
    int32_t x[256];
    for (int64_t j = 0; j < nJ; ++j)
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t i = 0; i < nI; ++i) {
                for (int t = 0; t < nl; ++t) x[t] = left[i * nl + t];
                int64_t cc = c;
                for (int t = 0; t < M; ++t) { x[nl + t] = (int32_t)(cc % localdims[nl + t]) + 1; cc /= localdims[nl + t]; }
                for (int t = 0; t < nr; ++t) x[nl + M + t] = right[j * nr + t];
                out[i + nI * (c + nc * j)] = feval(f, x, L);
            }
}

/* ===================================================================== */
/* Multi-index sets (Vector{MultiIndex})                                  */
/* ===================================================================== */

typedef struct mset {
    int32_t len, n, cap;
    int32_t *v;
} mset;

static void ms_init(mset *s, int32_t len) { s->len = len; s->n = 0; s->cap = 0; s->v = NULL; }
static void ms_free(mset *s) { free(s->v); s->v = NULL; s->n = s->cap = 0; }
static void ms_reserve(mset *s, int32_t n)
{
    if (n <= s->cap) return;
    int32_t c = s->cap ? s->cap : 4;
    while (c < n) c *= 2;
    s->v = (int32_t *)realloc(s->v, sizeof(int32_t) * (size_t)(c * (s->len > 0 ? s->len : 1)));
    s->cap = c;
}
static const int32_t *ms_at(const mset *s, int32_t i) { return s->v + (size_t)i * s->len; }
static int ms_eq(const int32_t *a, const int32_t *b, int32_t len)
{
    for (int t = 0; t < len; ++t) if (a[t] != b[t]) return 0;
    return 1;
}
static int32_t ms_find(const mset *s, const int32_t *x)
{
    for (int32_t i = 0; i < s->n; ++i) if (ms_eq(ms_at(s, i), x, s->len)) return i;
    return -1;
}
static void ms_push(mset *s, const int32_t *x)
{
    ms_reserve(s, s->n + 1);
    if (s->len > 0) memcpy(s->v + (size_t)s->n * s->len, x, sizeof(int32_t) * (size_t)s->len);
    s->n++;
}
/* pushunique! (util.jl:16-20) */
static void ms_pushunique(mset *s, const int32_t *x) { if (ms_find(s, x) < 0) ms_push(s, x); }
static void ms_copy(mset *dst, const mset *src)
{
    ms_free(dst);
    ms_init(dst, src->len);
    ms_reserve(dst, src->n);
    if (src->n && src->len) memcpy(dst->v, src->v, sizeof(int32_t) * (size_t)(src->n * src->len));
    dst->n = src->n;
}
/* kronecker(Iset, d) (tensorci2.jl:315-320): [is..., j], is fastest */
static void ms_kron_right(mset *out, const mset *I, int32_t d)
{
    ms_init(out, I->len + 1);
    ms_reserve(out, I->n * d);
    int32_t x[256];
    for (int32_t j = 1; j <= d; ++j)
        for (int32_t i = 0; i < I->n; ++i) {
            memcpy(x, ms_at(I, i), sizeof(int32_t) * (size_t)I->len);
            x[I->len] = j;
            ms_push(out, x);
        }
}
/* kronecker(d, Jset) (tensorci2.jl:322-327): [i, js...], i fastest */
static void ms_kron_left(mset *out, int32_t d, const mset *J)
{
    ms_init(out, J->len + 1);
    ms_reserve(out, J->n * d);
    int32_t x[256];
    for (int32_t j = 0; j < J->n; ++j)
        for (int32_t i = 1; i <= d; ++i) {
            x[0] = i;
            memcpy(x + 1, ms_at(J, j), sizeof(int32_t) * (size_t)J->len);
            ms_push(out, x);
        }
}
/* union(a, b): first occurrence order (Base.union on vectors) */
static void ms_union_inplace(mset *a, const mset *b)
{
    mset u;
    ms_init(&u, a->len);
    for (int32_t i = 0; i < a->n; ++i) ms_pushunique(&u, ms_at(a, i));
    if (b && b->len == a->len)
        for (int32_t i = 0; i < b->n; ++i) ms_pushunique(&u, ms_at(b, i));
    ms_free(a);
    *a = u;
}
static void ms_select(mset *dst, const mset *src, const int64_t *perm1, int64_t p)
{
    mset o;
    ms_init(&o, src->len);
    ms_reserve(&o, (int32_t)p);
    for (int64_t k = 0; k < p; ++k) ms_push(&o, ms_at(src, (int32_t)(perm1[k] - 1)));
    ms_free(dst);
    *dst = o;
}

/* ===================================================================== */
/* TensorCI2 state — tensorci2.jl:6-40                                     */
/* ===================================================================== */

struct orc_tci {
    int32_t L;
    int32_t *d;
    mset *I, *J;                 /* I[b] prefixes (len b), J[b] suffixes (len L-1-b) */
    double **T;                  /* site tensors, NULL = invalidated */
    int64_t (*Td)[3];
    double *pivoterrors;
    int32_t npe, cape;
    double *bonderrors;          /* L-1 */
    double maxsample;
    int32_t has_hist;
    mset *hI, *hJ;               /* Iset_history[end], Jset_history[end] */
    int64_t *ranks, *nglob;
    double *errors;
    int32_t niter, capiter;
    double ws[6];
    uint64_t rng[4];
    int32_t pivotsearch;         /* 0 :full, 1 :rook (updatepivots!) */
};

orc_tci *orc_tci_new(int32_t L, const int32_t *localdims)
{
    if (L < 2) { fail(ORC_ERR_GENERIC, "localdims should have at least 2 elements!"); return NULL; }
    orc_tci *t = (orc_tci *)calloc(1, sizeof(orc_tci));
    t->L = L;
    t->d = (int32_t *)malloc(sizeof(int32_t) * (size_t)L);
    memcpy(t->d, localdims, sizeof(int32_t) * (size_t)L);
    t->I = (mset *)calloc((size_t)L, sizeof(mset));
    t->J = (mset *)calloc((size_t)L, sizeof(mset));
    t->hI = (mset *)calloc((size_t)L, sizeof(mset));
    t->hJ = (mset *)calloc((size_t)L, sizeof(mset));
    for (int b = 0; b < L; ++b) {
        ms_init(&t->I[b], b);
        ms_init(&t->J[b], L - 1 - b);
        ms_init(&t->hI[b], b);
        ms_init(&t->hJ[b], L - 1 - b);
    }
    t->T = (double **)calloc((size_t)L, sizeof(double *));
    t->Td = (int64_t(*)[3])calloc((size_t)L, sizeof(int64_t[3]));
    t->bonderrors = (double *)calloc((size_t)(L - 1), sizeof(double));
    orc_rng_seed(t->rng, 0);
    return t;
}

void orc_tci_free(orc_tci *t)
{
    if (!t) return;
    for (int b = 0; b < t->L; ++b) {
        ms_free(&t->I[b]); ms_free(&t->J[b]); ms_free(&t->hI[b]); ms_free(&t->hJ[b]);
        free(t->T[b]);
    }
    free(t->I); free(t->J); free(t->hI); free(t->hJ); free(t->T); free(t->Td);
    free(t->d); free(t->bonderrors); free(t->pivoterrors);
    free(t->ranks); free(t->nglob); free(t->errors);
    free(t);
}

static void invalidatesitetensors(orc_tci *t)   /* tensorci2.jl:90-95 */
{
    for (int b = 0; b < t->L; ++b) {
        free(t->T[b]);
        t->T[b] = NULL;
        t->Td[b][0] = t->Td[b][1] = t->Td[b][2] = 0;
    }
}

/* updatemaxsample! / maxabs (tensorci2.jl:397-399, util.jl:1-10) */
static void updatemaxsample(orc_tci *t, const double *v, int64_t n)
{
    double m = t->maxsample;
    for (int64_t i = 0; i < n; ++i) m = jl_max(fabs(m), fabs(v[i]));
    t->maxsample = m;
}

/* updateerrors! / updatepivoterror! (tensorci2.jl:132-169) */
static void updateerrors(orc_tci *t, int32_t b0, const double *errs, int32_t ne)
{
    t->bonderrors[b0] = errs[ne - 1];
    int32_t n = t->npe > ne ? t->npe : ne;
    if (n > t->cape) {
        t->cape = n * 2;
        t->pivoterrors = (double *)realloc(t->pivoterrors, sizeof(double) * (size_t)t->cape);
    }
    for (int32_t i = 0; i < n; ++i) {
        double a = i < t->npe ? t->pivoterrors[i] : 0.0;
        double b = i < ne ? errs[i] : 0.0;
        t->pivoterrors[i] = jl_max(a, b);
    }
    t->npe = n;
}

/* addglobalpivots! (tensorci2.jl:193-213) */
int orc_addglobalpivots(orc_tci *t, int32_t npiv, const int32_t *pivots)
{
    for (int32_t q = 0; q < npiv; ++q) {
        const int32_t *p = pivots + (size_t)q * t->L;
        for (int b = 0; b < t->L; ++b) {
            ms_pushunique(&t->I[b], p);
            ms_pushunique(&t->J[b], p + b + 1);
        }
    }
    if (npiv > 0) invalidatesitetensors(t);
    return ORC_OK;
}

int orc_tci_init(orc_tci *t, const orc_functor *f, int32_t npiv, const int32_t *pivots)
{
    orc_addglobalpivots(t, npiv, pivots);
    /* maxsamplevalue = maximum(abs, f(x) for x in initialpivots) (49-50) */
    double m = -INFINITY;
    for (int32_t q = 0; q < npiv; ++q) {
        double v = fabs(feval(f, pivots + (size_t)q * t->L, t->L));
        m = (q == 0) ? v : jl_max(m, v);
    }
    t->maxsample = m;
    if (!(fabs(m) > 0.0)) return fail(ORC_ERR_ZERO_MAXSAMPLE, "maxsamplevalue is zero!");
    invalidatesitetensors(t);
    return ORC_OK;
}

/* filltensor (tensorci2.jl:290-312) over sets; out must hold nI·nc·nJ */
static void filltensor(orc_tci *t, const orc_functor *f, const mset *I, const mset *J,
                       double *out, int64_t *nc_out)
{
    int32_t M = t->L - I->len - J->len;
    int64_t nc = 1;
    for (int q = 0; q < M; ++q) nc *= t->d[I->len + q];
    *nc_out = nc;
    if (I->n == 0 || J->n == 0) return;
    orc_batcheval(f, t->L, t->d, I->v, I->n, I->len, J->v, J->n, J->len, out);
    t->ws[0] += (double)I->n * (double)nc * (double)J->n;
    t->ws[5] += 1.0;
}

static void log_rrlu(orc_tci *t, int64_t m, int64_t n, int64_t p)
{
    t->ws[1] += 8.0 * g_stat_arg;
    t->ws[2] += 8.0 * (double)m * (double)n + 16.0 * g_stat_schur + 8.0 * (double)p * (double)(m + n);
    t->ws[3] += 2.0 * g_stat_schur;
    t->ws[4] += 1.0;
}

typedef struct luci_res {
    int64_t m, n, p;
    int64_t *rowperm, *colperm;
    double *L, *U;
    int64_t ldu;
    double error;
} luci_res;

static int run_luci(orc_tci *t, const double *Pi, int64_t m, int64_t n, double reltol,
                    double abstol, int64_t maxrank, int leftorth, luci_res *r)
{
    int64_t cap = m < n ? m : n;
    r->m = m; r->n = n;
    r->rowperm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    r->colperm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    r->L = (double *)malloc(sizeof(double) * (size_t)(m * cap + 1));
    r->U = (double *)malloc(sizeof(double) * (size_t)(cap * n + 1));
    r->ldu = cap > 0 ? cap : 1;
    int st = orc_rrlu(Pi, m, n, maxrank, reltol, abstol, leftorth, r->rowperm, r->colperm,
                      r->L, r->U, r->ldu, &r->p, &r->error);
    log_rrlu(t, m, n, r->p);
    return st;
}
static void free_luci(luci_res *r) { free(r->rowperm); free(r->colperm); free(r->L); free(r->U); }

/* pivoterrors(lu) = [|diag|..., error] (matrixlu.jl:394-416): the diagonal
 * of U (leftorth) or of L (otherwise) before the unit diagonal is imposed,
 * i.e. the pivot values, which rrlu leaves on the non-unit factor. */
static double *luci_pivoterrors(const luci_res *r, int leftorth)
{
    double *e = (double *)malloc(sizeof(double) * (size_t)(r->p + 1));
    for (int64_t k = 0; k < r->p; ++k)
        e[k] = fabs(leftorth ? r->U[k + r->ldu * k] : r->L[k + r->m * k]);
    e[r->p] = r->error;
    return e;
}

/* ===================================================================== */
/* Rook pivot search — arrlu (matrixlu.jl:227-293) over a lazily         */
/* evaluated matrix (SubMatrix, tensorci2.jl:476-503)                    */
/* ===================================================================== */

static int32_t rng_randint(uint64_t s[4], int32_t d);
typedef void (*log_fn)(void *ctx, int64_t m, int64_t n, int64_t p);

/* out[i + nr·j] = A(rows[i], cols[j]) (0-based positions), |out| folded
 * into the caller's running maximum */
typedef void (*sub_eval_fn)(void *ctx, const int64_t *rows, int64_t nr, const int64_t *cols, int64_t nc,
                            double *out);

/* pushrandomsubset!(S, 0:N-1, max(1, |S|)) (util.jl:36-58): randomsubset of
 * setdiff(1:N, S) (ascending), each draw rand(1:length(c)) then deleteat! */
static void push_random_subset(uint64_t rng[4], int64_t *S, int64_t *ns, int64_t N)
{
    int64_t k = *ns > 1 ? *ns : 1;
    char *in = (char *)calloc((size_t)(N + 1), 1);
    for (int64_t e = 0; e < *ns; ++e) in[S[e]] = 1;
    int64_t *c = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int64_t nc = 0;
    for (int64_t q = 0; q < N; ++q)
        if (!in[q]) c[nc++] = q;
    if (k > nc) k = nc;
    for (int64_t i = 0; i < k; ++i) {
        const int64_t idx = rng_randint(rng, (int32_t)nc) - 1;
        S[(*ns)++] = c[idx];
        memmove(c + idx, c + idx + 1, sizeof(int64_t) * (size_t)(nc - idx - 1));
        nc -= 1;
    }
    free(c);
    free(in);
}

typedef struct rook_res {
    int64_t p;
    int64_t *rowperm, *colperm;   /* 1-based, m and n entries */
    double *pv;                   /* the p pivot values (diag of U or L) */
    double error;
    int64_t ms_last, ns_last;     /* shape of the last factored submatrix */
    double *L, *U;                /* optional: m×p and p×n (matrixlu.jl:270-290) */
} rook_res;

static void free_rook(rook_res *r) { free(r->rowperm); free(r->colperm); free(r->pv); free(r->L); free(r->U); }

/* arrlu: I0/J0 (0-based positions, capacity m and n) are updated in place.
 * The factorization's permutations persist between the submatrices as in
 * the reference (swaps of rowpermutation/colpermutation entries < the
 * submatrix size), so a submatrix whose rows or columns are not the first
 * entries of those permutations yields positions, as the reference does.
 * want_lu: also the full L and U of the finish (cols2Lmatrix!/rows2Umatrix!). */
static int arrlu_core(sub_eval_fn ev, void *ectx, uint64_t rng[4], int64_t m, int64_t n, int64_t *I0,
                      int64_t nI0, int64_t *J0, int64_t nJ0, int64_t maxrank, double reltol, double abstol,
                      int leftorth, int numrookiter, int want_lu, rook_res *r, log_fn lg, void *lctx)
{
    if (maxrank > m) maxrank = m;
    if (maxrank > n) maxrank = n;
    int64_t *RP = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    int64_t *CP = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (size_t)((m > n ? m : n) + 1));
    for (int64_t i = 0; i < m; ++i) RP[i] = i;
    for (int64_t j = 0; j < n; ++j) CP[j] = j;
    const int64_t cap = m < n ? m : n;
    double *S = (double *)malloc(sizeof(double) * (size_t)(m * n + 1));
    int64_t *rl = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    int64_t *cl = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    double *Lf = (double *)malloc(sizeof(double) * (size_t)(m * (cap > 0 ? cap : 1) + 1));
    double *Uf = (double *)malloc(sizeof(double) * (size_t)((cap > 0 ? cap : 1) * n + 1));
    double *pv = (double *)malloc(sizeof(double) * (size_t)(cap + 1));
    int islowrank = 0, st = ORC_OK;
    int64_t p = 0, ms = 0, ns = 0;
    double err = NAN;
    for (;;) {
        if (leftorth) push_random_subset(rng, J0, &nJ0, n);
        else push_random_subset(rng, I0, &nI0, m);
        for (int it = 1; it <= numrookiter; ++it) {
            const int colmove = ((it % 2 == 0) == (leftorth != 0));
            const int64_t *ri = colmove ? I0 : RP, *ci = colmove ? CP : J0;
            ms = colmove ? nI0 : m;
            ns = colmove ? n : nJ0;
            ev(ectx, ri, ms, ci, ns, S);
            const int64_t ldu = (ms < ns ? ms : ns) > 0 ? (ms < ns ? ms : ns) : 1;
            st = orc_rrlu(S, ms, ns, maxrank, reltol, abstol, leftorth, rl, cl, Lf, Uf, ldu, &p, &err);
            if (lg) lg(lctx, ms, ns, p);
            if (st != ORC_OK) goto done;
            for (int64_t k = 0; k < p; ++k) pv[k] = leftorth ? Uf[k + ldu * k] : Lf[k + ms * k];
            /* the swaps act on the persistent permutations' first ms / ns entries */
            for (int64_t i = 0; i < ms; ++i) tmp[i] = RP[rl[i] - 1];
            memcpy(RP, tmp, sizeof(int64_t) * (size_t)ms);
            for (int64_t j = 0; j < ns; ++j) tmp[j] = CP[cl[j] - 1];
            memcpy(CP, tmp, sizeof(int64_t) * (size_t)ns);
            islowrank |= p < (ms < ns ? ms : ns);
            int same = (p == nI0 && p == nJ0);
            for (int64_t k = 0; same && k < p; ++k) same = RP[k] == I0[k] && CP[k] == J0[k];
            if (same) break;
            for (int64_t k = 0; k < p; ++k) { J0[k] = CP[k]; I0[k] = RP[k]; }
            nJ0 = p;
            nI0 = p;
        }
        if (islowrank || nI0 >= maxrank) break;
    }
    /* the finish: rows outside I0 against J0, columns outside J0 against I0
     * (their values enter the running maximum as the reference's SubMatrix
     * calls do) */
    {
        double *L2 = NULL, *U2 = NULL;
        int64_t n2 = 0;
        char *in = (char *)calloc((size_t)((m > n ? m : n) + 1), 1);
        if (ms < m) {
            for (int64_t e = 0; e < nI0; ++e) in[I0[e]] = 1;
            for (int64_t i = 0; i < m; ++i) if (!in[i]) tmp[n2++] = i;
            L2 = (double *)malloc(sizeof(double) * (size_t)(n2 * nJ0 + 1));
            ev(ectx, tmp, n2, J0, nJ0, L2);
            for (int64_t i = 0; i < n2; ++i) RP[nI0 + i] = tmp[i];   /* rowpermutation = vcat(I0, I2) */
            for (int64_t k = 0; k < nI0; ++k) RP[k] = I0[k];
            if (want_lu) {
                /* cols2Lmatrix!(L2, U[1:p,1:p]) (matrixlu.jl:314-335) */
                const int64_t ldu = (ms < ns ? ms : ns) > 0 ? (ms < ns ? ms : ns) : 1;
                for (int64_t k = 0; k < p; ++k) {
                    for (int64_t i = 0; i < n2; ++i) L2[i + n2 * k] = L2[i + n2 * k] / Uf[k + ldu * k];
                    for (int64_t j = k + 1; j < p; ++j)
                        for (int64_t i = 0; i < n2; ++i)
                            L2[i + n2 * j] = L2[i + n2 * j] - L2[i + n2 * k] * Uf[k + ldu * j];
                }
                r->L = (double *)calloc((size_t)(m * p + 1), sizeof(double));
                for (int64_t j = 0; j < p; ++j) {
                    for (int64_t i = 0; i < p; ++i) r->L[i + m * j] = Lf[i + ms * j];
                    for (int64_t i = 0; i < n2; ++i) r->L[p + i + m * j] = L2[i + n2 * j];
                }
            }
            memset(in, 0, (size_t)((m > n ? m : n) + 1));
        } else if (want_lu) {
            r->L = (double *)calloc((size_t)(m * p + 1), sizeof(double));
            memcpy(r->L, Lf, sizeof(double) * (size_t)(m * p));
        }
        n2 = 0;
        if (ns < n) {
            for (int64_t e = 0; e < nJ0; ++e) in[J0[e]] = 1;
            for (int64_t j = 0; j < n; ++j) if (!in[j]) tmp[n2++] = j;
            U2 = (double *)malloc(sizeof(double) * (size_t)(nI0 * n2 + 1));
            ev(ectx, I0, nI0, tmp, n2, U2);
            for (int64_t j = 0; j < n2; ++j) CP[nJ0 + j] = tmp[j];
            for (int64_t k = 0; k < nJ0; ++k) CP[k] = J0[k];
            if (want_lu) {
                /* rows2Umatrix!(U2, L[1:p,1:p]) (matrixlu.jl:337-358) */
                for (int64_t k = 0; k < p; ++k) {
                    for (int64_t j = 0; j < n2; ++j) U2[k + p * j] = U2[k + p * j] / Lf[k + ms * k];
                    for (int64_t j = 0; j < n2; ++j)
                        for (int64_t i = k + 1; i < p; ++i)
                            U2[i + p * j] = U2[i + p * j] - Lf[i + ms * k] * U2[k + p * j];
                }
                const int64_t ldu = (ms < ns ? ms : ns) > 0 ? (ms < ns ? ms : ns) : 1;
                r->U = (double *)calloc((size_t)(p * n + 1), sizeof(double));
                for (int64_t j = 0; j < p; ++j)
                    for (int64_t i = 0; i < p; ++i) r->U[i + p * j] = Uf[i + ldu * j];
                for (int64_t j = 0; j < n2; ++j)
                    for (int64_t i = 0; i < p; ++i) r->U[i + p * (p + j)] = U2[i + p * j];
            }
        } else if (want_lu) {
            const int64_t ldu = (ms < ns ? ms : ns) > 0 ? (ms < ns ? ms : ns) : 1;
            r->U = (double *)calloc((size_t)(p * n + 1), sizeof(double));
            for (int64_t j = 0; j < n; ++j)
                for (int64_t i = 0; i < p; ++i) r->U[i + p * j] = Uf[i + ldu * j];
        }
        free(in); free(L2); free(U2);
    }
done:
    r->p = p;
    r->error = err;
    r->ms_last = ms;
    r->ns_last = ns;
    r->rowperm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    r->colperm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < m; ++i) r->rowperm[i] = RP[i] + 1;
    for (int64_t j = 0; j < n; ++j) r->colperm[j] = CP[j] + 1;
    r->pv = pv;
    free(RP); free(CP); free(tmp); free(S); free(rl); free(cl); free(Lf); free(Uf);
    return st;
}

/* the explicit-matrix evaluation of the standalone arrlu */
typedef struct { const double *A; int64_t m; double maxabs; } mat_ctx;
static void mat_eval(void *c, const int64_t *rows, int64_t nr, const int64_t *cols, int64_t nc, double *out)
{
    mat_ctx *x = (mat_ctx *)c;
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i) {
            const double v = x->A[rows[i] + x->m * cols[j]];
            out[i + nr * j] = v;
            x->maxabs = jl_max(x->maxabs, fabs(v));
        }
}

int orc_arrlu(const double *A, int64_t m, int64_t n, const int64_t *I0in, int64_t nI0, const int64_t *J0in,
              int64_t nJ0, int64_t maxrank, double reltol, double abstol, int32_t leftorth, int32_t numrookiter,
              uint64_t seed, int64_t *rowperm, int64_t *colperm, double *Lout, double *Uout, int64_t *npivot,
              double *error)
{
    uint64_t rng[4];
    orc_rng_seed(rng, seed);
    int64_t *I0 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    int64_t *J0 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t e = 0; e < nI0; ++e) I0[e] = I0in[e] - 1;
    for (int64_t e = 0; e < nJ0; ++e) J0[e] = J0in[e] - 1;
    mat_ctx c = {A, m, 0.0};
    rook_res r;
    memset(&r, 0, sizeof r);
    int st = arrlu_core(mat_eval, &c, rng, m, n, I0, nI0, J0, nJ0, maxrank, reltol, abstol, leftorth,
                        numrookiter, 1, &r, NULL, NULL);
    if (st == ORC_OK) {
        memcpy(rowperm, r.rowperm, sizeof(int64_t) * (size_t)m);
        memcpy(colperm, r.colperm, sizeof(int64_t) * (size_t)n);
        if (r.L) memcpy(Lout, r.L, sizeof(double) * (size_t)(m * r.p));
        if (r.U) memcpy(Uout, r.U, sizeof(double) * (size_t)(r.p * n));
        *npivot = r.p;
        *error = r.error;
    }
    free_rook(&r);
    free(I0);
    free(J0);
    return st;
}

/* updatepivots! with pivotsearch=:full (tensorci2.jl:510-551, 599-607).
 * The setsitetensor! calls at 601-604 are made only when there are no
 * extras; their tensors are invalidated by the next updatepivots! or
 * recomputed by fillsitetensors!, so this restatement omits them. */
/* Pi entries f(Icombined[i] ⊕ Jcombined[j]) of the rook search, folded into
 * maxsamplevalue as SubMatrix does (tensorci2.jl:499-503 and 581) */
typedef struct { orc_tci *t; const orc_functor *f; const mset *I, *J; } tci_sub_ctx;
static void tci_sub_eval(void *c, const int64_t *rows, int64_t nr, const int64_t *cols, int64_t nc, double *out)
{
    tci_sub_ctx *x = (tci_sub_ctx *)c;
    const int32_t L = x->t->L, nl = x->I->len;
    int32_t *pt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(L + 1));
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t i = 0; i < nr; ++i) {
            memcpy(pt, ms_at(x->I, (int32_t)rows[i]), sizeof(int32_t) * (size_t)nl);
            memcpy(pt + nl, ms_at(x->J, (int32_t)cols[j]), sizeof(int32_t) * (size_t)(L - nl));
            out[i + nr * j] = feval(x->f, pt, L);
        }
    free(pt);
    updatemaxsample(x->t, out, nr * nc);
    x->t->ws[0] += (double)(nr * nc);
    x->t->ws[5] += 1.0;
}

static void tci_log(void *c, int64_t m, int64_t n, int64_t p) { log_rrlu((orc_tci *)c, m, n, p); }

static int updatepivots_x(orc_tci *t, int32_t b0, const orc_functor *f, int leftorth,
                          double reltol, double abstol, int64_t maxbonddim,
                          const mset *extraI, const mset *extraJ)
{
    invalidatesitetensors(t);
    mset Ic, Jc;
    ms_kron_right(&Ic, &t->I[b0], t->d[b0]);
    ms_union_inplace(&Ic, extraI);
    ms_kron_left(&Jc, t->d[b0 + 1], &t->J[b0 + 1]);
    ms_union_inplace(&Jc, extraJ);
    int64_t m = Ic.n, n = Jc.n, nc;
    if (t->pivotsearch == 1) {
        /* :rook (tensorci2.jl:552-595): I0/J0 = the positions of Iset[b+1] /
         * Jset[b] in Icombined / Jcombined (findfirst, misses dropped) */
        int64_t *I0 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
        int64_t *J0 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
        int64_t nI0 = 0, nJ0 = 0;
        for (int32_t e = 0; e < t->I[b0 + 1].n; ++e) {
            const int32_t i = ms_find(&Ic, ms_at(&t->I[b0 + 1], e));
            if (i >= 0) I0[nI0++] = i;
        }
        for (int32_t e = 0; e < t->J[b0].n; ++e) {
            const int32_t j = ms_find(&Jc, ms_at(&t->J[b0], e));
            if (j >= 0) J0[nJ0++] = j;
        }
        tci_sub_ctx ec = {t, f, &Ic, &Jc};
        rook_res rr;
        memset(&rr, 0, sizeof rr);
        int st = arrlu_core(tci_sub_eval, &ec, t->rng, m, n, I0, nI0, J0, nJ0, maxbonddim, reltol, abstol,
                            leftorth, 5, 0, &rr, tci_log, t);
        free(I0);
        free(J0);
        if (st != ORC_OK) { free_rook(&rr); ms_free(&Ic); ms_free(&Jc); return st; }
        if (rr.p > 0) {
            ms_select(&t->I[b0 + 1], &Ic, rr.rowperm, rr.p);
            ms_select(&t->J[b0], &Jc, rr.colperm, rr.p);
            double *e = (double *)malloc(sizeof(double) * (size_t)(rr.p + 1));
            for (int64_t k = 0; k < rr.p; ++k) e[k] = fabs(rr.pv[k]);
            e[rr.p] = rr.error;
            updateerrors(t, b0, e, (int32_t)rr.p + 1);
            free(e);
            free_rook(&rr);
            ms_free(&Ic);
            ms_free(&Jc);
            return ORC_OK;
        }
        free_rook(&rr);   /* no pivot: the full search (tensorci2.jl:584-595) */
    }
    double *Pi = (double *)malloc(sizeof(double) * (size_t)(m * n + 1));
    filltensor(t, f, &Ic, &Jc, Pi, &nc);
    updatemaxsample(t, Pi, m * n);
    luci_res r;
    int st = run_luci(t, Pi, m, n, reltol, abstol, maxbonddim, leftorth, &r);
    free(Pi);
    if (st != ORC_OK) { free_luci(&r); ms_free(&Ic); ms_free(&Jc); return st; }
    ms_select(&t->I[b0 + 1], &Ic, r.rowperm, r.p);
    ms_select(&t->J[b0], &Jc, r.colperm, r.p);
    double *e = luci_pivoterrors(&r, leftorth);
    updateerrors(t, b0, e, (int32_t)r.p + 1);
    free(e);
    free_luci(&r);
    ms_free(&Ic);
    ms_free(&Jc);
    return ORC_OK;
}

int orc_updatepivots(orc_tci *t, int32_t b0, const orc_functor *f, int32_t leftorth,
                     double reltol, double abstol, int64_t maxbonddim)
{
    return updatepivots_x(t, b0, f, leftorth, reltol, abstol, maxbonddim, NULL, NULL);
}

/* T = Pi1 · P⁻¹ via transpose(P) \ transpose(Pi1) (tensorci2.jl:391):
 * LU with partial pivoting of Pᵀ (first maximum |.| in the column), then
 * forward/back substitution for each row of Pi1.  r rows, c = order of P. */
static int solve_right(const double *Pi1, int64_t r, const double *P, int64_t c, double *X)
{
    double *A = (double *)malloc(sizeof(double) * (size_t)(c * c + 1));
    int64_t *piv = (int64_t *)malloc(sizeof(int64_t) * (size_t)(c + 1));
    double *y = (double *)malloc(sizeof(double) * (size_t)(c + 1));
    for (int64_t i = 0; i < c; ++i)
        for (int64_t j = 0; j < c; ++j) A[i + c * j] = P[j + c * i];   /* A = Pᵀ */
    int singular = 0;
    for (int64_t k = 0; k < c; ++k) {
        int64_t pr = k;
        double best = fabs(A[k + c * k]);
        for (int64_t i = k + 1; i < c; ++i) {
            double v = fabs(A[i + c * k]);
            if (v > best) { best = v; pr = i; }
        }
        piv[k] = pr;
        if (pr != k)
            for (int64_t j = 0; j < c; ++j) {
                double x = A[k + c * j]; A[k + c * j] = A[pr + c * j]; A[pr + c * j] = x;
            }
        double d = A[k + c * k];
        if (d == 0.0) singular = 1;
        for (int64_t i = k + 1; i < c; ++i) A[i + c * k] = A[i + c * k] / d;
        for (int64_t j = k + 1; j < c; ++j) {
            double u = A[k + c * j];
            for (int64_t i = k + 1; i < c; ++i) A[i + c * j] = A[i + c * j] - A[i + c * k] * u;
        }
    }
    if (!singular) {
        for (int64_t q = 0; q < r; ++q) {
            for (int64_t i = 0; i < c; ++i) y[i] = Pi1[q + r * i];
            for (int64_t k = 0; k < c; ++k)
                if (piv[k] != k) { double x = y[k]; y[k] = y[piv[k]]; y[piv[k]] = x; }
            for (int64_t i = 0; i < c; ++i) {
                double s = y[i];
                for (int64_t t = 0; t < i; ++t) s = s - A[i + c * t] * y[t];
                y[i] = s;
            }
            for (int64_t i = c - 1; i >= 0; --i) {
                double s = y[i];
                for (int64_t t = i + 1; t < c; ++t) s = s - A[i + c * t] * y[t];
                y[i] = s / A[i + c * i];
            }
            for (int64_t i = 0; i < c; ++i) X[q + r * i] = y[i];
        }
    }
    free(A); free(piv); free(y);
    return singular ? ORC_ERR_SINGULAR : ORC_OK;
}

static void set_T(orc_tci *t, int32_t b0, double *data, int64_t d0, int64_t d1, int64_t d2)
{
    free(t->T[b0]);
    t->T[b0] = data;
    t->Td[b0][0] = d0; t->Td[b0][1] = d1; t->Td[b0][2] = d2;
}

/* setsitetensor!(tci, f, b) (tensorci2.jl:367-394) with leftorthogonal=true */
static int setsitetensor_f(orc_tci *t, const orc_functor *f, int32_t b0)
{
    const mset *Ib = &t->I[b0], *Jb = &t->J[b0];
    int64_t nI = Ib->n, d = t->d[b0], nJ = Jb->n, nc;
    double *Pi1 = (double *)malloc(sizeof(double) * (size_t)(nI * d * nJ + 1));
    filltensor(t, f, Ib, Jb, Pi1, &nc);
    updatemaxsample(t, Pi1, nI * d * nJ);
    if (b0 == t->L - 1) { set_T(t, b0, Pi1, nI, d, nJ); return ORC_OK; }
    const mset *In = &t->I[b0 + 1];
    int64_t np = In->n;
    double *P = (double *)malloc(sizeof(double) * (size_t)(np * nJ + 1));
    filltensor(t, f, In, Jb, P, &nc);
    if (np != nJ) {
        free(P); free(Pi1);
        return fail(ORC_ERR_NONSQUARE, "Pivot matrix at bond %d is not square!", b0 + 1);
    }
    double *X = (double *)malloc(sizeof(double) * (size_t)(nI * d * np + 1));
    int st = solve_right(Pi1, nI * d, P, np, X);
    free(P); free(Pi1);
    if (st != ORC_OK) { free(X); return fail(st, "SingularException in P \\ Pi1 at bond %d", b0 + 1); }
    set_T(t, b0, X, nI, d, np);
    return ORC_OK;
}

/* fillsitetensors! (globalsearch.jl:97-104) */
int orc_fillsitetensors(orc_tci *t, const orc_functor *f)
{
    for (int32_t b0 = 0; b0 < t->L; ++b0) {
        int st = setsitetensor_f(t, f, b0);
        if (st != ORC_OK) return st;
    }
    return ORC_OK;
}

/* sweep1site! (tensorci2.jl:402-461) */
int orc_sweep1site(orc_tci *t, const orc_functor *f, int32_t forward, double reltol,
                   double abstol, int64_t maxbonddim, int32_t updatetensors)
{
    t->npe = 0;                       /* flushpivoterror! */
    invalidatesitetensors(t);
    const int32_t L = t->L;
    for (int32_t s = 0; s < L - 1; ++s) {
        int32_t b0 = forward ? s : (L - 1 - s);
        mset Is, Js;
        if (forward) { ms_kron_right(&Is, &t->I[b0], t->d[b0]); ms_init(&Js, 0); ms_copy(&Js, &t->J[b0]); }
        else { ms_init(&Is, 0); ms_copy(&Is, &t->I[b0]); ms_kron_left(&Js, t->d[b0], &t->J[b0]); }
        int64_t m = Is.n, n = Js.n, nc;
        double *Pi = (double *)malloc(sizeof(double) * (size_t)(m * n + 1));
        filltensor(t, f, &t->I[b0], &t->J[b0], Pi, &nc);
        updatemaxsample(t, Pi, m * n);
        luci_res r;
        int st = run_luci(t, Pi, m, n, reltol, abstol, maxbonddim, forward, &r);
        free(Pi);
        if (st != ORC_OK) { free_luci(&r); ms_free(&Is); ms_free(&Js); return st; }
        if (forward) {
            ms_select(&t->I[b0 + 1], &Is, r.rowperm, r.p);
            ms_select(&t->J[b0], &Js, r.colperm, r.p);
        } else {
            ms_select(&t->I[b0], &Is, r.rowperm, r.p);
            ms_select(&t->J[b0 - 1], &Js, r.colperm, r.p);
        }
        if (updatetensors) {
            if (forward) {
                double *Lm = (double *)malloc(sizeof(double) * (size_t)(m * r.p + 1));
                orc_luci_left(m, n, r.p, 1, r.rowperm, r.colperm, r.L, r.U, r.ldu, Lm);
                set_T(t, b0, Lm, t->I[b0].n, t->d[b0], t->J[b0].n);
            } else {
                double *Rm = (double *)malloc(sizeof(double) * (size_t)(r.p * n + 1));
                orc_luci_right(m, n, r.p, 0, r.rowperm, r.colperm, r.L, r.U, r.ldu, Rm);
                set_T(t, b0, Rm, t->I[b0].n, t->d[b0], t->J[b0].n);
            }
            int64_t cnt = t->Td[b0][0] * t->Td[b0][1] * t->Td[b0][2];
            for (int64_t q = 0; q < cnt; ++q)
                if (t->T[b0][q] != t->T[b0][q]) {
                    free_luci(&r); ms_free(&Is); ms_free(&Js);
                    return fail(ORC_ERR_NAN_TENSOR, "Error: NaN in tensor T[%d]", b0 + 1);
                }
        }
        double *e = luci_pivoterrors(&r, forward);
        updateerrors(t, forward ? b0 : b0 - 1, e, (int32_t)r.p + 1);
        free(e);
        free_luci(&r);
        ms_free(&Is);
        ms_free(&Js);
    }
    if (updatetensors) {
        int32_t b0 = forward ? L - 1 : 0;
        int64_t nI = t->I[b0].n, nJ = t->J[b0].n, d = t->d[b0], nc;
        double *X = (double *)malloc(sizeof(double) * (size_t)(nI * d * nJ + 1));
        filltensor(t, f, &t->I[b0], &t->J[b0], X, &nc);
        set_T(t, b0, X, nI, d, nJ);
    }
    return ORC_OK;
}

/* forwardsweep (sweepstrategies.jl:1-6) */
static int forwardsweep(int32_t strategy, int32_t iter)
{
    return strategy == 1 || (strategy == 0 && (iter & 1));
}

/* sweep2site! (tensorci2.jl:855-916) */
int orc_sweep2site(orc_tci *t, const orc_functor *f, int32_t niter, double abstol,
                   int64_t maxbonddim, int32_t sweepstrategy, int32_t strictlynested,
                   int32_t fillsitetensors)
{
    invalidatesitetensors(t);
    const int32_t L = t->L;
    mset *eI = (mset *)calloc((size_t)L, sizeof(mset));
    mset *eJ = (mset *)calloc((size_t)L, sizeof(mset));
    for (int32_t iter = 1; iter <= niter; ++iter) {
        int use_extra = !strictlynested && t->has_hist;
        for (int b = 0; b < L; ++b) {
            ms_free(&eI[b]); ms_free(&eJ[b]);
            ms_init(&eI[b], b); ms_init(&eJ[b], L - 1 - b);
            if (use_extra) { ms_copy(&eI[b], &t->hI[b]); ms_copy(&eJ[b], &t->hJ[b]); }
        }
        for (int b = 0; b < L; ++b) { ms_copy(&t->hI[b], &t->I[b]); ms_copy(&t->hJ[b], &t->J[b]); }
        t->has_hist = 1;
        t->npe = 0;                   /* flushpivoterror! */
        int fwd = forwardsweep(sweepstrategy, iter);
        for (int32_t s = 0; s < L - 1; ++s) {
            int32_t b0 = fwd ? s : (L - 2 - s);
            int st = updatepivots_x(t, b0, f, fwd, 1e-14, abstol, maxbonddim, &eI[b0 + 1], &eJ[b0]);
            if (st != ORC_OK) {
                for (int b = 0; b < L; ++b) { ms_free(&eI[b]); ms_free(&eJ[b]); }
                free(eI); free(eJ);
                return st;
            }
        }
    }
    for (int b = 0; b < L; ++b) { ms_free(&eI[b]); ms_free(&eJ[b]); }
    free(eI); free(eJ);
    if (fillsitetensors) return orc_fillsitetensors(t, f);
    return ORC_OK;
}

/* TT evaluation: prod(T[:, x_ℓ, :]) left to right (abstracttensortrain.jl:131);
 * each vector-matrix product summed in ascending bond order. */
double orc_tci_evaluate(const orc_tci *t, const int32_t *x)
{
    int64_t chl = t->Td[0][0];
    double *v = (double *)malloc(sizeof(double) * (size_t)(chl > 0 ? chl : 1));
    for (int64_t i = 0; i < chl; ++i) v[i] = 1.0;
    for (int32_t b = 0; b < t->L; ++b) {
        int64_t d0 = t->Td[b][0], d = t->Td[b][1], d2 = t->Td[b][2];
        int64_t s = x[b] - 1;
        double *w = (double *)malloc(sizeof(double) * (size_t)(d2 > 0 ? d2 : 1));
        for (int64_t j = 0; j < d2; ++j) {
            double acc = 0.0;
            for (int64_t i = 0; i < d0; ++i) acc = acc + v[i] * t->T[b][i + d0 * (s + d * j)];
            w[j] = acc;
        }
        free(v);
        v = w;
        chl = d2;
    }
    double r = chl > 0 ? v[0] : 0.0;
    free(v);
    return r;
}

void orc_eval_points(const orc_functor *f, int32_t L, int64_t npts, const int32_t *pts, double *out)
{
    for (int64_t q = 0; q < npts; ++q) out[q] = feval(f, pts + q * L, L);
}

void orc_tt_evaluate(int32_t L, const int64_t *dims3, const double *cores, int64_t npts,
                     const int32_t *pts, double *out)
{
    int64_t chimax = 1;
    for (int32_t b = 0; b < L; ++b) {
        if (dims3[3 * b] > chimax) chimax = dims3[3 * b];
        if (dims3[3 * b + 2] > chimax) chimax = dims3[3 * b + 2];
    }
    double *v = (double *)malloc(sizeof(double) * (size_t)chimax);
    double *w = (double *)malloc(sizeof(double) * (size_t)chimax);
    for (int64_t q = 0; q < npts; ++q) {
        const int32_t *x = pts + q * L;
        int64_t chl = dims3[0];
        for (int64_t i = 0; i < chl; ++i) v[i] = 1.0;
        const double *T = cores;
        for (int32_t b = 0; b < L; ++b) {
            const int64_t d0 = dims3[3 * b], d = dims3[3 * b + 1], d2 = dims3[3 * b + 2];
            const int64_t s = x[b] - 1;
            for (int64_t j = 0; j < d2; ++j) {
                double acc = 0.0;
                for (int64_t i = 0; i < d0; ++i) acc = acc + v[i] * T[i + d0 * (s + d * j)];
                w[j] = acc;
            }
            double *tmp = v; v = w; w = tmp;
            chl = d2;
            T += d0 * d * d2;
        }
        out[q] = chl > 0 ? v[0] : 0.0;
    }
    free(v);
    free(w);
}

/* Julia's isless on Float64 for the argmax of _floatingzone: NaN sorts
 * above every number (and equal to itself) */
static int jl_isless(double a, double b)
{
    if (a != a) return 0;
    if (b != b) return 1;
    return a < b;
}

/* _floatingzone (globalsearch.jl:43-95) from each start point; TT values
 * through the left row vector l and the right environment R of the current
 * point: TT_σ = Σ_k (Σ_j l[j] T_p[j,σ,k]) R_{p+1}[k], sums ascending.  A NaN
 * error ends the search (the reference would not stop). */
void orc_floatingzone(const orc_functor *f, int32_t L, const int64_t *dims3, const double *cores, int64_t npts,
                      int32_t *pts, double *errs, int64_t nsweeps, double earlystoptol)
{
    int64_t *coff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(L + 1));
    int64_t chimax = 1, dmax = 1;
    coff[0] = 0;
    for (int32_t b = 0; b < L; ++b) {
        coff[b + 1] = coff[b] + dims3[3 * b] * dims3[3 * b + 1] * dims3[3 * b + 2];
        if (dims3[3 * b + 2] > chimax) chimax = dims3[3 * b + 2];
        if (dims3[3 * b] > chimax) chimax = dims3[3 * b];
        if (dims3[3 * b + 1] > dmax) dmax = dims3[3 * b + 1];
    }
    /* E[b] (b = 1..L): the product of sites b..L-1 at the point, χ_b long */
    double **E = (double **)malloc(sizeof(double *) * (size_t)(L + 1));
    for (int32_t b = 0; b <= L; ++b) E[b] = (double *)malloc(sizeof(double) * (size_t)chimax);
    double *l = (double *)malloc(sizeof(double) * (size_t)chimax);
    double *w = (double *)malloc(sizeof(double) * (size_t)(dmax * chimax));
    double *err = (double *)malloc(sizeof(double) * (size_t)dmax);
    int32_t *x = (int32_t *)malloc(sizeof(int32_t) * (size_t)L);
    for (int64_t q = 0; q < npts; ++q) {
        int32_t *piv = pts + q * L;
        double maxerror = 0.0;
        for (int64_t sweep = 0;; ++sweep) {
            E[L][0] = 1.0;
            for (int32_t b = L - 1; b >= 1; --b) {
                const int64_t d0 = dims3[3 * b], d = dims3[3 * b + 1], d2 = dims3[3 * b + 2];
                const double *T = cores + coff[b];
                for (int64_t j = 0; j < d0; ++j) {
                    double acc = 0.0;
                    for (int64_t k = 0; k < d2; ++k) acc = acc + T[j + d0 * ((piv[b] - 1) + d * k)] * E[b + 1][k];
                    E[b][j] = acc;
                }
            }
            if (sweep == 0) {
                /* maxerror = abs(f(pivot) - ttcache(pivot)) (globalsearch.jl:62) */
                const int64_t d = dims3[1], d2 = dims3[2];
                double acc = 0.0;
                for (int64_t k = 0; k < d2; ++k) acc = acc + (0.0 + 1.0 * cores[(piv[0] - 1) + d * k]) * E[1][k];
                maxerror = fabs(feval(f, piv, L) - acc);
            }
            const double prev = maxerror;
            l[0] = 1.0;
            for (int32_t p = 0; p < L; ++p) {
                const int64_t d0 = dims3[3 * p], d = dims3[3 * p + 1], d2 = dims3[3 * p + 2];
                const double *T = cores + coff[p];
                for (int64_t s = 0; s < d; ++s)
                    for (int64_t k = 0; k < d2; ++k) {
                        double acc = 0.0;
                        for (int64_t j = 0; j < d0; ++j) acc = acc + l[j] * T[j + d0 * (s + d * k)];
                        w[s * chimax + k] = acc;
                    }
                for (int64_t s = 0; s < d; ++s) {
                    double acc = 0.0;
                    for (int64_t k = 0; k < d2; ++k) acc = acc + w[s * chimax + k] * E[p + 1][k];
                    memcpy(x, piv, sizeof(int32_t) * (size_t)L);
                    x[p] = (int32_t)(s + 1);
                    err[s] = fabs(feval(f, x, L) - acc);
                }
                /* pivot[ipos] = argmax(err); maxerror = max(maximum(err), maxerror) */
                int64_t best = 0;
                double mx = err[0];
                for (int64_t s = 1; s < d; ++s) {
                    if (jl_isless(err[best], err[s])) best = s;
                    mx = jl_max(mx, err[s]);
                }
                piv[p] = (int32_t)(best + 1);
                maxerror = jl_max(mx, maxerror);
                for (int64_t k = 0; k < d2; ++k) l[k] = w[best * chimax + k];
            }
            if (maxerror == prev || maxerror > earlystoptol || maxerror != maxerror || sweep + 1 >= nsweeps) break;
        }
        errs[q] = maxerror;
    }
    for (int32_t b = 0; b <= L; ++b) free(E[b]);
    free(E); free(l); free(w); free(err); free(x); free(coff);
}

/* ===================================================================== */
/* Tensor-train sweeps — compress! (tensortrain.jl:95-183) and            */
/* sweep1sitegetindices! (conversion.jl:73-156)                           */
/* ===================================================================== */

/* C = A·B (M×K · K×N, column-major, lda = M, ldb = K), each sum in
 * ascending k from 0.0; the reference calls BLAS for these products */
static void gemm_fixed(int64_t M, int64_t N, int64_t K, const double *A, const double *B, double *Cm)
{
    for (int64_t j = 0; j < N; ++j)
        for (int64_t i = 0; i < M; ++i) {
            double acc = 0.0;
            for (int64_t k = 0; k < K; ++k) acc = acc + A[i + M * k] * B[k + K * j];
            Cm[i + M * j] = acc;
        }
}

/* The factors of one grouped site matrix (m×n): for method 0 left(lu),
 * right(lu) with the permutations applied (matrixlu.jl:374-392), for
 * method 1 left(luci), right(luci) (matrixluci.jl:70-84).  The diagonal of
 * the factored matrix goes to diagp (|U_kk| or |L_kk|, pivoterrors at
 * matrixlu.jl:394-416). */
static int tt_factor(const double *A, int64_t m, int64_t n, int32_t method, int32_t leftorth,
                     int64_t maxrank, double reltol, double abstol, int64_t **rp, int64_t **cp,
                     double **left, double **right, int64_t *p_out, double **pe)
{
    int64_t cap = m < n ? m : n;
    if (maxrank < cap) cap = maxrank;
    int64_t *rowperm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    int64_t *colperm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    double *Lf = (double *)calloc((size_t)(m * cap + 1), sizeof(double));
    double *Uf = (double *)calloc((size_t)(cap * n + 1), sizeof(double));
    int64_t p = 0;
    double err = 0.0;
    int st = orc_rrlu(A, m, n, maxrank, reltol, abstol, leftorth, rowperm, colperm, Lf, Uf,
                      cap > 0 ? cap : 1, &p, &err);
    if (st != ORC_OK) {
        free(rowperm); free(colperm); free(Lf); free(Uf);
        return st;
    }
    double *l = (double *)calloc((size_t)(m * p + 1), sizeof(double));
    double *r = (double *)calloc((size_t)(p * n + 1), sizeof(double));
    const int64_t ldu = cap > 0 ? cap : 1;
    if (method == 0) {
        for (int64_t j = 0; j < p; ++j)
            for (int64_t i = 0; i < m; ++i) l[(rowperm[i] - 1) + m * j] = Lf[i + m * j];
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < p; ++i) r[i + p * (colperm[j] - 1)] = Uf[i + ldu * j];
    } else {
        orc_luci_left(m, n, p, leftorth, rowperm, colperm, Lf, Uf, ldu, l);
        orc_luci_right(m, n, p, leftorth, rowperm, colperm, Lf, Uf, ldu, r);
    }
    /* pivoterrors(lu) = [|diag(lu)|..., lu.error]: diag of U with
     * leftorthogonal, of L without (matrixlu.jl:394-400) */
    double *e = (double *)malloc(sizeof(double) * (size_t)(p + 1));
    for (int64_t k = 0; k < p; ++k) e[k] = fabs(leftorth ? Uf[k + ldu * k] : Lf[k + m * k]);
    e[p] = err;
    free(Lf); free(Uf);
    *rp = rowperm; *cp = colperm; *left = l; *right = r; *p_out = p; *pe = e;
    return ORC_OK;
}

int orc_tt_sweep(int32_t L, int64_t *linkdims, const int32_t *localdims, double *cores, int32_t method,
                 int32_t forward, int64_t maxrank, double reltol, double abstol, int32_t *pivots,
                 double *pivoterrors)
{
    if (L < 1) return fail(ORC_ERR_ARG, "a tensor train needs at least one site");
    if (method != 0 && method != 1) return fail(ORC_ERR_ARG, "method must be :LU or :CI");
    int64_t R = 0;
    for (int32_t b = 1; b < L; ++b) R = linkdims[b] > R ? linkdims[b] : R;
    if (pivoterrors)
        for (int64_t k = 0; k <= R; ++k) pivoterrors[k] = 0.0;   /* zeros(rank(tt) + 1) */
    double **T = (double **)malloc(sizeof(double *) * (size_t)L);
    int64_t off = 0;
    for (int32_t b = 0; b < L; ++b) {
        const int64_t sz = linkdims[b] * localdims[b] * linkdims[b + 1];
        T[b] = (double *)malloc(sizeof(double) * (size_t)(sz + 1));
        memcpy(T[b], cores + off, sizeof(double) * (size_t)sz);
        off += sz;
    }
    int64_t *perms = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * R + 2) * (size_t)(L > 1 ? L - 1 : 1));
    int64_t *npiv = (int64_t *)calloc((size_t)(L > 1 ? L - 1 : 1), sizeof(int64_t));
    int st = ORC_OK;
    for (int32_t i = 0; i < L - 1 && st == ORC_OK; ++i) {
        /* the bond b sits between sites b and b+1; forward factors site b
         * as (χ_b d_b)×χ_{b+1}, backward site b+1 as χ_{b+1}×(d_{b+1} χ_{b+2}) */
        const int32_t b = forward ? i : L - 2 - i;
        const int32_t s = forward ? b : b + 1;
        const int64_t m = forward ? linkdims[b] * localdims[b] : linkdims[b + 1];
        const int64_t n = forward ? linkdims[b + 1] : localdims[b + 1] * linkdims[b + 2];
        int64_t *rp, *cp, p;
        double *l, *r, *e;
        st = tt_factor(T[s], m, n, method, forward, maxrank, reltol, abstol, &rp, &cp, &l, &r, &p, &e);
        if (st != ORC_OK) break;
        if (forward) {
            /* site b ← left; site b+1 ← right · (χ_{b+1} × d_{b+1} χ_{b+2}) */
            const int64_t nn = localdims[b + 1] * linkdims[b + 2];
            double *nx = (double *)malloc(sizeof(double) * (size_t)(p * nn + 1));
            gemm_fixed(p, nn, n, r, T[b + 1], nx);
            free(T[b]); free(T[b + 1]); free(r);
            T[b] = l;
            T[b + 1] = nx;
        } else {
            /* site b+1 ← right; site b ← (χ_b d_b × χ_{b+1}) · left */
            const int64_t mm = linkdims[b] * localdims[b];
            double *nx = (double *)malloc(sizeof(double) * (size_t)(mm * p + 1));
            gemm_fixed(mm, p, m, T[b], l, nx);
            free(T[b]); free(T[b + 1]); free(l);
            T[b + 1] = r;
            T[b] = nx;
        }
        for (int64_t k = 0; k < p; ++k) {
            perms[(size_t)b * (size_t)(2 * R + 2) + (size_t)k] = rp[k];
            perms[(size_t)b * (size_t)(2 * R + 2) + (size_t)(p + k)] = cp[k];
        }
        linkdims[b + 1] = p;
        npiv[b] = p;
        if (pivoterrors)
            for (int64_t k = 0; k <= p; ++k) pivoterrors[k] = jl_max(pivoterrors[k], e[k]);
        free(rp); free(cp); free(e);
    }
    if (st == ORC_OK) {
        off = 0;
        for (int32_t b = 0; b < L; ++b) {
            const int64_t sz = linkdims[b] * localdims[b] * linkdims[b + 1];
            memcpy(cores + off, T[b], sizeof(double) * (size_t)sz);
            off += sz;
        }
        if (pivots) {
            int64_t q = 0;
            for (int32_t b = 0; b < L - 1; ++b)
                for (int64_t k = 0; k < 2 * npiv[b]; ++k)
                    pivots[q++] = (int32_t)perms[(size_t)b * (size_t)(2 * R + 2) + (size_t)k];
        }
    }
    for (int32_t b = 0; b < L; ++b) free(T[b]);
    free(T);
    free(perms);
    free(npiv);
    return st;
}

int orc_tt_compress(int32_t L, int64_t *linkdims, const int32_t *localdims, double *cores, int32_t method,
                    double tolerance, int64_t maxbonddim, int32_t normalizeerror)
{
    /* left to right without truncation: _factorize with tolerance 0.0 and
     * the default normalizeerror, so reltol = abstol = 0 */
    int st = orc_tt_sweep(L, linkdims, localdims, cores, method, 1, INT64_MAX, 0.0, 0.0, NULL, NULL);
    if (st != ORC_OK) return st;
    /* right to left: reltol = tolerance (normalizeerror) or reltol = 1e-14
     * and abstol = tolerance (tensortrain.jl:98-104) */
    const double reltol = normalizeerror ? tolerance : 1e-14;
    const double abstol = normalizeerror ? 0.0 : tolerance;
    return orc_tt_sweep(L, linkdims, localdims, cores, method, 0, maxbonddim, reltol, abstol, NULL, NULL);
}

/* the functor math routines, exported for their accuracy tests */
double orc_exp(double x) { return tci_exp(x); }
double orc_cos(double x) { return tci_cos(x); }

/* ---- project PRNG ---- */
static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
void orc_rng_seed(uint64_t s[4], uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(&x);
}
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
uint64_t orc_rng_next(uint64_t s[4])
{
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}
/* rand(1:d) */
static int32_t rng_randint(uint64_t s[4], int32_t d)
{
    return 1 + (int32_t)((orc_rng_next(s) >> 32) % (uint64_t)d);
}

/* DefaultGlobalPivotFinder (globalpivotfinder.jl:143-195). Returns count;
 * found holds up to nsearch pivots (L each). */
static int32_t globalpivotfinder(orc_tci *t, const orc_functor *f, double abstol,
                                 int32_t nsearch, int32_t maxng, double margin, int32_t *found)
{
    const int32_t L = t->L;
    int32_t *pts = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nsearch * L + 1));
    for (int32_t q = 0; q < nsearch; ++q)
        for (int32_t p = 0; p < L; ++p) pts[q * L + p] = rng_randint(t->rng, t->d[p]);
    int32_t nf = 0;
    int32_t cur[256], best[256];
    for (int32_t q = 0; q < nsearch; ++q) {
        const int32_t *pt = pts + q * L;
        memcpy(cur, pt, sizeof(int32_t) * (size_t)L);
        memcpy(best, pt, sizeof(int32_t) * (size_t)L);
        double best_err = 0.0;
        for (int32_t p = 0; p < L; ++p) {
            for (int32_t v = 1; v <= t->d[p]; ++v) {
                cur[p] = v;
                double err = fabs(feval(f, cur, L) - orc_tci_evaluate(t, cur));
                if (err > best_err) { best_err = err; memcpy(best, cur, sizeof(int32_t) * (size_t)L); }
            }
            cur[p] = pt[p];
        }
        if (best_err > abstol * margin) { memcpy(found + nf * L, best, sizeof(int32_t) * (size_t)L); nf++; }
    }
    free(pts);
    if (nf > maxng) nf = maxng;
    return nf;
}

/* the global pivot finder on the current state with its PRNG seeded by
 * `seed` (for the brute-force checks of the selection rule) */
int32_t orc_globalpivotfinder(orc_tci *t, const orc_functor *f, double abstol, int32_t nsearch,
                              int32_t maxng, double margin, uint64_t seed, int32_t *found)
{
    orc_rng_seed(t->rng, seed);
    return globalpivotfinder(t, f, abstol, nsearch, maxng, margin, found);
}

/* convergencecriterion (tensorci2.jl:609-628) */
int orc_convergencecriterion(int32_t n, const int64_t *ranks, const double *errors,
                             const int64_t *nglobal, double tol, int64_t maxbonddim,
                             int32_t nch, int32_t checkconvglobalpivot)
{
    if (n < nch) return 0;
    int all_err = 1, all_ng = 1, all_max = 1;
    int64_t minrank = ranks[n - nch];
    for (int32_t i = n - nch; i < n; ++i) {
        if (!(errors[i] < tol)) all_err = 0;
        if (nglobal[i] != 0) all_ng = 0;
        if (ranks[i] < minrank) minrank = ranks[i];
        if (!(ranks[i] >= maxbonddim)) all_max = 0;
    }
    if (!checkconvglobalpivot) all_ng = 1;
    return (all_err && all_ng && minrank == ranks[n - 1]) || all_max;
}

static int64_t tci_rank(const orc_tci *t)   /* linkdims / rank (tensorci2.jl:83-85) */
{
    int64_t r = 0;
    for (int b = 1; b < t->L; ++b) if (t->I[b].n > r) r = t->I[b].n;
    return r;
}

static void push_iter(orc_tci *t, int64_t rank, double err, int64_t ng)
{
    if (t->niter + 1 > t->capiter) {
        t->capiter = t->capiter ? t->capiter * 2 : 16;
        t->ranks = (int64_t *)realloc(t->ranks, sizeof(int64_t) * (size_t)t->capiter);
        t->nglob = (int64_t *)realloc(t->nglob, sizeof(int64_t) * (size_t)t->capiter);
        t->errors = (double *)realloc(t->errors, sizeof(double) * (size_t)t->capiter);
    }
    t->ranks[t->niter] = rank; t->errors[t->niter] = err; t->nglob[t->niter] = ng;
    t->niter++;
}

/* optimize! (tensorci2.jl:700-850) */
int orc_optimize(orc_tci *t, const orc_functor *f, const orc_opts *o)
{
    t->niter = 0;
    if (o->nsearchglobalpivot > 0 && o->nsearchglobalpivot < o->maxnglobalpivot)
        return fail(ORC_ERR_GENERIC, "nsearchglobalpivot < maxnglobalpivot!");
    double tol = o->tolerance;
    if (o->maxbonddim >= INT64_MAX && tol <= 0)
        return fail(ORC_ERR_ARG, "Specify either tolerance > 0 or some maxbonddim; otherwise, "
                                 "the convergence criterion is not reachable!");
    orc_rng_seed(t->rng, o->seed);
    t->pivotsearch = o->pivotsearch;
    const int32_t L = t->L;
    int32_t *found = (int32_t *)malloc(sizeof(int32_t) * (size_t)((o->nsearchglobalpivot + 1) * L));
    int st = ORC_OK;
    for (int32_t iter = 1; iter <= o->maxiter; ++iter) {
        double norm = o->normalizeerror ? t->maxsample : 1.0;
        double abstol = tol * norm;
        st = orc_sweep2site(t, f, 2, abstol, o->maxbonddim, o->sweepstrategy, o->strictlynested, 1);
        if (st != ORC_OK) break;
        double perr = t->bonderrors[0];
        for (int b = 1; b < L - 1; ++b) perr = jl_max(perr, t->bonderrors[b]);
        int32_t ng = 0;
        if (o->nsearchglobalpivot > 0)
            ng = globalpivotfinder(t, f, abstol, o->nsearchglobalpivot, o->maxnglobalpivot,
                                   o->tolmarginglobalsearch, found);
        orc_addglobalpivots(t, ng, found);
        push_iter(t, tci_rank(t), perr, ng);
        if (orc_convergencecriterion(t->niter, t->ranks, t->errors, t->nglob, abstol, o->maxbonddim,
                                     o->ncheckhistory, o->checkconvglobalpivot))
            break;
    }
    free(found);
    if (st != ORC_OK) return st;
    double norm = o->normalizeerror ? t->maxsample : 1.0;
    double abstol = tol * norm;
    st = orc_sweep1site(t, f, 1, 1e-14, abstol, o->maxbonddim, 1);
    if (st != ORC_OK) return st;
    /* _sanitycheck (globalsearch.jl:106-111) */
    for (int b = 0; b < L - 1; ++b)
        if (t->I[b + 1].n != t->J[b].n)
            return fail(ORC_ERR_NONSQUARE, "Pivot matrix at bond %d is not square!", b + 1);
    for (int32_t i = 0; i < t->niter; ++i) t->errors[i] = t->errors[i] / norm;
    return ORC_OK;
}

int orc_crossinterpolate2(orc_tci *t, const orc_functor *f, int32_t npiv,
                          const int32_t *pivots, const orc_opts *o)
{
    int st = orc_tci_init(t, f, npiv, pivots);
    if (st != ORC_OK) return st;
    return orc_optimize(t, f, o);
}

/* ---- accessors ---- */
int32_t orc_tci_len(const orc_tci *t) { return t->L; }
int32_t orc_tci_setsize(const orc_tci *t, int32_t which, int32_t b0)
{
    return which ? t->J[b0].n : t->I[b0].n;
}
void orc_tci_getset(const orc_tci *t, int32_t which, int32_t b0, int32_t *out)
{
    const mset *s = which ? &t->J[b0] : &t->I[b0];
    if (s->n && s->len) memcpy(out, s->v, sizeof(int32_t) * (size_t)(s->n * s->len));
}
int orc_tci_setset(orc_tci *t, int32_t which, int32_t b0, int32_t n, const int32_t *in)
{
    mset *s = which ? &t->J[b0] : &t->I[b0];
    s->n = 0;
    for (int32_t i = 0; i < n; ++i) ms_push(s, in + (size_t)i * s->len);
    return ORC_OK;
}
void orc_tci_sitetensor_dims(const orc_tci *t, int32_t b0, int64_t *dims3)
{
    dims3[0] = t->Td[b0][0]; dims3[1] = t->Td[b0][1]; dims3[2] = t->Td[b0][2];
}
void orc_tci_sitetensor(const orc_tci *t, int32_t b0, double *out)
{
    int64_t n = t->Td[b0][0] * t->Td[b0][1] * t->Td[b0][2];
    if (n > 0) memcpy(out, t->T[b0], sizeof(double) * (size_t)n);
}
double orc_tci_maxsample(const orc_tci *t) { return t->maxsample; }
void orc_tci_set_maxsample(orc_tci *t, double v) { t->maxsample = v; }
int32_t orc_tci_npivoterrors(const orc_tci *t) { return t->npe; }
void orc_tci_pivoterrors(const orc_tci *t, double *out)
{
    if (t->npe) memcpy(out, t->pivoterrors, sizeof(double) * (size_t)t->npe);
}
void orc_tci_bonderrors(const orc_tci *t, double *out)
{
    memcpy(out, t->bonderrors, sizeof(double) * (size_t)(t->L - 1));
}
int32_t orc_tci_niter(const orc_tci *t) { return t->niter; }
void orc_tci_history(const orc_tci *t, int64_t *ranks, double *errors, int64_t *nglobal)
{
    for (int32_t i = 0; i < t->niter; ++i) {
        if (ranks) ranks[i] = t->ranks[i];
        if (errors) errors[i] = t->errors[i];
        if (nglobal) nglobal[i] = t->nglob[i];
    }
}
void orc_tci_workstats(const orc_tci *t, double *out6)
{
    for (int i = 0; i < 6; ++i) out6[i] = t->ws[i];
}
