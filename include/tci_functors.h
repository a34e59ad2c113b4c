/*
 * tci_functors.h — the functions f(σ₁,…,σ_L) that a TCI solve interpolates.
 *
 * The reference passes an arbitrary Julia callable `f(::MultiIndex)::T` to
 * `crossinterpolate2` (tensorci2.jl:943-953) and evaluates it element by
 * element in `_batchevaluate_dispatch` (batcheval.jl:32-61).  A GPU cannot
 * run a Julia closure, so here f is a *device functor*: a kind tag plus a
 * parameter block, evaluated by the same inline code on the host and on the
 * device.  This header is the problem definition: the CPU oracle and the HIP
 * engine both include it, exactly as the reference's tests hand the same
 * closure to every code path.
 *
 * Arithmetic rules (needed for bit-identical host/device values):
 *   - every file that includes this header is compiled with -ffp-contract=off,
 *     so a*b+c is two roundings everywhere;
 *   - exp and cos are the project's own routines below (Cody–Waite argument
 *     reduction with split constants derived from ln 2 and π, then Horner
 *     evaluation of the Taylor series with coefficients 1/n!); no vendor
 *     math library is called, so host and device produce the same bits.
 *
 * Multi-indices are 1-based, as in the reference (`MultiIndex = Vector{Int}`,
 * abstracttensortrain.jl:6-7).  A point is passed as two byte strings, the
 * row part (sites 1..na) and the column part (sites na+1..L), which is how a
 * Pi-matrix element (row multi-index ⊕ column multi-index) is formed
 * (batcheval.jl:53-56).
 */
#ifndef TCI_FUNCTORS_H
#define TCI_FUNCTORS_H

#include <stdint.h>

#if defined(__HIPCC__)
#define TCI_HD __host__ __device__
#else
#define TCI_HD
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Functor kinds.  Parameter layouts are given per kind. */
enum tci_functor_kind {
    /* f = a / (1 + Σ σ_i²), integer sum then one FP64 division.
     * params[0] = a.  (README.md:24-28; test_tensorci2.jl:249 with coeff a.) */
    TCI_F_LORENTZ_INT = 1,
    /* f = 1 / (1 + Σ_i c_i (σ_i/D)²), summed i = 1..L, unfused.
     * params[0] = D, params[1..L] = c_i.  (high-dimensional workload) */
    TCI_F_LORENTZ_WEIGHTED = 2,
    /* quantics: x = Σ_i (σ_i−1) 2^{−i} (site 1 most significant);
     * f = Σ_k a_k exp(−(x−c_k)²/(2 w_k²)) + A cos(F x).
     * params[0] = K, then K triples (a_k, c_k, w_k), then A, F. */
    TCI_F_QUANTICS_GAUSS_COS = 3,
    /* x_i = (σ_i−1)/(d−1); f = Σ_k a_k exp(−Σ_i (x_i−c_{k,i})² / (2 w_k²)).
     * params[0] = K, params[1] = d, then a[K], w[K], then c[K][L] (k-major). */
    TCI_F_GAUSS_SUM = 4,
    /* f = s · Σ σ_i  (batcheval closed form, test_batcheval.jl:13-35).
     * params[0] = s. */
    TCI_F_SUM = 5,
    /* dense table over a small domain: f = params[Σ_i (σ_i−1)·Π_{j<i} d_j]
     * (column-major, first index fastest).  iparams[0..L-1] = d_i. */
    TCI_F_TABLE = 6,
    /* tensor train: f = T_1[1,σ₁,:]·T_2[:,σ₂,:]···T_L[:,σ_L,1], left to right,
     * each product summed in ascending bond order.  iparams[0..L] = bond dims
     * χ_0..χ_L (χ_0 = χ_L = 1), iparams[L+1..2L] = local dims; params = cores
     * concatenated, each χ_{i-1}×d_i×χ_i column-major. */
    TCI_F_TT = 7,
    /* quantics exponentials: x as in kind 3;
     * f = α₁ exp(β₁ x) + α₂ exp(β₂ x).  params = α₁, β₁, α₂, β₂. */
    TCI_F_QUANTICS_EXP = 8
};

/* ---------------------------------------------------------------------- */
/* Project-owned exp / cos.                                                  */
/* ---------------------------------------------------------------------- */

typedef union { double d; uint64_t u; } tci_f64bits;

static inline TCI_HD double tci_pow2i(int k)
{
    /* exact 2^k for -1022 <= k <= 1023 */
    tci_f64bits b;
    b.u = (uint64_t)(k + 1023) << 52;
    return b.d;
}

static inline TCI_HD double tci_round_half_away(double x)
{
    /* nearest integer, ties away from zero; |x| < 2^52 */
    return x >= 0.0 ? (double)(int64_t)(x + 0.5) : -(double)(int64_t)(0.5 - x);
}

/* exp(x): x = k·ln2 + r with ln2 split as hi (42 significant bits, so k·hi
 * is exact for |k| < 2^11) + lo; exp(r) for |r| <= ln2/2 by the Taylor series
 * to r^15 (truncation < 2^-60); result scaled by 2^k. */
static inline TCI_HD double tci_exp(double x)
{
    const double LN2_HI = 0.6931471805598903;        /* bits 0x3fe62e42fefa3800 */
    const double LN2_LO = 5.497923018708371e-14;     /* ln2 - LN2_HI */
    const double INV_LN2 = 1.4426950408889634;
    if (x != x) return x;
    if (x > 709.782712893384) return 1.0 / 0.0;
    if (x < -745.1332191019412) return 0.0;
    double kd = tci_round_half_away(x * INV_LN2);
    int k = (int)kd;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    /* Horner on Σ_{n=0}^{15} r^n / n! */
    double p = 7.647163731819816e-13;                /* 1/15! */
    p = p * r + 1.1470745597729725e-11;              /* 1/14! */
    p = p * r + 1.6059043836821613e-10;              /* 1/13! */
    p = p * r + 2.08767569878681e-09;                /* 1/12! */
    p = p * r + 2.505210838544172e-08;               /* 1/11! */
    p = p * r + 2.755731922398589e-07;               /* 1/10! */
    p = p * r + 2.7557319223985893e-06;              /* 1/9!  */
    p = p * r + 2.48015873015873e-05;                /* 1/8!  */
    p = p * r + 0.0001984126984126984;               /* 1/7!  */
    p = p * r + 0.001388888888888889;                /* 1/6!  */
    p = p * r + 0.008333333333333333;                /* 1/5!  */
    p = p * r + 0.041666666666666664;                /* 1/4!  */
    p = p * r + 0.16666666666666666;                 /* 1/3!  */
    p = p * r + 0.5;                                 /* 1/2!  */
    p = p * r + 1.0;
    p = p * r + 1.0;
    if (k >= -1022) {
        if (k <= 1023) return p * tci_pow2i(k);
        return (p * tci_pow2i(1023)) * tci_pow2i(k - 1023);
    }
    return (p * tci_pow2i(k + 1000)) * tci_pow2i(-1000);
}

/* sin(r), cos(r) for |r| <= π/4 by Taylor series (to r^17 / r^18). */
static inline TCI_HD double tci_sin_kernel(double r)
{
    double r2 = r * r;
    double p = 2.8114572543455206e-15;               /* 1/17! */
    p = p * r2 - 7.647163731819816e-13;              /* 1/15! */
    p = p * r2 + 1.6059043836821613e-10;             /* 1/13! */
    p = p * r2 - 2.505210838544172e-08;              /* 1/11! */
    p = p * r2 + 2.7557319223985893e-06;             /* 1/9!  */
    p = p * r2 - 0.0001984126984126984;              /* 1/7!  */
    p = p * r2 + 0.008333333333333333;               /* 1/5!  */
    p = p * r2 - 0.16666666666666666;                /* 1/3!  */
    p = p * r2;
    return r + r * p;
}

static inline TCI_HD double tci_cos_kernel(double r)
{
    double r2 = r * r;
    double p = 1.5619206968586225e-16;               /* 1/18! */
    p = p * r2 - 4.779477332387385e-14;              /* 1/16! */
    p = p * r2 + 1.1470745597729725e-11;             /* 1/14! */
    p = p * r2 - 2.08767569878681e-09;               /* 1/12! */
    p = p * r2 + 2.755731922398589e-07;              /* 1/10! */
    p = p * r2 - 2.48015873015873e-05;               /* 1/8!  */
    p = p * r2 + 0.001388888888888889;               /* 1/6!  */
    p = p * r2 - 0.041666666666666664;               /* 1/4!  */
    p = p * r2 + 0.5;                                /* 1/2!  */
    return 1.0 - r2 * p;
}

/* cos(x) for |x| < 2^30: x = q·π/2 + r, π/2 split in three 33-bit parts so
 * q·part is exact for |q| < 2^20 (the reduction stays accurate beyond that
 * only to about |x|·2^-86). */
static inline TCI_HD double tci_cos(double x)
{
    const double PIO2_1 = 1.5707963267341256;
    const double PIO2_2 = 6.077100506303966e-11;
    const double PIO2_3 = 2.0222662487959506e-21;
    const double TWO_OVER_PI = 0.6366197723675814;
    if (x != x) return x;
    double qd = tci_round_half_away(x * TWO_OVER_PI);
    double r = ((x - qd * PIO2_1) - qd * PIO2_2) - qd * PIO2_3;
    int64_t q = (int64_t)qd;
    switch ((int)(q & 3)) {
    case 0: return tci_cos_kernel(r);
    case 1: return -tci_sin_kernel(r);
    case 2: return -tci_cos_kernel(r);
    default: return tci_sin_kernel(r);
    }
}

/* ---------------------------------------------------------------------- */
/* Evaluation                                                               */
/* ---------------------------------------------------------------------- */

/* A point σ: sites 0..na-1 from `a`, sites na..L-1 from `b` (1-based values). */
typedef struct tci_point {
    const uint8_t *a;
    const uint8_t *b;
    int32_t na;
    int32_t L;
} tci_point;

static inline TCI_HD int tci_pt(const tci_point *x, int t)
{
    return t < x->na ? (int)x->a[t] : (int)x->b[t - x->na];
}

/* quantics coordinate Σ (σ_i−1) 2^{−i}, exact for L <= 52 */
static inline TCI_HD double tci_quantics_x(const tci_point *x)
{
    int64_t acc = 0;
    for (int t = 0; t < x->L; ++t) acc = acc * 2 + (tci_pt(x, t) - 1);
    double s = (double)acc;
    int L = x->L;
    while (L > 0) { int c = L > 60 ? 60 : L; s = s * (1.0 / (double)((int64_t)1 << c)); L -= c; }
    return s;
}

static inline TCI_HD double tci_functor_eval(int kind, const double *p, const int32_t *ip,
                                             const tci_point *x)
{
    const int L = x->L;
    switch (kind) {
    case TCI_F_LORENTZ_INT: {
        int64_t s = 0;
        for (int t = 0; t < L; ++t) { int64_t v = tci_pt(x, t); s += v * v; }
        return p[0] / (double)(s + 1);
    }
    case TCI_F_LORENTZ_WEIGHTED: {
        double D = p[0], s = 0.0;
        for (int t = 0; t < L; ++t) {
            double u = (double)tci_pt(x, t) / D;
            s = s + p[1 + t] * (u * u);
        }
        return 1.0 / (1.0 + s);
    }
    case TCI_F_QUANTICS_GAUSS_COS: {
        double xc = tci_quantics_x(x);
        int K = (int)p[0];
        double f = 0.0;
        for (int k = 0; k < K; ++k) {
            double a = p[1 + 3 * k], c = p[2 + 3 * k], w = p[3 + 3 * k];
            double u = xc - c;
            f = f + a * tci_exp(-(u * u) / (2.0 * (w * w)));
        }
        double A = p[1 + 3 * K], F = p[2 + 3 * K];
        return f + A * tci_cos(F * xc);
    }
    case TCI_F_GAUSS_SUM: {
        int K = (int)p[0];
        double dm1 = p[1] - 1.0;
        const double *a = p + 2, *w = p + 2 + K, *c = p + 2 + 2 * K;
        double f = 0.0;
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int t = 0; t < L; ++t) {
                double xi = (double)(tci_pt(x, t) - 1) / dm1;
                double u = xi - c[k * L + t];
                s = s + u * u;
            }
            f = f + a[k] * tci_exp(-s / (2.0 * (w[k] * w[k])));
        }
        return f;
    }
    case TCI_F_SUM: {
        int64_t s = 0;
        for (int t = 0; t < L; ++t) s += tci_pt(x, t);
        return p[0] * (double)s;
    }
    case TCI_F_TABLE: {
        int64_t off = 0, stride = 1;
        for (int t = 0; t < L; ++t) { off += (int64_t)(tci_pt(x, t) - 1) * stride; stride *= ip[t]; }
        return p[off];
    }
    case TCI_F_TT: {
        /* running row vector v (length χ_t), at most 64 wide */
        double v[64], w[64];
        const double *core = p;
        int chl = ip[0];
        v[0] = 1.0;
        for (int t = 0; t < L; ++t) {
            int chr = ip[t + 1], d = ip[L + 1 + t], s = tci_pt(x, t) - 1;
            for (int j = 0; j < chr; ++j) {
                double acc = 0.0;
                for (int i = 0; i < chl; ++i)
                    acc = acc + v[i] * core[i + chl * (s + d * j)];
                w[j] = acc;
            }
            for (int j = 0; j < chr; ++j) v[j] = w[j];
            core += (int64_t)chl * d * chr;
            chl = chr;
        }
        return v[0];
    }
    case TCI_F_QUANTICS_EXP: {
        double xc = tci_quantics_x(x);
        return p[0] * tci_exp(p[1] * xc) + p[2] * tci_exp(p[3] * xc);
    }
    default:
        return 0.0 / 0.0;
    }
}

#ifdef __cplusplus
}
#endif

#endif /* TCI_FUNCTORS_H */
