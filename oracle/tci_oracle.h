/*
 * tci_oracle.h — CPU restatement of the reference TCI2 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libtci_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C with IEEE double arithmetic and no FMA contraction
 * (-ffp-contract=off), the Julia reference TensorCrossInterpolation.jl v0.9.19:
 *   rrLU                matrixlu.jl:1-32, 98-204, 217-225, 360-424
 *   MatrixLUCI          matrixluci.jl:1-92
 *   batch evaluation    batcheval.jl:32-61; filltensor tensorci2.jl:290-312
 *   TCI2 sweep          tensorci2.jl:6-72, 83-213, 315-338, 367-461, 510-628,
 *                       700-916, 943-953; util.jl:1-26; sweepstrategies.jl:1-6;
 *                       globalsearch.jl:97-111; globalpivotfinder.jl:143-195;
 *                       abstracttensortrain.jl:69-71, 124-132
 *
 * Where the reference calls BLAS/LAPACK (the T = Pi1·P⁻¹ solve at
 * tensorci2.jl:391, the trsm/gemm in matrixluci.jl:41-64, the matrix chain
 * in abstracttensortrain.jl:131) this restatement uses fixed-order loops; the
 * reference's own results there are only pinned to rounding tolerance.
 * The global pivot finder draws from the project PRNG (SplitMix64-seeded
 * xoshiro256**) instead of Julia's default RNG.
 */
#ifndef TCI_ORACLE_H
#define TCI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (same numeric values as tci_status in include/tci.h) */
enum {
    ORC_OK = 0,
    ORC_ERR_NAN_L = 1,          /* ErrorException("lu.L contains NaNs")        */
    ORC_ERR_NAN_U = 2,          /* ErrorException("lu.U contains NaNs")        */
    ORC_ERR_ARG = 3,            /* ArgumentError                               */
    ORC_ERR_ZERO_MAXSAMPLE = 4, /* ErrorException("maxsamplevalue is zero!")   */
    ORC_ERR_NAN_TENSOR = 5,     /* ErrorException("Error: NaN in tensor T[b]") */
    ORC_ERR_NONSQUARE = 6,      /* ErrorException("Pivot matrix ... not square") */
    ORC_ERR_DIM = 7,            /* DimensionMismatch                           */
    ORC_ERR_GENERIC = 8,        /* other ErrorException                        */
    ORC_ERR_SINGULAR = 9        /* SingularException in P \ Pi1                */
};

typedef struct orc_functor {
    int32_t kind;               /* tci_functor_kind, or 0 = callback           */
    int32_t nparams;
    const double *params;
    int32_t niparams;
    const int32_t *iparams;
    /* kind 0: f(x[0..L-1] 1-based, L, user) */
    double (*callback)(const int32_t *x, int32_t L, void *user);
    void *user;
} orc_functor;

typedef struct orc_opts {
    double tolerance;
    int64_t maxbonddim;         /* INT64_MAX = unlimited                       */
    int32_t maxiter;            /* default 20 (tensorci2.jl:706)               */
    int32_t sweepstrategy;      /* 0 :backandforth, 1 :forward, 2 :backward     */
    int32_t normalizeerror;
    int32_t ncheckhistory;
    int32_t maxnglobalpivot;
    int32_t nsearchglobalpivot;
    double tolmarginglobalsearch;
    int32_t strictlynested;
    int32_t checkconvglobalpivot;
    uint64_t seed;              /* project PRNG seed for the global search     */
    int32_t pivotsearch;        /* 0 :full, 1 :rook (tensorci2.jl:552-595)     */
} orc_opts;

const char *orc_last_error(void);

/* ---- rrLU (matrixlu.jl) ---- */
void orc_submatrixargmax(const double *A, int64_t lda, int64_t m, int64_t n, int64_t k,
                         int64_t *r, int64_t *c);
int orc_rrlu(const double *A, int64_t m, int64_t n, int64_t maxrank, double reltol,
             double abstol, int32_t leftorth, int64_t *rowperm, int64_t *colperm,
             double *L, double *U, int64_t ldu, int64_t *npivot, double *error);
/* arrlu (matrixlu.jl:227-293) of an explicit matrix: I0/J0 the starting
 * rows/columns (1-based), randomsubset draws from the project PRNG seeded
 * with `seed`; L (m×p) and U (p×n) as the reference builds them (the last
 * submatrix's factors, then cols2Lmatrix!/rows2Umatrix! of the rest),
 * rowperm/colperm 1-based. */
int orc_arrlu(const double *A, int64_t m, int64_t n, const int64_t *I0, int64_t nI0, const int64_t *J0,
              int64_t nJ0, int64_t maxrank, double reltol, double abstol, int32_t leftorth, int32_t numrookiter,
              uint64_t seed, int64_t *rowperm, int64_t *colperm, double *L, double *U, int64_t *npivot,
              double *error);
/* MatrixLUCI left/right factors from an rrLU result (matrixluci.jl:40-84) */
void orc_luci_left(int64_t m, int64_t n, int64_t p, int32_t leftorth, const int64_t *rowperm,
                   const int64_t *colperm, const double *L, const double *U, int64_t ldu,
                   double *out /* m×p */);
void orc_luci_right(int64_t m, int64_t n, int64_t p, int32_t leftorth, const int64_t *rowperm,
                    const int64_t *colperm, const double *L, const double *U, int64_t ldu,
                    double *out /* p×n */);

/* ---- batch evaluation (batcheval.jl:32-61) ----
 * left: nl-long 1-based multi-indices (nI of them, row-major nI×nl);
 * right: nr-long (nJ of them); M = L - nl - nr central indices.
 * out: (nI, d_{nl+1}, …, d_{nl+M}, nJ) column-major. */
void orc_batcheval(const orc_functor *f, int32_t L, const int32_t *localdims,
                   const int32_t *left, int64_t nI, int32_t nl, const int32_t *right,
                   int64_t nJ, int32_t nr, double *out);

/* ---- TensorCI2 ---- */
typedef struct orc_tci orc_tci;

orc_tci *orc_tci_new(int32_t L, const int32_t *localdims);
void orc_tci_free(orc_tci *t);
/* TensorCI2{T}(f, localdims, initialpivots) (tensorci2.jl:42-53) */
int orc_tci_init(orc_tci *t, const orc_functor *f, int32_t npiv, const int32_t *pivots);
int orc_addglobalpivots(orc_tci *t, int32_t npiv, const int32_t *pivots);
int orc_updatepivots(orc_tci *t, int32_t b0, const orc_functor *f, int32_t leftorth,
                     double reltol, double abstol, int64_t maxbonddim);
int orc_sweep1site(orc_tci *t, const orc_functor *f, int32_t forward, double reltol,
                   double abstol, int64_t maxbonddim, int32_t updatetensors);
int orc_sweep2site(orc_tci *t, const orc_functor *f, int32_t niter, double abstol,
                   int64_t maxbonddim, int32_t sweepstrategy, int32_t strictlynested,
                   int32_t fillsitetensors);
int orc_fillsitetensors(orc_tci *t, const orc_functor *f);
int orc_optimize(orc_tci *t, const orc_functor *f, const orc_opts *o);
/* crossinterpolate2 = new + init + optimize */
int orc_crossinterpolate2(orc_tci *t, const orc_functor *f, int32_t npiv,
                          const int32_t *pivots, const orc_opts *o);

int32_t orc_tci_len(const orc_tci *t);
int32_t orc_tci_setsize(const orc_tci *t, int32_t which /*0 I,1 J*/, int32_t b0);
void orc_tci_getset(const orc_tci *t, int32_t which, int32_t b0, int32_t *out);
int orc_tci_setset(orc_tci *t, int32_t which, int32_t b0, int32_t n, const int32_t *in);
void orc_tci_sitetensor_dims(const orc_tci *t, int32_t b0, int64_t *dims3);
void orc_tci_sitetensor(const orc_tci *t, int32_t b0, double *out);
double orc_tci_maxsample(const orc_tci *t);
void orc_tci_set_maxsample(orc_tci *t, double v);
int32_t orc_tci_npivoterrors(const orc_tci *t);
void orc_tci_pivoterrors(const orc_tci *t, double *out);
void orc_tci_bonderrors(const orc_tci *t, double *out);
int32_t orc_tci_niter(const orc_tci *t);
void orc_tci_history(const orc_tci *t, int64_t *ranks, double *errors, int64_t *nglobal);
/* DefaultGlobalPivotFinder (globalpivotfinder.jl:143-195) on the current
 * site tensors, PRNG seeded by `seed`; found: nsearch × L; returns count */
int32_t orc_globalpivotfinder(orc_tci *t, const orc_functor *f, double abstol, int32_t nsearch,
                              int32_t maxng, double margin, uint64_t seed, int32_t *found);
/* TT evaluation of the current site tensors (abstracttensortrain.jl:124-132) */
double orc_tci_evaluate(const orc_tci *t, const int32_t *x);
/* f at npts points (npts×L 1-based, row-major), host FP64 */
void orc_eval_points(const orc_functor *f, int32_t L, int64_t npts, const int32_t *pts, double *out);
/* evaluate(tt, x) (abstracttensortrain.jl:124-132) of a TT given by its
 * cores (site b: dims3[3b]×dims3[3b+1]×dims3[3b+2], column-major, stored
 * back to back) at npts points: the row vector times each slice, each
 * product summed in ascending bond order */
void orc_tt_evaluate(int32_t L, const int64_t *dims3, const double *cores, int64_t npts,
                     const int32_t *pts, double *out);
/* _floatingzone (globalsearch.jl:43-95) from each of npts start points
 * (pts: npts×L, 1-based, row-major; the found pivots on return), the
 * largest error seen in errs; the per-point part of estimatetrueerror
 * (globalsearch.jl:18-41).  TT cores as orc_tt_evaluate. */
void orc_floatingzone(const orc_functor *f, int32_t L, const int64_t *dims3, const double *cores, int64_t npts,
                      int32_t *pts, double *errs, int64_t nsweeps, double earlystoptol);
/* One sweep over a tensor train whose cores are stored back to back
 * (site b: linkdims[b]×localdims[b]×linkdims[b+1], column-major), as
 * compress! does in each direction (tensortrain.jl:147-183) and
 * sweep1sitegetindices! does (conversion.jl:73-156).  method 0 :LU
 * (left/right of rrlu), 1 :CI (left/right of MatrixLUCI); leftorthogonal =
 * forward.  linkdims and cores are updated in place (cores repacked with the
 * new bonds).  pivots (optional): per bond in bond order, rowindices then
 * colindices of the factored site (site b forward, b+1 backward), each the
 * new bond dimension's worth of 1-based indices.  pivoterrors (optional, rank(tt)+1
 * entries): the running max of pivoterrors(lu) over the bonds. */
int orc_tt_sweep(int32_t L, int64_t *linkdims, const int32_t *localdims, double *cores, int32_t method,
                 int32_t forward, int64_t maxrank, double reltol, double abstol, int32_t *pivots,
                 double *pivoterrors);
/* compress!(tt, method; tolerance, maxbonddim, normalizeerror)
 * (tensortrain.jl:95-183) for method :LU (0) or :CI (1) */
int orc_tt_compress(int32_t L, int64_t *linkdims, const int32_t *localdims, double *cores, int32_t method,
                    double tolerance, int64_t maxbonddim, int32_t normalizeerror);
/* algorithmic work counters: [0] Pi elements evaluated, [1] rrLU argmax bytes,
 * [2] fused rrLU bytes, [3] Schur flops, [4] number of rrLU, [5] number of evals */
void orc_tci_workstats(const orc_tci *t, double *out6);

int orc_convergencecriterion(int32_t n, const int64_t *ranks, const double *errors,
                             const int64_t *nglobal, double tol, int64_t maxbonddim,
                             int32_t ncheckhistory, int32_t checkconvglobalpivot);

/* the project exp/cos of include/tci_functors.h (for accuracy tests) */
double orc_exp(double x);
double orc_cos(double x);

/* project PRNG (xoshiro256** seeded by SplitMix64) for fixture generation */
void orc_rng_seed(uint64_t s[4], uint64_t seed);
uint64_t orc_rng_next(uint64_t s[4]);

#ifdef __cplusplus
}
#endif
#endif
