/*
 * tci.h — C ABI of libtci_hip.so, the MI355X (gfx950) TCI2 sweep.
 *
 * Everything is extern "C", returns a tci_status, throws nothing, and takes
 * plain pointers and sizes.  A Julia `ccall` shim (or ctypes, or C++) binds it;
 * each entry point names the reference interface it replaces
 * (TensorCrossInterpolation.jl v0.9.19, paths relative to its repository).
 *
 * Threading: calls on distinct contexts are independent; calls on one
 * context are serialized by the caller.  Every context owns one HIP stream.
 * There is no CPU fallback: without a usable gfx950 device every compute
 * entry point returns TCI_ERR_DEVICE.
 */
#ifndef TCI_H
#define TCI_H

#include <stdint.h>
#include "tci_functors.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; 1..9 map one-to-one onto the reference's exceptions. */
typedef enum tci_status {
    TCI_OK = 0,
    TCI_ERR_NAN_L = 1,          /* ErrorException("lu.L contains NaNs")    matrixlu.jl:164-166 */
    TCI_ERR_NAN_U = 2,          /* ErrorException("lu.U contains NaNs")    matrixlu.jl:167-169 */
    TCI_ERR_ARG = 3,            /* ArgumentError                           tensorci2.jl:736-754 */
    TCI_ERR_ZERO_MAXSAMPLE = 4, /* ErrorException("maxsamplevalue is zero!") tensorci2.jl:50  */
    TCI_ERR_NAN_TENSOR = 5,     /* ErrorException("Error: NaN in tensor T[b]") tensorci2.jl:439 */
    TCI_ERR_NONSQUARE = 6,      /* "Pivot matrix at bond b is not square!" tensorci2.jl:388    */
    TCI_ERR_DIM = 7,            /* DimensionMismatch                       tensorci2.jl:197-199 */
    TCI_ERR_GENERIC = 8,        /* other ErrorException                    tensorci2.jl:730-732 */
    TCI_ERR_SINGULAR = 9,       /* SingularException from P \ Pi1          tensorci2.jl:391    */
    TCI_ERR_CAPACITY = 10,      /* a set or Pi matrix outgrew the capacities chosen at submit */
    TCI_ERR_OOM = 11,           /* device memory / HBM budget exhausted                       */
    TCI_ERR_DEVICE = 12,        /* no gfx950 device, or a HIP runtime error                   */
    TCI_ERR_PENDING = 13        /* result not ready yet (non-blocking queries)                */
} tci_status;

#define TCI_UNLIMITED INT64_MAX

typedef struct tci_ctx tci_ctx;

/* Library / context ---------------------------------------------------- */
const char *tci_last_error(void);             /* thread-local message of the last failure */
const char *tci_version(void);
/* One context = one device + one HIP stream + an HBM budget for solve state. */
tci_status tci_ctx_create(tci_ctx **out, int32_t device, uint64_t hbm_budget_bytes);
tci_status tci_ctx_destroy(tci_ctx *ctx);
tci_status tci_ctx_synchronize(tci_ctx *ctx);
/* Serving (no counterpart in the reference, whose solves run one at a time):
 * give the context n HIP streams (default-priority) so that batches submitted
 * while others run execute concurrently.  tci_solve_submit puts a batch on
 * an idle stream when there is one.  The HBM budget is shared by all. */
tci_status tci_ctx_set_streams(tci_ctx *ctx, int32_t n);
/* The context's first HIP stream (a hipStream_t), on which the synchronous
 * calls and tci_rrlu_batched_f64 run, so that a caller can order its own
 * work or timing events with them (no counterpart in the reference). */
tci_status tci_ctx_stream(tci_ctx *ctx, void **stream);

/* rrLU ----------------------------------------------------------------- */
/* Replaces rrlu(A; maxrank, reltol, abstol, leftorthogonal)::rrLU{Float64}
 * (matrixlu.jl:217-225, i.e. rrlu! 194-204 and _optimizerrlu! 141-181).
 * Synchronous; every pointer is host memory owned by the caller.
 * A: m×n column-major with leading dimension lda (not modified).
 * Out: rowperm[m], colperm[n] (1-based, rrLU.rowpermutation/colpermutation),
 *      L (m×p, leading dim ldl, cap columns), U (p×n, leading dim ldu >= cap),
 *      *npivot = p <= cap = min(m, n, maxrank), *error = rrLU.error.
 * Pivot order, L, U and error are bit-identical to the reference algorithm. */
tci_status tci_rrlu_f64(tci_ctx *ctx, const double *A, int64_t lda, int64_t m, int64_t n,
                        int64_t maxrank, double reltol, double abstol, int32_t leftorth,
                        int64_t *rowperm, int64_t *colperm, double *L, int64_t ldl,
                        double *U, int64_t ldu, int64_t *npivot, double *error);

/* Replaces MatrixLUCI(A; maxrank, reltol, abstol, leftorthogonal) with its
 * left(luci) and right(luci) (matrixluci.jl:5-7, 40-84): with leftorth,
 * left = colstimespivotinv (m×p, rows of [I; L21·L11⁻¹] in original order)
 * and right = rowmatrix (L11·U, p×n); otherwise left = colmatrix (L·U11,
 * m×p) and right = pivotinvtimesrows ([I U11⁻¹·U12], columns in original
 * order).  The factorization is tci_rrlu_f64's; rowindices/colindices are
 * rowperm[1..p]/colperm[1..p].  Synchronous; host buffers; left has leading
 * dimension ldleft >= m, right ldright >= min(m, n, maxrank). */
tci_status tci_luci_f64(tci_ctx *ctx, const double *A, int64_t lda, int64_t m, int64_t n,
                        int64_t maxrank, double reltol, double abstol, int32_t leftorth,
                        int64_t *rowperm, int64_t *colperm, double *left, int64_t ldleft,
                        double *right, int64_t ldright, int64_t *npivot, double *error);

/* C = A·B (M×K times K×N, column-major, host buffers) on the device, each
 * sum in ascending k.  The matrix products of compress! and of the TT →
 * TensorCI2 conversion (tensortrain.jl:161-180, conversion.jl:122-126),
 * which the reference does with BLAS.  Synchronous. */
tci_status tci_gemm_f64(tci_ctx *ctx, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                        const double *B, int64_t ldb, double *C, int64_t ldc);

/* ---- tensor trains (tensortrain.jl, conversion.jl) ----
 * A tensor train is given by L+1 bond dimensions (linkdims[0] and
 * linkdims[L] are the outer ones, left unchanged) and L local dimensions,
 * with the cores back to back: site b (0-based) is linkdims[b] ×
 * localdims[b] × linkdims[b+1], column-major.  Both entries upload the
 * train once, sweep on the device, and write linkdims and cores back in
 * place (the cores repacked with the new bonds; no site grows, so the
 * buffer the caller passed always holds the result). */
#define TCI_FACTOR_LU 0    /* :LU — left/right of rrlu (matrixlu.jl:374-392)      */
#define TCI_FACTOR_CI 1    /* :CI — left/right of MatrixLUCI (matrixluci.jl:70-84) */
#define TCI_FACTOR_SVD 2   /* :SVD — refused with TCI_ERR_ARG                      */

/* compress!(tt, method; tolerance, maxbonddim, normalizeerror)
 * (tensortrain.jl:147-183 with _factorize at 95-145): left to right with
 * no truncation, then right to left with reltol = tolerance
 * (normalizeerror) or abstol = tolerance, maxrank = maxbonddim. */
tci_status tci_tt_compress_f64(tci_ctx *ctx, int32_t L, int64_t *linkdims, const int32_t *localdims,
                               double *cores, int32_t method, double tolerance, int64_t maxbonddim,
                               int32_t normalizeerror);

/* One sweep of sweep1sitegetindices! (conversion.jl:73-156) — and of each
 * direction of compress! — with the given rrLU options and
 * leftorthogonal = forward.  pivots (optional, 2·Σ_{b=1}^{L-1} linkdims[b]
 * entries on entry): per bond in bond order, rowindices(luci) then
 * colindices(luci) of the factored site (the new bond dimension's worth of
 * 1-based indices each; the site is b when forward, b+1 when backward).
 * pivoterrors (optional, rank(tt)+1
 * entries with rank(tt) the largest inner bond on entry): the running
 * maximum of pivoterrors(luci) over the bonds (conversion.jl:141). */
tci_status tci_tt_sweep_f64(tci_ctx *ctx, int32_t L, int64_t *linkdims, const int32_t *localdims, double *cores,
                            int32_t method, int32_t forward, int64_t maxrank, double reltol, double abstol,
                            int32_t *pivots, double *pivoterrors);

/* One factorization job of a batch (device memory).  A is overwritten with
 * the packed factors exactly as rrlu!(A) leaves it (matrixlu.jl:194-204):
 * tril(A[:,1:p]) holds L (before its unit diagonal is imposed when
 * leftorth), triu(A[1:p,:]) holds U. */
typedef struct tci_lu_job {
    double *A;             /* device, m×n column-major, lda = m, in/out       */
    int32_t *rowperm;      /* device, m entries, 1-based on return             */
    int32_t *colperm;      /* device, n entries, 1-based on return             */
    int64_t m, n;
    int64_t maxrank;
    double reltol, abstol;
    int32_t leftorth;
    int32_t status;        /* out: tci_status of this job                      */
    int64_t npivot;        /* out                                              */
    double error;          /* out                                              */
} tci_lu_job;

/* Batched rrLU of `count` independent matrices of uneven sizes in ONE kernel
 * launch (no counterpart in the reference, which factors one matrix per
 * call; each job's result equals rrlu! on that matrix).  `jobs` is a device
 * array; the call is asynchronous on the context stream. */
tci_status tci_rrlu_batched_f64(tci_ctx *ctx, tci_lu_job *jobs_device, int64_t count);

/* Which kernel the context's standalone rrLU calls (tci_rrlu_f64,
 * tci_rrlu_batched_f64) use; no counterpart in the reference, and the results
 * are the same bits either way.  AUTO (the default): the register-tile
 * factorization (one workgroup) for matrices up to 128×128; the cooperative
 * whole-chip factorization (one workgroup per CU) for matrices of at least
 * 384² elements with sides up to 8192 when a batch holds at most 8 of them;
 * the streaming one-workgroup factorization otherwise.  STREAM: always
 * streaming; TILE: always the register tile (jobs larger than 128×128 are
 * refused with TCI_ERR_ARG); GRID: the cooperative factorization for every
 * job with sides up to 8192, one job after the other.  The no-allocation
 * entry below never uses the cooperative kernel. */
enum { TCI_RRLU_AUTO = 0, TCI_RRLU_STREAM = 1, TCI_RRLU_TILE = 2, TCI_RRLU_GRID = 3 };
tci_status tci_ctx_set_rrlu_algo(tci_ctx *ctx, int32_t algo);

/* The autotuner of the standalone rrLU (no counterpart in the reference).
 * For square matrices of each bucket side (32, 64, ..., 4096; rank
 * min(side/2, 64), reltol 1e-10) it times on this context's device: the
 * register tile against the streaming kernel for a batch of 512 (sides up
 * to 128), the cooperative kernel at 256, 512 and 1024 threads for one
 * matrix, and k cooperative runs one after the other against one streaming
 * launch of k matrices (k = 1, 2, 4, ..., 64).  The winners are stored in
 * the context, and AUTO then follows them: a job of largest side s uses the
 * first entry whose side is at least s (the last one beyond it).  The
 * register tile is used when that entry's tile is set.  The cooperative
 * kernel (with that entry's team) is used when the batch holds no more such
 * jobs than every one of their entries' grid_max_jobs.  Choices change
 * speed only: every kernel gives the same bits.  Writes up to cap entries
 * to table (may be NULL) and their number to *count.  Takes a few seconds
 * and up to about 9 GB of the HBM budget while it runs.
 * tci_ctx_set_rrlu_tuning installs a table (one measured earlier and saved);
 * n = 0 returns AUTO to its fixed rules. */
typedef struct tci_rrlu_tune {
    int32_t side;            /* bucket: square matrices of this side           */
    int32_t tile;            /* 1: the register tile for these (side <= 128)   */
    int32_t grid_team;       /* threads per workgroup of the cooperative kernel */
    int32_t grid_max_jobs;   /* most such jobs the cooperative kernel takes (0: none) */
    float ms_stream;         /* one matrix, streaming kernel                   */
    float ms_grid;           /* one matrix, cooperative kernel at grid_team    */
    float ms_tile_batch;     /* 512 matrices, register tile (side <= 128)      */
    float ms_stream_batch;   /* 512 matrices, streaming kernel (side <= 128)   */
} tci_rrlu_tune;
tci_status tci_ctx_autotune_rrlu(tci_ctx *ctx, tci_rrlu_tune *table, int32_t cap, int32_t *count);
tci_status tci_ctx_set_rrlu_tuning(tci_ctx *ctx, const tci_rrlu_tune *table, int32_t n);

/* tci_rrlu_batched_f64 that allocates no device memory.  No counterpart in
 * the reference: it is for callers that cannot call hipMalloc (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  The
 * job table, permutations and matrices are the caller's, as in
 * tci_rrlu_batched_f64, and the results are the same bits.  Its working data
 * (job order, staging vectors, register-tile scratch) lives in a fixed
 * 32 MiB device array that each GPU holds until the process ends; a batch
 * needing more than that for its order and staging vectors is refused with
 * TCI_ERR_ARG (tile jobs that do not fit the rest go to the streaming
 * kernel).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (different streams or contexts) race; calls on one context are
 * ordered.  Returns after the job shapes are read and the kernels are
 * queued on the context stream. */
tci_status tci_rrlu_batched_noalloc_f64(tci_ctx *ctx, tci_lu_job *jobs_device, int64_t count);

/* Pi-tensor evaluation ------------------------------------------------- */
/* Replaces _batchevaluate_dispatch(Float64, f, localdims, Iset, Jset, Val(M))
 * (batcheval.jl:32-61) for a device functor f (include/tci_functors.h).
 * left: nI multi-indices of length nl (nI×nl bytes, row-major, 1-based);
 * right: nJ multi-indices of length nr; M = L - nl - nr central indices.
 * out: (nI, d_{nl+1}, …, d_{nl+M}, nJ) column-major, host memory.
 * *maxabs_inout = max(|*maxabs_inout|, max|out|) with NaN propagation
 * (updatemaxsample!, tensorci2.jl:397-399).  Synchronous. */
typedef struct tci_functor {
    int32_t kind;          /* tci_functor_kind                                 */
    int32_t nparams;
    const double *params;  /* host                                             */
    int32_t niparams;
    const int32_t *iparams;/* host                                             */
} tci_functor;

tci_status tci_pi_eval_f64(tci_ctx *ctx, const tci_functor *f, int32_t L,
                           const int32_t *localdims, const uint8_t *left, int64_t nI,
                           int32_t nl, const uint8_t *right, int64_t nJ, int32_t nr,
                           double *out, double *maxabs_inout);

/* Tensor-train evaluation ---------------------------------------------- */
/* Replaces evaluate(tt::TensorTrain{Float64}, indexset) / tt(indexset)
 * (abstracttensortrain.jl:124-132), batched over npts points on the device.
 * linkdims[0..L] (linkdims[0] = linkdims[L] = 1), localdims[L]; cores: the L
 * site tensors (linkdims[b] × localdims[b] × linkdims[b+1], column-major)
 * back to back, host memory; pts: npts × L bytes (row-major, 1-based).
 * out[npts].  Each product is summed in ascending bond order.  Synchronous. */
tci_status tci_tt_evaluate_f64(tci_ctx *ctx, int32_t L, const int64_t *linkdims,
                               const int32_t *localdims, const double *cores, int64_t npts,
                               const uint8_t *pts, double *out);

/* The per-point search of estimatetrueerror(tt, f; nsearch, initialpoints)
 * (globalsearch.jl:18-41): _floatingzone(ttcache, f; earlystoptol,
 * nsweeps, initp) (globalsearch.jl:43-95) from each of npts start points,
 * one workgroup per point.  points: npts × L bytes (1-based), the start
 * points on entry and the found pivots on return; errors[npts]: the
 * largest |f − TT| seen from each start.  The caller orders the pairs by
 * descending error and drops repeats (globalsearch.jl:37-40).  TT values
 * are l·T_p[:,σ,:]·R with the left row vector and the right environment of
 * the current point, each sum in ascending bond order.  A NaN error ends
 * that point's search (the reference would sweep without end).  nsweeps:
 * at most that many sweeps (>= 1; TCI_UNLIMITED for the reference's
 * default).  Synchronous; host buffers. */
tci_status tci_tt_floatingzone_f64(tci_ctx *ctx, const tci_functor *f, int32_t L, const int64_t *linkdims,
                                   const int32_t *localdims, const double *cores, int64_t npts, uint8_t *points,
                                   double *errors, int64_t nsweeps, double earlystoptol);

/* Whole TCI2 solves ----------------------------------------------------- */
/* optimize! keyword arguments (tensorci2.jl:700-720) plus two device
 * capacities.  Defaults: tci_options_default(). */
typedef struct tci_options {
    double tolerance;              /* 1e-8                                    */
    int64_t maxbonddim;            /* TCI_UNLIMITED                           */
    int32_t maxiter;               /* 20                                      */
    int32_t sweepstrategy;         /* 0 :backandforth, 1 :forward, 2 :backward */
    int32_t normalizeerror;        /* 1                                       */
    int32_t ncheckhistory;         /* 3                                       */
    int32_t maxnglobalpivot;       /* 5                                       */
    int32_t nsearchglobalpivot;    /* 5                                       */
    double tolmarginglobalsearch;  /* 10.0                                    */
    int32_t strictlynested;        /* 0                                       */
    int32_t checkconvglobalpivot;  /* 1                                       */
    uint64_t seed;                 /* global pivot search PRNG seed            */
    int32_t chi_cap;               /* device capacity for any |Iset|/|Jset|;
                                      0 = min(maxbonddim, rank bound, 32); see
                                      tci_capacity_plan                        */
    int32_t flags;                 /* TCI_FLAG_* bits, 0 by default           */
    /* f memoized on the device as CachedFunction{Float64,K}(f, localdims)
     * does (cachedfunction.jl:8-203): 0 = off, else a table of 2^cache_log2
     * entries (4..26) per solve, keyed by Σ_i coeff_i·(σ_i - 1) with
     * coeff_1 = 1, coeff_i = d_{i-1}·coeff_{i-1} (177-184) held in 256 bits
     * (domains of more than 2^256 points are refused, as the reference
     * refuses a key type that overflows, 19-21).  Values are f's own bits,
     * so results are the same with or without the table. */
    int32_t cache_log2;
    /* updatepivots!'s pivot search (tensorci2.jl:510-595): 0 :full (the
     * whole Pi matrix), 1 :rook (arrlu, matrixlu.jl:227-293, on Pi entries
     * evaluated as the search asks for them; its randomsubset draws come
     * from the solve's PRNG, seeded with `seed`) */
    int32_t pivotsearch;
} tci_options;

/* Diagnostic (no counterpart in the reference): end optimize! after its last
 * iteration without the final sweep1site! (and its _sanitycheck), so the
 * site tensors are those the last fillsitetensors! built (T_b = Pi1·P⁻¹,
 * tensorci2.jl:367-394); lets a caller inspect them. */
#define TCI_FLAG_NO_FINAL_SWEEP 1

void tci_options_default(tci_options *o);

/* One TCI problem: crossinterpolate2(Float64, f, localdims, initialpivots;
 * opts...) (tensorci2.jl:943-953).  Pointers are host memory, copied at
 * submit time. */
/* A solve's starting state instead of initial pivots: the pivot sets of
 * TensorCI2{T}(f, localdims, Iset, Jset) (tensorci2.jl:58-72), or the state
 * of an existing TensorCI2 that optimize! continues from (tensorci2.jl:700;
 * only the last history entry is ever read, sweep2site! 874-877).  Host
 * memory, copied at submit.  Set b (0-based site) holds nI[b] multi-indices
 * of length b (Iset) and nJ[b] of length L-1-b (Jset), 1-based bytes,
 * entries back to back and sets back to back in site order. */
typedef struct tci_state {
    const int32_t *nI;             /* L counts                                 */
    const uint8_t *Iset;
    const int32_t *nJ;             /* L counts                                 */
    const uint8_t *Jset;
    const int32_t *nI_hist;        /* Iset_history[end], Jset_history[end] in  */
    const uint8_t *Iset_hist;      /* the same layout; NULL counts: no history */
    const int32_t *nJ_hist;
    const uint8_t *Jset_hist;
    double maxsamplevalue;         /* > 0: continue with it (optimize!); 0: the
                                      constructor's max |f| over every
                                      Iset[b] ⊕ σ ⊕ Jset[b] (tensorci2.jl:67-69) */
} tci_state;

typedef struct tci_problem {
    int32_t L;
    const int32_t *localdims;
    tci_functor f;
    int32_t npivots;
    const int32_t *initialpivots;  /* npivots × L, 1-based; NULL = [ones(L)]   */
    tci_options opts;
    const tci_state *state;        /* NULL: start from the initial pivots
                                      (tensorci2.jl:42-53); else from this state
                                      and the initial pivots are ignored      */
} tci_problem;

typedef struct tci_batch tci_batch;

/* Capacity schedule of a problem (host only, no device needed).  *chi_used:
 * the set capacity tci_solve_submit gives it (opts.chi_cap, or when that is
 * 0 the default min(maxbonddim, rank bound, 32), the rank bound being
 * max_b min(Π_{i<=b} d_i, Π_{i>b} d_i), lowered to no less than 8 while the
 * solve's Pi buffer would exceed 8 MiB); *chi_limit: min(maxbonddim, rank
 * bound), the capacity at which TCI_ERR_CAPACITY can no longer occur.  The
 * reference has no capacities (tensorci2.jl grows its Vectors): a caller
 * that gets TCI_ERR_CAPACITY for a solve resubmits it with chi_cap doubled,
 * up to chi_limit, which reproduces the reference's behaviour for every
 * maxbonddim (the crossinterpolate2 mirrors in Python and Julia do). */
tci_status tci_capacity_plan(const tci_problem *p, int64_t *chi_used, int64_t *chi_limit);

/* Submit `count` independent solves as one launch: one workgroup per solve,
 * every solve's pivot sets, Pi buffers and site tensors resident in HBM for
 * the whole solve.  Asynchronous on the context stream. */
tci_status tci_solve_submit(tci_ctx *ctx, int32_t count, const tci_problem *problems,
                            tci_batch **out);

/* The solve engine's team autotuner (no counterpart in the reference).  A
 * batch of solves runs with 256 threads per solve by default (two solves per
 * CU); larger Pi matrices can run faster with 512 or 1024, small ones with a
 * single wave of 64 (eight solves per CU), and the best team for one solve
 * alone is not the best for thousands.  This times `batch` copies of
 * `sample` with each team (a warm-up launch and a timed one, kernel time in
 * ms4[0..3] for 64, 256, 512, 1024; may be NULL), keeps the fastest (*team,
 * may be NULL) and records it for batches of at least `batch` solves whose
 * largest Pi-row capacity (the `mrcap` of their largest problem) is at least
 * the sample's.  A batch takes the team of the entry with the largest such
 * capacity and then the largest such batch size, the default team without
 * one.  Results are the same bits with every team.  tci_ctx_set_solve_tuning
 * installs n entries (min_mrcap[i], min_count[i], team[i]) instead (n = 0:
 * the default team for every batch). */
tci_status tci_ctx_autotune_solve(tci_ctx *ctx, const tci_problem *sample, int32_t batch, float *ms4,
                                  int32_t *team);
tci_status tci_ctx_set_solve_tuning(tci_ctx *ctx, const int32_t *min_mrcap, const int32_t *min_count,
                                    const int32_t *team, int32_t n);
/* TCI_OK when finished, TCI_ERR_PENDING while running. */
tci_status tci_batch_query(tci_batch *b);
tci_status tci_batch_wait(tci_batch *b);
/* Fetch results of solve i after the batch finished (host copies). */
typedef struct tci_result {
    int32_t status;                /* tci_status of this solve                */
    int32_t niter;                 /* length of ranks/errors                  */
    int32_t L;
    int32_t npivoterrors;
    double maxsamplevalue;
    double workstats[6];           /* Pi elements, argmax bytes, fused rrLU bytes,
                                      Schur flops, #rrLU, #evaluations        */
} tci_result;
tci_status tci_batch_result(tci_batch *b, int32_t i, tci_result *out);
/* ranks[niter], errors[niter] (normalized as optimize! returns them),
 * nglobalpivots[niter]; any pointer may be NULL. */
tci_status tci_batch_history(tci_batch *b, int32_t i, int64_t *ranks, double *errors,
                             int64_t *nglobalpivots);
/* linkdims[L-1] = |Iset[b+1]| (tensorci2.jl:83-85) */
tci_status tci_batch_linkdims(tci_batch *b, int32_t i, int64_t *linkdims);
/* which = 0: Iset[site] (entries of length site-1), 1: Jset[site] (length
 * L-site), 2 and 3: Iset_history[end][site] and Jset_history[end][site]
 * (tensorci2.jl:12-13; count 0 when there is no history); site is 1-based.
 * out receives count×len bytes, 1-based values. */
tci_status tci_batch_indexset(tci_batch *b, int32_t i, int32_t which, int32_t site,
                              int32_t *count, uint8_t *out, int64_t cap_bytes);
/* site tensor T[site] (χ_{site-1} × d × χ_site column-major). */
tci_status tci_batch_sitetensor(tci_batch *b, int32_t i, int32_t site, int64_t *dims3,
                                double *out, int64_t cap);
tci_status tci_batch_pivoterrors(tci_batch *b, int32_t i, double *out, int64_t cap);
tci_status tci_batch_bonderrors(tci_batch *b, int32_t i, double *out);
/* Evaluates solve i's tensor train where it lies in device memory (its
 * TensorTrain(tci), tensortrain.jl:62-64, then evaluate at each point,
 * abstracttensortrain.jl:124-132), and f through its device functor, at
 * npts points (npts × L bytes, row-major, 1-based, host memory).  Any output
 * may be NULL: tt_out[npts], f_out[npts], *maxerr = max|TT - f| and
 * *maxabsf = max|f| over the points (NaN-propagating).  max|TT - f|/max|f|
 * is the normalized error of the reference (tensorci2.jl:769-770, 849)
 * measured on the points instead of the pivots.  Synchronous. */
tci_status tci_batch_evaluate(tci_batch *b, int32_t i, int64_t npts, const uint8_t *pts,
                              double *tt_out, double *f_out, double *maxerr, double *maxabsf);
/* Solve i's memo table of f (opts.cache_log2): its size, the evaluations
 * asked of it and how many it answered (the reference keeps no counts; this
 * shows what the CachedFunction wrapper saved). */
tci_status tci_batch_cachestats(tci_batch *b, int32_t i, int64_t *entries, int64_t *lookups,
                                int64_t *hits);
/* cachedata(cf) (cachedfunction.jl:198-203) of solve i's memo table: the
 * cached points (cap × L bytes, 1-based, decoded from their keys as
 * decodecachekey does, 188-196) and values.  *count = the entries held;
 * TCI_ERR_ARG when they exceed cap (pass NULL buffers to ask the count). */
tci_status tci_batch_cachedata(tci_batch *b, int32_t i, int64_t cap, uint8_t *points, double *values,
                               int64_t *count);
tci_status tci_batch_free(tci_batch *b);

/* Slot-level serving (no counterpart in the reference, whose solves run one
 * at a time).  A server keeps a persistent grid on the device (two
 * workgroups per CU, a few slots left free for the runtime's copies) that
 * takes published requests one at a time from a ring of `slots` entries, so
 * a request starts as soon as a workgroup is free instead of waiting for a
 * batch.  `shape` is the largest problem expected (it sizes the grid's LDS
 * plan; a larger one still runs, with less on chip).  The grid leaves after
 * `idle_ms` without a running or new request and is started again by the
 * next submit or poll.  team: threads per solve, 256 (two solves per CU,
 * the most requests in flight) or 512 (one per CU, each solve faster; for
 * streams that keep fewer requests in flight than there are CUs); 0: the
 * context's default team.  Its stream has a hardware queue of its own, so the
 * library's own copies never wait behind it; work on the legacy default
 * stream, which waits for every blocking stream, waits for the grid to leave
 * (environment TCI_SERVE_QUEUE=shared: a plain non-blocking stream instead,
 * for processes with spare hardware queues).
 * tci_server_submit stages one problem in pinned host memory and publishes
 * it, with no allocation or stream operation once the server's size classes
 * are warm (the claiming workgroup copies it to the device); the returned
 * tci_batch (one solve) is read with the tci_batch_* accessors
 * (tci_batch_query polls its done mark; tci_batch_elapsed_ms is refused)
 * and freed with tci_batch_free, which returns its memory to the server.  A
 * full ring, or an HBM budget that cannot hold a new block, returns
 * TCI_ERR_OOM.
 * tci_server_poll writes up to cap finished requests not returned by an
 * earlier poll to out (*count of them) and starts the grid again if requests
 * wait and it has left.
 * tci_server_stop runs every published request to completion, copies their
 * reports to the host and releases the server; the batches stay valid.  If
 * the grid fails first, stop returns TCI_ERR_DEVICE and every request that
 * had not run reports TCI_ERR_DEVICE from its queries, waits and results
 * (tci_batch_free still releases it).
 * Calls on one server, and on its batches, are serialized by the caller.
 * One server per device at a time: a second server's grid becomes resident
 * only as the first one's workgroups leave (after its idle_ms without work),
 * and large solve launches on the device share it the same way. */
typedef struct tci_server tci_server;
tci_status tci_server_start(tci_ctx *ctx, const tci_problem *shape, int32_t slots, int32_t idle_ms,
                            int32_t team, tci_server **out);
tci_status tci_server_submit(tci_server *s, const tci_problem *p, tci_batch **out);
tci_status tci_server_poll(tci_server *s, tci_batch **out, int32_t cap, int32_t *count);
/* tci_server_submit for problems[0..count) in order, published together:
 * out[q] receives each request; *accepted = how many were taken.  A full
 * ring or budget ends the burst early with TCI_OK (retire, then submit the
 * rest); any other failure returns its status. */
tci_status tci_server_submit_many(tci_server *s, int32_t count, const tci_problem *problems, tci_batch **out,
                                  int32_t *accepted);
tci_status tci_server_stop(tci_server *s);
/* Device time of the batch's solve kernel (HIP events), milliseconds. */
tci_status tci_batch_elapsed_ms(tci_batch *b, float *ms);
/* Diagnostic (no counterpart in the reference, whose only timing split is the
 * verbosity>2 print at tensorci2.jl:547-550): microseconds solve i spent in
 * each phase — [0] Pi evaluation, [1] rrLU, [2] pivot-set updates in the
 * 2-site sweeps, [3] fillsitetensors!, [4] global pivot search, [5] final
 * sweep1site!, [6] whole solve, [7] the Pi1·P⁻¹ solves inside [3].  [0] and
 * [1] also count the evaluations and rrLUs inside [3] and [5]. */
tci_status tci_batch_profile(tci_batch *b, int32_t i, double *us8);

#ifdef __cplusplus
}
#endif
#endif /* TCI_H */
