// tci_desc.h — per-solve descriptor shared by the host API and the device
// solve engine.  Plain data: the host fills it, copies it to device memory,
// and the engine reads problem data and writes results through its pointers.
//
// The fields mirror the TensorCI2 state (tensorci2.jl:6-40) and the
// optimize! keyword arguments (tensorci2.jl:700-720).
#pragma once
#include <stdint.h>

namespace tci {

// set table ids inside one solve's set arena
//   I[b] = b, J[b] = L + b, history A = 2L.., history B = 4L..
enum { SET_I = 0, SET_J = 1, SET_HA_I = 2, SET_HA_J = 3, SET_HB_I = 4, SET_HB_J = 5, NSETKIND = 6 };

// integer slots in desc.ints
enum {
    INT_STATUS = 0,     // tci_status of the solve
    INT_NITER = 1,      // number of optimize! iterations run
    INT_NPE = 2,        // length of pivoterrors
    INT_HIST = 3,       // 0 no history yet, 1 history in A, 2 history in B
    INT_ERRSITE = 4,    // bond/site that raised an error (1-based), 0 if none
    INT_DONE = 5,       // set to 1 when the engine has finished
    NINTS = 8
};

// double slots in desc.scal
enum {
    SCAL_MAXSAMPLE = 0,
    SCAL_WS_PI = 1,       // Pi elements evaluated
    SCAL_WS_ARG = 2,      // rrLU argmax bytes  8·Σ_{k<s}(m-k)(n-k)
    SCAL_WS_RRLU = 3,     // fused rrLU bytes   8mn + 16·Σ_{k=1..p}(m-k)(n-k) + 8p(m+n)
    SCAL_WS_SCHUR = 4,    // Schur flops        2·Σ_{k=1..p}(m-k)(n-k)
    SCAL_WS_NRRLU = 5,
    SCAL_WS_NEVAL = 6,
    SCAL_TPH = 8,         // 8 phase timers (device clock ticks): eval, rrlu, sets,
                          // fill, global search, sweep1site, total
    SCAL_CACHE_LOOKUPS = 16,   // f evaluations asked of the memo table
    SCAL_CACHE_HITS = 17,      // of which answered from it
    NSCAL = 18
};

// the memo table of f (CachedFunction): one state word per entry (empty,
// being written, ready), the 256-bit key as four 64-bit limbs, the value
enum { CACHE_EMPTY = 0, CACHE_BUSY = 1, CACHE_READY = 2, CACHE_LIMBS = 4 };

// bits of SolveDesc::optflags (the values of tci.h's TCI_FLAG_*)
enum { OPT_NO_FINAL_SWEEP = 1 };

struct SolveDesc {
    // ---- problem ----
    int32_t L;
    int32_t fkind;
    int32_t setcap;          // capacity of every I/J set (entries)
    int32_t mrcap;           // capacity of a Pi matrix side (rows and cols)
    int32_t npivinit;        // initial pivots
    int32_t maxiter;
    int32_t sweepstrategy;   // 0 :backandforth, 1 :forward, 2 :backward
    int32_t normalizeerror;
    int32_t ncheckhistory;
    int32_t maxnglobalpivot;
    int32_t nsearchglobalpivot;
    int32_t strictlynested;
    int32_t checkconvglobalpivot;
    int32_t dmax;
    int32_t sumd;            // Σ_b d_b
    int32_t maxiter_cap;     // capacity of ranks/errors/nglob
    int32_t optflags;        // TCI_FLAG_* of tci_options
    int32_t pivotsearch;     // 0 :full, 1 :rook
    double tolerance;
    double tolmargin;
    int64_t maxbonddim;
    uint64_t seed;
    // ---- inputs (device) ----
    const int32_t *localdims;
    const double *params;
    const int32_t *iparams;
    const uint8_t *initpivots;   // npivinit × L, 1-based
    const uint8_t *initsets;     // starting pivot sets in the set arena's layout,
                                 // then their NSETKIND·L int32 counts; null: none
    double init_maxsample;       // > 0: the starting maxsamplevalue
    int32_t init_hist;           // the starting sets include a history entry (slot A)
    int32_t cache_log2;          // memo table of f: 2^cache_log2 entries, 0 = none
    const uint64_t *cache_coef;  // L × CACHE_LIMBS: the key coefficient of each site
    uint32_t *cache_state;       // 2^cache_log2
    uint64_t *cache_keys;        // 2^cache_log2 × CACHE_LIMBS
    double *cache_vals;          // 2^cache_log2
    int32_t *rook;               // :rook work arrays, 6 × mrcap (RP, CP, I0, J0, a copy, candidates)
    // ---- state (device) ----
    uint8_t *sets;           // 6L sets × setcap entries × L bytes
    int32_t *setn;           // 6L counts
    uint8_t *rows;           // mrcap × L
    uint8_t *cols;           // mrcap × L
    uint8_t *flags;          // mrcap (extras membership)
    double *pi;              // mrcap × mrcap
    double *pmat;            // setcap × setcap (pivot matrix P for Pi1·P⁻¹)
    int32_t *rowperm;        // mrcap
    int32_t *colperm;        // mrcap
    int32_t *ipiv;           // 2·setcap (partial pivoting of Pᵀ, then its permutation)
    double *pivval;          // mrcap
    double *T;               // L site-tensor slots of tstride doubles
    int64_t tstride;
    int64_t *Tdims;          // 3L
    double *work;            // global-search scratch
    int64_t workcap;         // doubles in work
    uint8_t *gpts;           // global-search points: nsearch × L (+ found: nsearch × L)
    double *stage;           // rrLU staging vectors in HBM: 2 × mrcap
    // ---- outputs (device) ----
    int64_t *ranks;
    int64_t *nglob;
    double *errors;
    double *pivoterrors;     // setcap + 1
    double *bonderrors;      // L - 1
    double *scal;            // NSCAL
    int32_t *ints;           // NINTS
};

}  // namespace tci
