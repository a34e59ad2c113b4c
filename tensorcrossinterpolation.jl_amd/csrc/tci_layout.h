// tci_layout.h — memory layout of one solve's device state.
//
// A solve owns three regions:
//   * its inputs (local dims, functor parameters, initial pivots), packed
//     with every other solve's inputs so a batch uploads them in one copy;
//   * the state arena (pivot sets, Pi buffer, site tensors, scratch), sized
//     from the capacities fixed at submit time;
//   * a small "report" block (status, history, errors, set counts) so that a
//     whole batch's reports come back in one copy.
// Offsets inside the input and state regions are 256-byte aligned.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "tci_desc.h"

namespace tci {

struct SolveCaps {
    int32_t L;
    int32_t dmax;
    int32_t sumd;
    int32_t setcap;
    int32_t mrcap;
    int32_t npivinit;
    int32_t nparams;
    int32_t niparams;
    int32_t nsearch;
    int32_t maxiter;
    int32_t has_init;   // the solve starts from given pivot sets (tci_state)
    int32_t cache_log2; // memo table of f: 2^cache_log2 entries, 0 = none
    int32_t rook;       // pivotsearch = :rook (its work arrays)
};

// bytes of a starting-state image: the pivot sets in the set arena's own
// layout (NSETKIND·L sets of setcap entries of L bytes) and their counts
inline size_t init_image_bytes(const SolveCaps &c)
{
    return (size_t)NSETKIND * c.L * c.setcap * c.L + 4 * (size_t)NSETKIND * c.L;
}

inline size_t lay_align(size_t x) { return (x + 255) & ~(size_t)255; }

// Sizes of the three regions of one solve.  When `in`/`base`/`rep` are
// non-null the descriptor's pointers into them are assigned.
inline void layout_solve(const SolveCaps &c, SolveDesc *D, uint8_t *in, uint8_t *base, uint8_t *rep,
                         size_t *input_bytes, size_t *state_bytes, size_t *report_bytes)
{
    const size_t L = (size_t)c.L, sc = (size_t)c.setcap, mr = (size_t)c.mrcap;
    const size_t ncand = (size_t)c.nsearch * (size_t)c.sumd;
    const size_t tstride = sc * (size_t)c.dmax * sc;

    size_t i = 0;
    auto itake = [&](size_t bytes) { size_t at = i; i = lay_align(i + (bytes ? bytes : 1)); return at; };
    size_t o_ld = itake(L * 4);
    size_t o_par = itake((size_t)c.nparams * 8);
    size_t o_ipar = itake((size_t)c.niparams * 4);
    size_t o_piv = itake((size_t)c.npivinit * L);
    size_t o_init = itake(c.has_init ? init_image_bytes(c) : 0);
    size_t o_coef = itake(c.cache_log2 ? L * CACHE_LIMBS * 8 : 0);
    *input_bytes = i;

    size_t o = 0;
    auto take = [&](size_t bytes) { size_t at = o; o = lay_align(o + (bytes ? bytes : 1)); return at; };
    size_t o_sets = take((size_t)NSETKIND * L * sc * L);
    size_t o_rows = take(mr * L);
    size_t o_cols = take(mr * L);
    size_t o_flags = take(mr > sc ? mr : sc);
    size_t o_pi = take(mr * mr * 8);
    size_t o_pmat = take(sc * sc * 8);
    size_t o_rp = take(mr * 4);
    size_t o_cp = take(mr * 4);
    size_t o_ipiv = take(2 * sc * 4);
    size_t o_pv = take(mr * 8);
    size_t o_T = take(L * tstride * 8);
    size_t o_Td = take(3 * L * 8);
    size_t o_work = take(ncand * (1 + 2 * sc) * 8);
    size_t o_gpts = take((2 * (size_t)c.nsearch + ncand) * L);
    size_t o_stage = take(2 * mr * 8);
    const size_t ncache = c.cache_log2 ? (size_t)1 << c.cache_log2 : 0;
    size_t o_cst = take(ncache * 4);
    size_t o_ckey = take(ncache * CACHE_LIMBS * 8);
    size_t o_cval = take(ncache * 8);
    size_t o_rook = take(c.rook ? 6 * mr * 4 : 0);
    *state_bytes = o;

    size_t r = 0;
    auto rtake = [&](size_t bytes) { size_t at = r; r += (bytes + 7) & ~(size_t)7; return at; };
    size_t r_ints = rtake(NINTS * 4);
    size_t r_scal = rtake(NSCAL * 8);
    size_t r_ranks = rtake((size_t)c.maxiter * 8);
    size_t r_nglob = rtake((size_t)c.maxiter * 8);
    size_t r_err = rtake((size_t)c.maxiter * 8);
    size_t r_pe = rtake((sc + 2) * 8);
    size_t r_be = rtake(L * 8);
    size_t r_setn = rtake((size_t)NSETKIND * L * 4);
    *report_bytes = lay_align(r);

    if (!D) return;
    D->L = c.L;
    D->setcap = c.setcap;
    D->mrcap = c.mrcap;
    D->npivinit = c.npivinit;
    D->dmax = c.dmax;
    D->sumd = c.sumd;
    D->maxiter_cap = c.maxiter;
    D->tstride = (int64_t)tstride;
    D->workcap = (int64_t)(ncand * (1 + 2 * sc));
    if (in) {
        D->localdims = (const int32_t *)(in + o_ld);
        D->params = (const double *)(in + o_par);
        D->iparams = (const int32_t *)(in + o_ipar);
        D->initpivots = in + o_piv;
        D->initsets = c.has_init ? in + o_init : nullptr;
        D->cache_coef = c.cache_log2 ? (const uint64_t *)(in + o_coef) : nullptr;
    }
    D->cache_log2 = c.cache_log2;
    if (base && c.cache_log2) {
        D->cache_state = (uint32_t *)(base + o_cst);
        D->cache_keys = (uint64_t *)(base + o_ckey);
        D->cache_vals = (double *)(base + o_cval);
    }
    if (base) {
        D->sets = base + o_sets;
        D->rows = base + o_rows;
        D->cols = base + o_cols;
        D->flags = base + o_flags;
        D->pi = (double *)(base + o_pi);
        D->pmat = (double *)(base + o_pmat);
        D->rowperm = (int32_t *)(base + o_rp);
        D->colperm = (int32_t *)(base + o_cp);
        D->ipiv = (int32_t *)(base + o_ipiv);
        D->pivval = (double *)(base + o_pv);
        D->T = (double *)(base + o_T);
        D->Tdims = (int64_t *)(base + o_Td);
        D->work = (double *)(base + o_work);
        D->gpts = base + o_gpts;
        D->stage = (double *)(base + o_stage);
        D->rook = c.rook ? (int32_t *)(base + o_rook) : nullptr;
    }
    if (rep) {
        D->ints = (int32_t *)(rep + r_ints);
        D->scal = (double *)(rep + r_scal);
        D->ranks = (int64_t *)(rep + r_ranks);
        D->nglob = (int64_t *)(rep + r_nglob);
        D->errors = (double *)(rep + r_err);
        D->pivoterrors = (double *)(rep + r_pe);
        D->bonderrors = (double *)(rep + r_be);
        D->setn = (int32_t *)(rep + r_setn);
    }
}

}  // namespace tci
