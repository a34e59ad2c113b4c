// tci_problem.h — validation of a tci_problem and its device capacities.
// Host-side plumbing shared by the library's submit path (and the tests'
// CPU build of the engine): checks every argument that would otherwise make
// a device read go out of bounds, then fixes the capacities of the solve.
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>

#include "../../include/tci.h"
#include "tci_layout.h"

namespace tci {

// Largest bond dimension any solve of these local dims can reach: the rank
// at bond b is at most min(Π_{i<=b} d_i, Π_{i>b} d_i) (saturated at 2^30).
inline int64_t rank_bound(int L, const int32_t *localdims)
{
    const int64_t sat = (int64_t)1 << 30;
    int64_t best = 1;
    for (int b = 0; b + 1 < L; ++b) {
        int64_t lp = 1, rp = 1;
        for (int i = 0; i <= b && lp < sat; ++i) lp *= localdims[i];
        for (int i = b + 1; i < L && rp < sat; ++i) rp *= localdims[i];
        int64_t r = lp < rp ? lp : rp;
        if (r > sat) r = sat;
        best = r > best ? r : best;
    }
    return best;
}

// Default capacity of every |Iset|/|Jset| (before the global-pivot extras):
// min(maxbonddim, the rank bound, kDefaultChi), lowered (not below
// kMinDefaultChi) until the solve's Pi buffer, mrcap² doubles with
// mrcap = (dmax + 1)·(chi + extras), fits kDefaultPiBytes, so that a batch
// of many solves over large local dimensions stays small in HBM.  A solve
// whose sets outgrow it ends with TCI_ERR_CAPACITY; the host side then
// resubmits it at twice the capacity (up to min(maxbonddim, rank bound)), so
// the default options succeed wherever the reference does.
constexpr int64_t kDefaultChi = 32, kMinDefaultChi = 8;
constexpr int64_t kDefaultPiBytes = (int64_t)8 << 20;
inline int64_t default_chi_cap(const tci_problem &p)
{
    int64_t chi = kDefaultChi;
    if (p.opts.maxbonddim < chi) chi = p.opts.maxbonddim;
    const int64_t rb = rank_bound(p.L, p.localdims);
    if (rb < chi) chi = rb;
    int64_t dmax = 1;
    for (int b = 0; b < p.L; ++b) dmax = p.localdims[b] > dmax ? p.localdims[b] : dmax;
    const int64_t npiv = p.initialpivots ? p.npivots : 1;
    const int64_t extra = p.opts.maxnglobalpivot > npiv ? p.opts.maxnglobalpivot : npiv;
    while (chi > kMinDefaultChi) {
        const int64_t mr = (dmax + 1) * (chi + extra);
        if (mr * mr * 8 <= kDefaultPiBytes) break;
        --chi;
    }
    return chi;
}

// Checks a functor against the local dims it is evaluated on (any L >= 1):
// the dims are in 1..255 and the parameter blocks cover every read the
// functor makes.  Returns TCI_OK or an error with *msg set.
inline int check_functor(const tci_functor &f, int L, const int32_t *localdims, std::string *msg)
{
    auto bad = [&](int code, const char *m) { if (msg) *msg = m; return code; };
    if (L < 1 || !localdims) return bad(TCI_ERR_ARG, "localdims is empty");
    for (int b = 0; b < L; ++b)
        if (localdims[b] < 1 || localdims[b] > 255) return bad(TCI_ERR_ARG, "local dimensions must be in 1..255");
    const int np = f.nparams, nip = f.niparams;
    if (np < 0 || nip < 0) return bad(TCI_ERR_ARG, "negative parameter count");
    if (np > 0 && !f.params) return bad(TCI_ERR_ARG, "functor params is NULL");
    if (nip > 0 && !f.iparams) return bad(TCI_ERR_ARG, "functor iparams is NULL");
    switch (f.kind) {
    case TCI_F_LORENTZ_INT:
    case TCI_F_SUM:
        if (np < 1) return bad(TCI_ERR_ARG, "functor needs 1 parameter");
        break;
    case TCI_F_LORENTZ_WEIGHTED:
        if (np < 1 + L) return bad(TCI_ERR_ARG, "weighted Lorentzian needs 1+L parameters");
        break;
    case TCI_F_QUANTICS_GAUSS_COS: {
        if (L > 52) return bad(TCI_ERR_ARG, "quantics functors support at most 52 sites");
        if (np < 1) return bad(TCI_ERR_ARG, "missing K");
        int K = (int)f.params[0];
        if (K < 0 || (double)K != f.params[0] || np < 1 + 3 * K + 2)
            return bad(TCI_ERR_ARG, "quantics Gaussian+cos parameter block too short");
        break;
    }
    case TCI_F_GAUSS_SUM: {
        if (np < 2) return bad(TCI_ERR_ARG, "missing K, d");
        int K = (int)f.params[0];
        if (K < 0 || (double)K != f.params[0] || (int64_t)np < 2 + 2 * (int64_t)K + (int64_t)K * L)
            return bad(TCI_ERR_ARG, "Gaussian-sum parameter block too short");
        break;
    }
    case TCI_F_TABLE: {
        if (nip < L) return bad(TCI_ERR_ARG, "table functor needs L dims");
        int64_t tot = 1;
        for (int b = 0; b < L; ++b) {
            if (f.iparams[b] != localdims[b]) return bad(TCI_ERR_ARG, "table dims differ from localdims");
            tot *= f.iparams[b];
            if (tot > ((int64_t)1 << 31)) return bad(TCI_ERR_ARG, "table too large");
        }
        if (np < tot) return bad(TCI_ERR_ARG, "table parameter block too short");
        break;
    }
    case TCI_F_TT: {
        if (nip < 2 * L + 1) return bad(TCI_ERR_ARG, "TT functor needs 2L+1 iparams");
        if (f.iparams[0] != 1 || f.iparams[L] != 1) return bad(TCI_ERR_ARG, "TT end bonds must be 1");
        int64_t tot = 0;
        for (int b = 0; b < L; ++b) {
            int cl = f.iparams[b], cr = f.iparams[b + 1], d = f.iparams[L + 1 + b];
            if (cl < 1 || cl > 64 || cr < 1 || cr > 64) return bad(TCI_ERR_ARG, "TT bond dims must be in 1..64");
            if (d != localdims[b]) return bad(TCI_ERR_ARG, "TT local dims differ from localdims");
            tot += (int64_t)cl * d * cr;
        }
        if (np < tot) return bad(TCI_ERR_ARG, "TT parameter block too short");
        break;
    }
    case TCI_F_QUANTICS_EXP:
        if (L > 52) return bad(TCI_ERR_ARG, "quantics functors support at most 52 sites");
        if (np < 4) return bad(TCI_ERR_ARG, "quantics exp needs 4 parameters");
        break;
    default:
        return bad(TCI_ERR_ARG, "unknown functor kind");
    }
    return TCI_OK;
}

// The key coefficients of the memo table (cachedfunction.jl:13-25):
// coef_1 = 1, coef_i = d_{i-1}·coef_{i-1}, as 256-bit numbers (four 64-bit
// limbs each, written to out when non-null).  False when the largest key,
// Π d_i - 1, would not stay below 2^256 - 1 (the reference's overflow check
// with a 256-bit key type, 19-21).
inline bool cache_key_coefs(int L, const int32_t *localdims, uint64_t *out)
{
    uint64_t c[5] = {1, 0, 0, 0, 0};
    for (int i = 0; i < L; ++i) {
        if (out)
            for (int q = 0; q < 4; ++q) out[4 * i + q] = c[q];
        // c *= d_i (five limbs, to see the overflow)
        uint64_t carry = 0;
        const uint64_t d = (uint64_t)localdims[i];
        for (int q = 0; q < 5; ++q) {
            const uint64_t lo = (c[q] & 0xffffffffULL) * d, hi = (c[q] >> 32) * d;
            const uint64_t mid = (lo >> 32) + (hi & 0xffffffffULL);
            const uint64_t plo = (lo & 0xffffffffULL) | (mid << 32);
            const uint64_t phi = (hi >> 32) + (mid >> 32);
            const uint64_t t = plo + carry;
            carry = phi + (t < plo);
            c[q] = t;
        }
        if (carry || c[4]) return false;   // Π d >= 2^256
    }
    return true;   // Π d < 2^256, so every key Π d - 1 < 2^256 - 1 fits
}

// The four set families of a starting state, in SET_I, SET_J, SET_HA_I,
// SET_HA_J order (history entries may be absent).
inline void state_sets(const tci_state &s, const int32_t *cnt[4], const uint8_t *bytes[4])
{
    cnt[0] = s.nI; bytes[0] = s.Iset;
    cnt[1] = s.nJ; bytes[1] = s.Jset;
    cnt[2] = s.nI_hist; bytes[2] = s.Iset_hist;
    cnt[3] = s.nJ_hist; bytes[3] = s.Jset_hist;
}

// The largest set of a starting state (the capacity must hold it).
inline int64_t largest_state_set(const tci_problem &p)
{
    const int32_t *cnt[4];
    const uint8_t *bytes[4];
    state_sets(*p.state, cnt, bytes);
    int64_t big = 0;
    for (int f = 0; f < 4; ++f)
        if (cnt[f])
            for (int b = 0; b < p.L; ++b) big = cnt[f][b] > big ? cnt[f][b] : big;
    return big;
}

// A starting state is well formed: counts >= 0 (at least one entry in Iset
// and Jset), the history given as a pair or not at all, every entry within
// 1..d of its site, maxsamplevalue >= 0.
inline int check_state(const tci_problem &p, std::string *msg)
{
    auto bad = [&](int code, const char *m) { if (msg) *msg = m; return code; };
    const tci_state &s = *p.state;
    const int L = p.L;
    if (!s.nI || !s.nJ) return bad(TCI_ERR_ARG, "state has no Iset/Jset counts");
    if ((s.nI_hist == nullptr) != (s.nJ_hist == nullptr)) return bad(TCI_ERR_ARG, "state history must hold both Iset and Jset");
    if (!(s.maxsamplevalue >= 0.0)) return bad(TCI_ERR_ARG, "state maxsamplevalue must be >= 0");
    const int32_t *cnt[4];
    const uint8_t *bytes[4];
    state_sets(s, cnt, bytes);
    for (int f = 0; f < 4; ++f) {
        if (!cnt[f]) continue;
        const bool iside = (f & 1) == 0;
        int64_t off = 0;
        for (int b = 0; b < L; ++b) {
            const int n = cnt[f][b];
            if (n < 0 || n > 8192) return bad(TCI_ERR_ARG, "state set size out of range");
            if (f < 2 && n < 1) return bad(TCI_ERR_ARG, "state Iset/Jset entries must be non-empty");
            const int len = iside ? b : L - 1 - b;
            if ((int64_t)n * len > 0 && !bytes[f]) return bad(TCI_ERR_ARG, "state set bytes are NULL");
            for (int64_t q = 0; q < (int64_t)n * len; ++q) {
                const int t = (int)(q % len), site = iside ? t : b + 1 + t;
                const int v = bytes[f][off + q];
                if (v < 1 || v > p.localdims[site]) return bad(TCI_ERR_ARG, "state index out of range");
            }
            off += (int64_t)n * len;
        }
    }
    return TCI_OK;
}

// Returns TCI_OK or an error with *msg set.
inline int check_problem(const tci_problem &p, SolveCaps &c, std::string *msg)
{
    auto bad = [&](int code, const char *m) { if (msg) *msg = m; return code; };
    if (p.L < 2) return bad(TCI_ERR_GENERIC, "localdims should have at least 2 elements!");
    if (p.L > 250) return bad(TCI_ERR_ARG, "at most 250 sites are supported");
    if (!p.localdims) return bad(TCI_ERR_ARG, "localdims is NULL");
    const int L = p.L;
    const int fst = check_functor(p.f, L, p.localdims, msg);
    if (fst != TCI_OK) return fst;
    int dmax = 0, sumd = 0;
    for (int b = 0; b < L; ++b) {
        int d = p.localdims[b];
        dmax = d > dmax ? d : dmax;
        sumd += d;
    }
    const int npiv = p.initialpivots ? p.npivots : 1;
    if (npiv < 1) return bad(TCI_ERR_ARG, "at least one initial pivot is required");
    if (p.initialpivots)
        for (int q = 0; q < npiv; ++q)
            for (int b = 0; b < L; ++b) {
                int v = p.initialpivots[(int64_t)q * L + b];
                if (v < 1 || v > p.localdims[b]) return bad(TCI_ERR_ARG, "initial pivot out of range");
            }
    const int np = p.f.nparams, nip = p.f.niparams;
    const tci_options &o = p.opts;
    if (o.maxiter < 0 || o.maxiter > 100000) return bad(TCI_ERR_ARG, "maxiter out of range");
    if (o.ncheckhistory < 1) return bad(TCI_ERR_ARG, "ncheckhistory must be >= 1");
    if (o.maxnglobalpivot < 0 || o.nsearchglobalpivot < 0 || o.nsearchglobalpivot > 4096)
        return bad(TCI_ERR_ARG, "global pivot search sizes out of range");
    if (o.sweepstrategy < 0 || o.sweepstrategy > 2) return bad(TCI_ERR_ARG, "unknown sweep strategy");
    if (o.maxbonddim < 0) return bad(TCI_ERR_ARG, "maxbonddim must be >= 0");
    if (o.chi_cap < 0) return bad(TCI_ERR_ARG, "chi_cap must be >= 0");
    int64_t chi = o.chi_cap > 0 ? o.chi_cap : default_chi_cap(p);
    if (chi < 1) chi = 1;
    if (p.state) {
        const int sst = check_state(p, msg);
        if (sst != TCI_OK) return sst;
        const int64_t big = largest_state_set(p);
        if (big > chi) chi = big;
    }
    if (chi > 8192) return bad(TCI_ERR_ARG, "chi_cap too large");
    int extra = o.maxnglobalpivot > npiv ? o.maxnglobalpivot : npiv;
    c.L = L;
    c.dmax = dmax;
    c.sumd = sumd;
    c.setcap = (int32_t)chi + extra;
    c.mrcap = dmax * c.setcap + c.setcap;
    c.npivinit = npiv;
    c.nparams = np;
    c.niparams = nip;
    c.nsearch = o.nsearchglobalpivot;
    c.maxiter = o.maxiter > 0 ? o.maxiter : 1;
    c.has_init = p.state ? 1 : 0;
    if (o.cache_log2 != 0 && (o.cache_log2 < 4 || o.cache_log2 > 26))
        return bad(TCI_ERR_ARG, "cache_log2 must be 0 or in 4..26");
    if (o.cache_log2 && !cache_key_coefs(L, p.localdims, nullptr))
        return bad(TCI_ERR_ARG, "Overflow in CachedFunction: the domain has more than 2^256 points");
    c.cache_log2 = o.cache_log2;
    if (o.pivotsearch != 0 && o.pivotsearch != 1)
        return bad(TCI_ERR_ARG, "Unknown pivot search strategy. Choose from :rook, :full.");
    c.rook = o.pivotsearch;
    return TCI_OK;
}

static_assert(OPT_NO_FINAL_SWEEP == TCI_FLAG_NO_FINAL_SWEEP, "option bits differ");

// Copy the problem-owned fields into a descriptor (pointers excluded).
inline void fill_desc_scalars(const tci_problem &p, SolveDesc &D)
{
    const tci_options &o = p.opts;
    D.fkind = p.f.kind;
    D.maxiter = o.maxiter;
    D.sweepstrategy = o.sweepstrategy;
    D.normalizeerror = o.normalizeerror;
    D.ncheckhistory = o.ncheckhistory;
    D.maxnglobalpivot = o.maxnglobalpivot;
    D.nsearchglobalpivot = o.nsearchglobalpivot;
    D.strictlynested = o.strictlynested;
    D.checkconvglobalpivot = o.checkconvglobalpivot;
    D.tolerance = o.tolerance;
    D.tolmargin = o.tolmarginglobalsearch;
    D.maxbonddim = o.maxbonddim;
    D.seed = o.seed;
    D.optflags = o.flags;
    D.pivotsearch = o.pivotsearch;
    D.init_maxsample = p.state ? p.state->maxsamplevalue : 0.0;
    D.init_hist = (p.state && p.state->nI_hist) ? 1 : 0;
}

// Pack the inputs (localdims, params, iparams, initial pivots) into a host
// image of the solve's input region: `dst_base` is the image of `dev_base`.
inline void pack_inputs(const tci_problem &p, const SolveCaps &c, const SolveDesc &Ddev,
                        uint8_t *dst_base, const uint8_t *dev_base)
{
    auto at = [&](const void *devp) { return dst_base + ((const uint8_t *)devp - dev_base); };
    memcpy(at(Ddev.localdims), p.localdims, sizeof(int32_t) * c.L);
    if (c.nparams) memcpy(at(Ddev.params), p.f.params, sizeof(double) * c.nparams);
    if (c.niparams) memcpy(at(Ddev.iparams), p.f.iparams, sizeof(int32_t) * c.niparams);
    uint8_t *piv = at(Ddev.initpivots);
    for (int q = 0; q < c.npivinit; ++q)
        for (int b = 0; b < c.L; ++b)
            piv[(int64_t)q * c.L + b] = p.initialpivots ? (uint8_t)p.initialpivots[(int64_t)q * c.L + b] : 1;
    if (c.cache_log2) cache_key_coefs(c.L, p.localdims, (uint64_t *)at(Ddev.cache_coef));
    if (c.has_init && p.state) {
        // the starting sets in the set arena's layout, then the counts
        uint8_t *img = at(Ddev.initsets);
        const int L = c.L;
        const size_t setbytes = (size_t)NSETKIND * L * c.setcap * L;
        memset(img, 0, init_image_bytes(c));
        int32_t *cnts = (int32_t *)(img + setbytes);
        const int32_t *cnt[4];
        const uint8_t *bytes[4];
        state_sets(*p.state, cnt, bytes);
        const int kinds[4] = {SET_I, SET_J, SET_HA_I, SET_HA_J};
        for (int f = 0; f < 4; ++f) {
            if (!cnt[f]) continue;
            const bool iside = (f & 1) == 0;
            int64_t off = 0;
            for (int b = 0; b < L; ++b) {
                const int n = cnt[f][b], len = iside ? b : L - 1 - b;
                uint8_t *dst = img + ((size_t)(kinds[f] * L + b) * c.setcap) * L;
                for (int e = 0; e < n; ++e)
                    for (int t = 0; t < len; ++t) dst[(size_t)e * L + t] = bytes[f][off + (int64_t)e * len + t];
                cnts[kinds[f] * L + b] = n;
                off += (int64_t)n * len;
            }
        }
    }
}

}  // namespace tci
