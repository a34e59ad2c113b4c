// tci_regtile.h — rrLU with the whole matrix in registers.
//
// The team is a TR×TC grid of threads; thread (tr, tc) keeps A[i, j] for
// i = tr + TR·a (a < RA) and j = tc + TC·b (b < CB) in a register array, so a
// matrix of up to (TR·RA)×(TC·CB) never touches memory while it is factored.
// Rows and columns are not moved: each thread tracks the current position of
// its rows and columns and applies the swap of step k (positions k ↔ pivot)
// to those positions locally.  The argmax key of an entry is its current
// column-major position colpos·m + rowpos, so ties resolve exactly as the
// reference's column-major scan of the physically swapped matrix does
// (matrixlu.jl:1-32), and every entry sees the same operation sequence as in
// addpivot! (matrixlu.jl:114-136): true division of the pivot column (or
// row), then a[i,j] - x[i]*y[j] with separate roundings.
//
// Per pivot step: one team argmax (1 barrier) and one staging barrier; LDS
// traffic per thread is O(RA + CB), not O(RA·CB).
#pragma once

#include "tci_common.h"

namespace tci {

template <class Team, int TR, int RA, int CB>
struct RegLU {
    static constexpr int TC = Team::NT / TR;   // thread columns of the grid
    double a[RA][CB];
    int rp[RA];     // current position of my rows (or INT32_MAX when i >= m)
    int cp[CB];     // current position of my columns
    int tr, tc;
    int m, n;

    ENG_INLINE void setup(Team &tm, int m_, int n_)
    {
        tr = tm.tid % TR;
        tc = tm.tid / TR;
        m = m_;
        n = n_;
#pragma unroll
        for (int i = 0; i < RA; ++i) {
            int r = tr + TR * i;
            rp[i] = r < m ? r : 0x7fffffff;
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            int c = tc + TC * j;
            cp[j] = c < n ? c : 0x7fffffff;
        }
    }

    ENG_INLINE static bool fits(const Team &, int m_, int n_)
    {
        return m_ <= TR * RA && n_ <= TC * CB;
    }

    ENG_INLINE void load(const double *A, int64_t lda)
    {
#pragma unroll
        for (int i = 0; i < RA; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const int r = tr + TR * i, c = tc + TC * j;
                a[i][j] = (r < m && c < n) ? A[r + lda * c] : 0.0 / 0.0;
            }
    }

    // f = c / (1 + rs[i] + cs[j]) with exact integer partial sums of v² over the
    // row and column parts of the point (TCI_F_LORENTZ_INT, whose value is
    // c / (double)(Σ v² + 1) with an exact integer sum); returns this
    // thread's NaN-propagating max|A|
    ENG_INLINE double fill_lorentz(double c, const double *rs, const double *cs)
    {
        double rsv[RA], csv[CB];
#pragma unroll
        for (int i = 0; i < RA; ++i) rsv[i] = rp[i] < m ? rs[rp[i]] : 0.0;
#pragma unroll
        for (int j = 0; j < CB; ++j) csv[j] = cp[j] < n ? cs[cp[j]] : 0.0;
        double lmax = 0.0;
#pragma unroll
        for (int i = 0; i < RA; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                const bool ok = rp[i] < m && cp[j] < n;
                const double v = c / (rsv[i] + csv[j] + 1.0);
                a[i][j] = ok ? v : 0.0 / 0.0;
                const double av = fabs(v);
                lmax = ok ? jl_max(lmax, av) : lmax;
            }
        return lmax;
    }

    // position-ordered columns 0..p-1 (the L part and above), m×p
    ENG_INLINE void store_left(double *Lpos, int p) const
    {
#pragma unroll
        for (int i = 0; i < RA; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j)
                if (rp[i] < m && cp[j] < p) Lpos[rp[i] + (int64_t)m * cp[j]] = a[i][j];
    }

    // The entry with the largest abs2 and, among equal ones, the smallest
    // current column-major position — the first maximum of the reference's
    // column-major scan over the trailing block (positions > lo).  vmax is
    // this thread's largest abs2 over that block (-1: none).  Only threads
    // holding the team maximum search their entries for its smallest position.
    ENG_INLINE ArgBest first_max(Team &tm, double vmax, int lo) const
    {
        const double g = tm.max_nan(vmax);
        ArgBest b;
        b.v = g;
        b.lin = INT64_MAX;
        b.a = 0.0;
        b.orig = 0;
        b.pad = 0;
        if (g >= 0.0 && vmax == g) {
#pragma unroll
            for (int i = 0; i < RA; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    const double x = a[i][j];
                    const int key = cp[j] * m + rp[i];   // < (TC·CB)·(TR·RA) < 2^31
                    const bool hit = rp[i] > lo && rp[i] < m && cp[j] > lo && cp[j] < n &&
                                     x * x == g && key < b.lin;
                    b.lin = hit ? key : b.lin;
                    b.a = hit ? x : b.a;
                    b.orig = hit ? (tr + TR * i) + 65536 * (tc + TC * j) : b.orig;
                }
        }
        return tm.argmin_key(b);
    }

    // Lean full-pivot rrLU for the solve engine (matrixlu.jl:141-181, same
    // pivots, pivot values, error and NaN verdicts as rrlu_team).  Entries of
    // pivoted rows and columns are turned into NaN by giving them NaN
    // multipliers, so the per-entry work is a[i,j] - x[i]*y[j], its square and
    // a maxNum — NaN never wins.  The pivot column/row values staged at step k
    // are final (matrixlu.jl:126-134 never touches them again); with Lorig
    // non-null the L column k is recorded there by original row
    // (Lorig[orig + m·k], m×p).  Requires invalid entries to hold NaN.
    // The staging vectors and the position→original maps live in the team's
    // LDS (tm.txs, tm.tys, tm.tperm: rows at [0, m), columns at [m, m+n)).
    ENG_INLINE int factor_lean(Team &tm, int64_t maxrank, double reltol, double abstol,
                               bool leftorth, double *pivval, double *perr, double *stats,
                               double *Lorig)
    {
        const double NaN = 0.0 / 0.0;
        for (int i = tm.tid; i < m; i += tm.nt) tm.tperm(i) = i;
        for (int j = tm.tid; j < n; j += tm.nt) tm.tperm(m + j) = j;
        double vmax0 = -1.0;
#pragma unroll
        for (int i = 0; i < RA; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j) vmax0 = fmax(vmax0, a[i][j] * a[i][j]);
        ArgBest best = first_max(tm, vmax0, -1);
        int64_t mr = maxrank;
        if (m < mr) mr = m;
        if (n < mr) mr = n;
        double maxerr = 0.0, err = NaN, sarg = 0.0, sschur = 0.0;
        int p = 0, nanx = 0, nany = 0;
        while (p < mr) {
            const int k = p;
            int pr = k, pc = k, orow, ocol;
            if (best.v >= 0.0) {
                pc = (int)(best.lin / m);
                pr = (int)(best.lin - (int64_t)pc * m);
                orow = best.orig & 65535;
                ocol = best.orig >> 16;
            } else {
                orow = tm.tperm(k);
                ocol = tm.tperm(m + k);
            }
            orow = tm.uniform(orow);
            ocol = tm.uniform(ocol);
            sarg += (double)(m - k) * (double)(n - k);
            const double av = best.v < 0.0 ? NaN : best.a;
            const double e = fabs(av);
            err = e;
            if ((e < reltol * maxerr || e < abstol) && p > 0) break;
            maxerr = jl_max(maxerr, e);
#pragma unroll
            for (int i = 0; i < RA; ++i) rp[i] = rp[i] == k ? pr : (rp[i] == pr ? k : rp[i]);
#pragma unroll
            for (int j = 0; j < CB; ++j) cp[j] = cp[j] == k ? pc : (cp[j] == pc ? k : cp[j]);
            if (tm.tid == 0) {
                int32_t t = tm.tperm(k); tm.tperm(k) = tm.tperm(pr); tm.tperm(pr) = t;
                t = tm.tperm(m + k); tm.tperm(m + k) = tm.tperm(m + pc); tm.tperm(m + pc) = t;
                if (pivval) pivval[k] = av;
            }
            if (av != av) { nanx = 1; nany = 1; }
            const double piv = av;
            // pivot column (original ocol) and row (original orow): the
            // register slot (jc, ir) is wave-uniform, so only its branch runs
            const int jc = ocol / TC, ir = orow / TR;
            double gv[RA > CB ? RA : CB];
            if (tc == ocol - TC * jc) {
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    if (j == jc) {
#pragma unroll
                        for (int i = 0; i < RA; ++i) gv[i] = a[i][j];
                    }
#pragma unroll
                for (int i = 0; i < RA; ++i) {
                    if (rp[i] > k && rp[i] < m) {
                        double x = gv[i];
                        if (leftorth) x = x / piv;
                        tm.txs(rp[i]) = x;
                        if (x != x) nanx = 1;
                        if (Lorig) Lorig[(tr + TR * i) + (int64_t)m * k] = x;
                    }
                }
            }
            if (tr == orow - TR * ir) {
#pragma unroll
                for (int i = 0; i < RA; ++i)
                    if (i == ir) {
#pragma unroll
                        for (int j = 0; j < CB; ++j) gv[j] = a[i][j];
                    }
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    if (cp[j] > k && cp[j] < n) {
                        double y = gv[j];
                        if (!leftorth) y = y / piv;
                        tm.tys(cp[j]) = y;
                        if (y != y) nany = 1;
                    }
                }
            }
            tm.sync();
            double xr[RA], yc[CB];
#pragma unroll
            for (int i = 0; i < RA; ++i) xr[i] = (rp[i] > k && rp[i] < m) ? tm.txs(rp[i]) : NaN;
#pragma unroll
            for (int j = 0; j < CB; ++j) yc[j] = (cp[j] > k && cp[j] < n) ? tm.tys(cp[j]) : NaN;
            double vmax = -1.0;
#pragma unroll
            for (int i = 0; i < RA; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    const double t = xr[i] * yc[j];
                    const double u = a[i][j] - t;
                    a[i][j] = u;
                    vmax = fmax(vmax, u * u);
                }
            best = first_max(tm, vmax, k);
            p = k + 1;
            sschur += (double)(m - p) * (double)(n - p);
        }
        // for leftorth L holds the staged x (and the pivots), U the staged y;
        // for !leftorth the other way round (matrixlu.jl:162-169)
        const int nanL = tm.any(leftorth ? nanx : nanx);
        const int nanU = tm.any(leftorth ? nany : nany);
        stats[0] = sarg;
        stats[1] = sschur;
        if (p >= (m < n ? m : n)) err = 0.0;
        *perr = err;
        if (nanL) return -ST_NAN_L;
        if (nanU) return -ST_NAN_U;
        return p;
    }
};

}  // namespace tci
