// tci_engine.h — the TCI2 solve engine: one team (a gfx950 workgroup) runs
// one whole crossinterpolate2 solve with every piece of state in device
// memory: pivot sets, Pi matrices, rrLU, site tensors, global pivot search.
//
// The engine is written once against a "Team" interface:
//   tm.tid(), tm.nt()            thread id / team size
//   tm.lane(), tm.wsize()        lane within a wave / wave width (64 on gfx950)
//   tm.wave(), tm.nwaves()       wave id / waves per team
//   tm.sync()                team barrier
//   tm.argmax(ArgBest)       team-wide argmax, result in every thread
//   tm.max_nan(double)       team-wide NaN-propagating max
//   tm.any(int)              team-wide logical or
//   tm.xs, tm.ys             staging vectors (mrcap doubles each)
//   tm.lpi, tm.lpi_cap       on-chip (LDS) buffer for Pi matrices, in doubles
//   tm.lrows, tm.lcols       on-chip index tables (null: use the HBM ones)
//   Team::NT, WS, NW         team size, wave width, waves (compile-time)
//   tm.uniform(x)            x as a wave-uniform value (all lanes hold the same x)
//   tm.wave_max/min/any/first/bcast, tm.wave_sync(), tm.wave_sync_lds()
//                            wave-level reductions, ballot, and ordering points
//                            (LDS and global, or LDS only) inside one wave
//   tm.tA(w), tB(w), tX(w, r), tY(w, j)  LDS slots of the register-tile rrLU
//   Team::RA, CB             register-tile shape (tci_regtile.h); RA = 0: none
//   tm.stage_lds             xs and ys are in LDS
//   tm.assume_lds(p)         a promise that p points into LDS (lets the device
//                            use LDS instructions; a no-op on the CPU)
// The device implementation lives in tci_kernels.hip.  Every parallel loop
// writes disjoint elements and every reduction is over a total order, so the
// result does not depend on the team size: a one-thread team on the CPU
// (used only by the tests to check the control logic) produces the same bits.
//
// Arithmetic: compiled with -ffp-contract=off; x*y and a-b round separately
// as in the reference (matrixlu.jl:132); divisions are IEEE divisions.
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/tci_functors.h"
#include "tci_desc.h"

#include "tci_common.h"
#include "tci_regtile.h"

namespace tci {

// rrlu!(A) with full pivoting (matrixlu.jl:114-181) on A (m×n, lda = m),
// run by the whole team.  Output: rowperm/colperm (0-based original index at
// each position), pivval[k] = k-th pivot (may be null), *perr = rrLU.error,
// stats[0] = Σ_{k<s}(m-k)(n-k) (argmax passes), stats[1] = Σ_{k=1..p}(m-k)(n-k)
// (Schur updates).  Returns p, or -status for the NaN errors.
template <class Team>
ENG int rrlu_team(Team &tm, double *A, int m, int n, int64_t maxrank, double reltol,
              double abstol, bool leftorth, int32_t *rowperm, int32_t *colperm,
              double *pivval, double *perr, double *stats)
{
    const int64_t lda = m;
    for (int i = tm.tid(); i < m; i += tm.nt()) rowperm[i] = i;
    for (int j = tm.tid(); j < n; j += tm.nt()) colperm[j] = j;
    // first argmax over the whole matrix (submatrixargmax, matrixlu.jl:1-32)
    ArgBest best;
    best.v = -1.0;
    best.lin = INT64_MAX;
    best.a = 0.0;
    for (int j = tm.wave(); j < n; j += tm.nwaves()) {
        const double *col = A + lda * j;
        for (int i = tm.lane(); i < m; i += tm.wsize()) {
            double a = col[i];
            double v = a * a;
            if (v == v) {
                ArgBest c;
                c.v = v;
                c.lin = i + lda * j;
                c.a = a;
                if (arg_better(c, best)) best = c;
            }
        }
    }
    best = tm.argmax(best);
    int64_t mr = maxrank;
    if (m < mr) mr = m;
    if (n < mr) mr = n;
    double maxerr = 0.0, err = 0.0 / 0.0;
    double sarg = 0.0, sschur = 0.0;
    int p = 0;
    while (p < mr) {
        const int k = p;
        int r, c;
        if (best.v < 0.0) { r = k; c = k; }
        else { c = (int)(best.lin / lda); r = (int)(best.lin - (int64_t)c * lda); }
        sarg += (double)(m - k) * (double)(n - k);
        // the candidate carries its value, so no thread re-reads A here
        const double e = best.v < 0.0 ? fabs(A[k + lda * k]) : fabs(best.a);
        err = e;
        if ((e < reltol * maxerr || e < abstol) && p > 0) break;
        maxerr = jl_max(maxerr, e);
        // swaprow!(k, r) then swapcol!(k, c) (matrixlu.jl:98-117), one pass:
        // A'[i,j] = A[σ(i), τ(j)] with σ = (k r), τ = (k c)
        if (r != k)
            for (int j = tm.tid(); j < n; j += tm.nt()) {
                if (j == k || j == c) continue;
                double t = A[k + lda * j];
                A[k + lda * j] = A[r + lda * j];
                A[r + lda * j] = t;
            }
        if (c != k)
            for (int i = tm.tid(); i < m; i += tm.nt()) {
                if (i == k || i == r) continue;
                double t = A[i + lda * k];
                A[i + lda * k] = A[i + lda * c];
                A[i + lda * c] = t;
            }
        if (tm.tid() == 0) {
            double akk = A[k + lda * k], akc = A[k + lda * c];
            double ark = A[r + lda * k], arc = A[r + lda * c];
            A[k + lda * k] = arc;
            A[k + lda * c] = ark;
            A[r + lda * k] = akc;
            A[r + lda * c] = akk;
            int32_t t = rowperm[k]; rowperm[k] = rowperm[r]; rowperm[r] = t;
            t = colperm[k]; colperm[k] = colperm[c]; colperm[c] = t;
        }
        tm.sync();
        const double piv = A[k + lda * k];
        if (tm.tid() == 0 && pivval) pivval[k] = piv;
        // scale the pivot column (leftorth) or row by true division, stage x, y
        if (leftorth) {
            for (int i = k + 1 + tm.tid(); i < m; i += tm.nt()) {
                double x = A[i + lda * k] / piv;
                A[i + lda * k] = x;
                tm.xs[i] = x;
            }
            for (int j = k + 1 + tm.tid(); j < n; j += tm.nt()) tm.ys[j] = A[k + lda * j];
        } else {
            for (int j = k + 1 + tm.tid(); j < n; j += tm.nt()) {
                double y = A[k + lda * j] / piv;
                A[k + lda * j] = y;
                tm.ys[j] = y;
            }
            for (int i = k + 1 + tm.tid(); i < m; i += tm.nt()) tm.xs[i] = A[i + lda * k];
        }
        tm.sync();
        // A[k+1:,k+1:] -= x·yᵀ, unfused, fused with the next argmax
        best.v = -1.0;
        best.lin = INT64_MAX;
        best.a = 0.0;
        // columns over waves, rows over lanes: coalesced, no index division;
        // four rows per lane and up to four columns per wave at a time, so
        // each wave has up to sixteen loads in flight (the candidate order is
        // total, so the visiting order is free)
        const int ws = tm.wsize();
        const int nw = tm.nwaves();
        int j = k + 1 + tm.wave();
        // four columns at a time while they exist (sixteen loads in flight)
        constexpr int NC = 4;
        for (; j + (NC - 1) * nw < n; j += NC * nw) {
            double *cl[NC];
            double yv[NC];
#pragma unroll
            for (int h = 0; h < NC; ++h) {
                cl[h] = A + lda * (j + h * nw);
                yv[h] = tm.ys[j + h * nw];
            }
            int i = k + 1 + tm.lane();
            for (; i + 3 * ws < m; i += 4 * ws) {
                double c[NC * 4], x[4];
#pragma unroll
                for (int h = 0; h < NC; ++h)
#pragma unroll
                    for (int u = 0; u < 4; ++u) c[4 * h + u] = cl[h][i + u * ws];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = tm.xs[i + u * ws];
#pragma unroll
                for (int h = 0; h < NC; ++h)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        double t = x[u] * yv[h];
                        double a = c[4 * h + u] - t;
                        cl[h][i + u * ws] = a;
                        double v = a * a;
                        if (v == v) {
                            ArgBest cnd;
                            cnd.v = v;
                            cnd.lin = (i + u * ws) + lda * (j + h * nw);
                            cnd.a = a;
                            if (arg_better(cnd, best)) best = cnd;
                        }
                    }
            }
            for (; i < m; i += ws)
#pragma unroll
                for (int h = 0; h < NC; ++h) {
                    double t = tm.xs[i] * yv[h];
                    double a = cl[h][i] - t;
                    cl[h][i] = a;
                    double v = a * a;
                    if (v == v) {
                        ArgBest cnd;
                        cnd.v = v;
                        cnd.lin = i + lda * (j + h * nw);
                        cnd.a = a;
                        if (arg_better(cnd, best)) best = cnd;
                    }
                }
        }
        for (; j < n; j += nw) {
            double *col = A + lda * j;
            const double yj = tm.ys[j];
            int i = k + 1 + tm.lane();
            for (; i + 3 * ws < m; i += 4 * ws) {
                double c[4], x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) c[u] = col[i + u * ws];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = tm.xs[i + u * ws];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    double t = x[u] * yj;
                    double a = c[u] - t;
                    col[i + u * ws] = a;
                    double v = a * a;
                    if (v == v) {
                        ArgBest cnd;
                        cnd.v = v;
                        cnd.lin = (i + u * ws) + lda * j;
                        cnd.a = a;
                        if (arg_better(cnd, best)) best = cnd;
                    }
                }
            }
            for (; i < m; i += ws) {
                double t = tm.xs[i] * yj;
                double a = col[i] - t;
                col[i] = a;
                double v = a * a;
                if (v == v) {
                    ArgBest cnd;
                    cnd.v = v;
                    cnd.lin = i + lda * j;
                    cnd.a = a;
                    if (arg_better(cnd, best)) best = cnd;
                }
            }
        }
        best = tm.argmax(best);
        p = k + 1;
        sschur += (double)(m - p) * (double)(n - p);
    }
    // L = tril(A[:,1:p]), U = triu(A[1:p,:]); NaN checks (matrixlu.jl:162-169)
    int nanL = 0, nanU = 0;
    for (int j = tm.wave(); j < p; j += tm.nwaves())
        for (int i = j + tm.lane(); i < m; i += tm.wsize()) {
            double v = A[i + lda * j];
            if (v != v) nanL = 1;
        }
    for (int j = tm.wave(); j < n; j += tm.nwaves()) {
        const int imax = j < p - 1 ? j : p - 1;
        for (int i = tm.lane(); i <= imax; i += tm.wsize()) {
            double v = A[i + lda * j];
            if (v != v) nanU = 1;
        }
    }
    nanL = tm.any(nanL);
    nanU = tm.any(nanU);
    stats[0] = sarg;
    stats[1] = sschur;
    if (p >= (m < n ? m : n)) err = 0.0;   // matrixlu.jl:176-178
    *perr = err;
    tm.sync();
    if (nanL) return -ST_NAN_L;
    if (nanU) return -ST_NAN_U;
    return p;
}


// phases timed by the engine (ticks of tm.clock(), accumulated by thread 0)
enum { PH_EVAL = 0, PH_RRLU = 1, PH_SETS = 2, PH_FILL = 3, PH_GSEARCH = 4, PH_SWEEP1 = 5, PH_TOTAL = 6,
       PH_SOLVE = 7 };

// FK > 0 compiles the engine for one functor kind (the switch folds away);
// FK = 0 dispatches on the descriptor's kind at run time.
template <class Team, int FK = 0>
struct Engine {
    static constexpr int TRA = Team::RA > 0 ? Team::RA : 1, TCB = Team::CB > 0 ? Team::CB : 1;
    static constexpr int TCBN = TCB >= 4 ? TCB / 4 : TCB;
    // the team's register tile, and narrower ones for thin matrices (one
    // block of rows per lane, a quarter of the columns, or both), whose
    // pivot steps cost proportionally less
    using Tile = TileLU<Team, TRA, TCB>;
    using TileR = TileLU<Team, 1, TCB>;
    using TileC = TileLU<Team, TRA, TCBN>;
    using TileRC = TileLU<Team, 1, TCBN>;
    Team &tm;
    const SolveDesc &D;
    EngShared &S;
    int L;

    uint8_t *rows, *cols;   // combined index tables (LDS when the team has them)
    uint8_t *sets;          // pivot sets (LDS when the team has room, else the HBM arena)
    int32_t *setcnt;        // their counts
    int32_t *rperm, *cperm; // rrLU permutations and pivots of the current bond
    double *pvals;          // (LDS when there is room, see perm_area)
    const double *lut;      // Lorentzian values by integer sum (LDS), or null

    ENG Engine(Team &t, const SolveDesc &d, EngShared &s)
        : tm(t), D(d), S(s), L(d.L), rows(t.lrows ? t.lrows : d.rows), cols(t.lcols ? t.lcols : d.cols),
          sets(t.lsets ? t.lsets : d.sets), setcnt(t.lsets ? t.lsetn : d.setn),
          rperm(d.rowperm), cperm(d.colperm), pvals(d.pivval), lut(nullptr) {}

    // lut[s] = c / (double)s, s = 1 .. Σ d², for the Lorentzian functor when
    // the team has room for it (the same division the functor performs)
    ENG void build_lut()
    {
        // (with a memo table every evaluation goes through it instead)
        if (!(FK == TCI_F_LORENTZ_INT || (FK == 0 && D.fkind == TCI_F_LORENTZ_INT)) || !tm.llut || cache_on())
            return;
        int64_t smax = 0;
        for (int t = 0; t < L; ++t) smax += (int64_t)D.localdims[t] * D.localdims[t];
        if (smax + 2 > tm.llut_cap) return;
        const double c = D.params[0];
        for (int64_t q = 1 + tm.tid(); q <= smax + 1; q += tm.nt()) tm.llut[q] = c / (double)q;
        tm.sync();
        lut = tm.llut;
    }

    // place the permutations and pivots of an m×n rrLU in the LDS Pi area at
    // offset off (doubles) when they fit, else in the arena; returns the
    // first free offset after them
    ENG_INLINE int64_t perm_area(int64_t off, int m, int n)
    {
        const int64_t ints = ((int64_t)m + n + 1) / 2;
        const int64_t need = (ints + (m < n ? m : n) + 2) & ~(int64_t)1;
        if (tm.lpi && off + need <= tm.lpi_cap) {
            rperm = (int32_t *)(tm.lpi + off);
            cperm = rperm + m;
            pvals = tm.lpi + off + ints;
            return off + need;
        }
        rperm = D.rowperm;
        cperm = D.colperm;
        pvals = D.pivval;
        return off;
    }

    // a Pi matrix of m×n doubles lives in LDS when it fits, else in HBM
    ENG double *pibuf(int m, int n) const
    {
        return (int64_t)m * n <= tm.lpi_cap ? tm.lpi : D.pi;
    }

    // ---------------------------------------------------------------- sets
    ENG uint8_t *setp(int kind, int b) const
    {
        return sets + ((int64_t)(kind * L + b) * D.setcap) * L;
    }
    ENG int32_t &setn(int kind, int b) const { return setcnt[kind * L + b]; }
    ENG static int ilen(int b) { return b; }                 // |Iset[b]| entries (0-based site b)
    ENG int jlen(int b) const { return L - 1 - b; }          // |Jset[b]| entries

    ENG static bool eq(const uint8_t *a, const uint8_t *b, int len)
    {
        for (int t = 0; t < len; ++t)
            if (a[t] != b[t]) return false;
        return true;
    }

    // addglobalpivots! (tensorci2.jl:193-213): for each pivot q in order and
    // each site b, pushunique!(Iset[b], q[1:b-1]) and pushunique!(Jset[b],
    // q[b+1:end]) (util.jl:16-20).  The 2L sets are independent, and a set
    // receives pivot q's part iff that part is neither in the set as it was
    // nor an earlier pivot's part; so the membership tests run in parallel
    // (one thread per set and pivot, a chunk of pivots at a time) and then
    // one thread per set appends in pivot order.
    ENG int addglobalpivots(const uint8_t *piv, int npiv)
    {
        const int nsets = 2 * L;
        int chunk = D.mrcap / nsets;          // flag bytes available per pivot chunk
        if (chunk < 1) chunk = 1;
        int caperr = 0;
        for (int q0 = 0; q0 < npiv; q0 += chunk) {
            const int nq = npiv - q0 < chunk ? npiv - q0 : chunk;
            for (int t = tm.tid(); t < nsets * nq; t += tm.nt()) {
                const int sidx = t / nq, q = q0 + (t - sidx * nq);
                const bool iside = sidx < L;
                const int b = iside ? sidx : sidx - L;
                const int kind = iside ? SET_I : SET_J;
                const int len = iside ? ilen(b) : jlen(b);
                const int off = iside ? 0 : b + 1;
                const uint8_t *x = piv + (int64_t)q * L + off;
                const uint8_t *base = setp(kind, b);
                const int n = setn(kind, b);
                bool dup = false;
                for (int e = 0; e < n && !dup; ++e) dup = eq(base + (int64_t)e * L, x, len);
                for (int q2 = 0; q2 < q && !dup; ++q2) dup = eq(piv + (int64_t)q2 * L + off, x, len);
                D.flags[t] = dup ? 0 : 1;
            }
            tm.sync();
            for (int sidx = tm.tid(); sidx < nsets; sidx += tm.nt()) {
                const bool iside = sidx < L;
                const int b = iside ? sidx : sidx - L;
                const int kind = iside ? SET_I : SET_J;
                const int len = iside ? ilen(b) : jlen(b);
                const int off = iside ? 0 : b + 1;
                uint8_t *base = setp(kind, b);
                int n = setn(kind, b);
                for (int q = 0; q < nq; ++q) {
                    if (!D.flags[sidx * nq + q]) continue;
                    if (n >= D.setcap) { caperr = 1; break; }
                    const uint8_t *x = piv + (int64_t)(q0 + q) * L + off;
                    for (int t = 0; t < len; ++t) base[(int64_t)n * L + t] = x[t];
                    ++n;
                }
                setn(kind, b) = n;
            }
            if (tm.any(caperr)) {
                if (tm.tid() == 0 && S.status == ST_OK) S.status = ST_CAPACITY;
                tm.sync();
                return S.status;
            }
        }
        return S.status;
    }

    // ------------------------------------------------------ index tables
    // rows = kronecker(Iset[b], d_b) (tensorci2.jl:315-320): entry q = i + nI·s
    ENG_INLINE int kron_rows(int b)
    {
        const int nI = setn(SET_I, b), d = D.localdims[b];
        const uint8_t *I = setp(SET_I, b);
        const int m = nI * d;
        if (m > D.mrcap) return m;   // caller reports ST_CAPACITY
        for (int q = tm.tid(); q < m; q += tm.nt()) {
            int s = q / nI, i = q - s * nI;
            uint8_t *r = rows + (int64_t)q * L;
            for (int t = 0; t < b; ++t) r[t] = I[(int64_t)i * L + t];
            r[b] = (uint8_t)(s + 1);
        }
        return m;
    }

    // cols = kronecker(d_{b+1}, Jset[b+1]) (tensorci2.jl:322-327): q = s + d·j
    ENG_INLINE int kron_cols(int b)
    {
        const int nJ = setn(SET_J, b + 1), d = D.localdims[b + 1];
        const uint8_t *J = setp(SET_J, b + 1);
        const int len = jlen(b + 1);
        const int n = nJ * d;
        if (n > D.mrcap) return n;   // caller reports ST_CAPACITY
        for (int q = tm.tid(); q < n; q += tm.nt()) {
            int j = q / d, s = q - j * d;
            uint8_t *c = cols + (int64_t)q * L;
            c[0] = (uint8_t)(s + 1);
            for (int t = 0; t < len; ++t) c[1 + t] = J[(int64_t)j * L + t];
        }
        return n;
    }

    // union(kron, extras) keeping first-occurrence order.  An extra prefix e
    // (length b+1) lies in kronecker(Iset[b], d_b) iff e[1:b] ∈ Iset[b]; an
    // extra suffix lies in kronecker(d, Jset[b+1]) iff e[2:end] ∈ Jset[b+1].
    // Extras come from an earlier Iset/Jset and are unique already.
    ENG_INLINE int union_extras(bool rowside, int b, int m0, int ekind)
    {
        const int ne = setn(ekind, rowside ? b + 1 : b);
        if (ne == 0) return m0;
        const uint8_t *E = setp(ekind, rowside ? b + 1 : b);
        const int refkind = rowside ? SET_I : SET_J;
        const int refb = rowside ? b : b + 1;
        const int nref = setn(refkind, refb);
        const uint8_t *R = setp(refkind, refb);
        const int cmplen = rowside ? b : jlen(b + 1);
        const int off = rowside ? 0 : 1;
        const int elen = rowside ? b + 1 : jlen(b);
        // membership of every (extra, reference entry) pair in parallel; the
        // flags live in the (idle) LDS Pi area when there is one
        uint8_t *fl = (tm.lpi && tm.lpi_cap * 8 >= ne) ? (uint8_t *)tm.lpi : D.flags;
        for (int e = tm.tid(); e < ne; e += tm.nt()) fl[e] = 1;
        tm.sync();
        for (int64_t t = tm.tid(); t < (int64_t)ne * nref; t += tm.nt()) {
            const int e = (int)(t / nref), i = (int)(t - (int64_t)e * nref);
            if (eq(R + (int64_t)i * L, E + (int64_t)e * L + off, cmplen)) fl[e] = 0;
        }
        tm.sync();
        uint8_t *dst = rowside ? rows : cols;
        int total = 0;
        if (ne <= tm.wsize()) {
            // the kept extras as a lane mask (every wave forms it); wave 0
            // appends them, each at the count of kept extras before it
            const bool keep = tm.lane() < ne && fl[tm.lane()];
            const uint64_t mask = tm.wave_ballot(keep);
            total = __builtin_popcountll(mask);
            if (m0 + total > D.mrcap) return fail(ST_CAPACITY);
            if (tm.wave() == 0 && keep) {
                const int e = tm.lane();
                const int pos = m0 + __builtin_popcountll(mask & (((uint64_t)1 << e) - 1));
                for (int t = 0; t < elen; ++t) dst[(int64_t)pos * L + t] = E[(int64_t)e * L + t];
            }
        } else {
            for (int e = 0; e < ne; ++e) total += fl[e];
            if (m0 + total > D.mrcap) return fail(ST_CAPACITY);
            for (int e = tm.tid(); e < ne; e += tm.nt()) {
                if (!fl[e]) continue;
                int pos = m0;
                for (int e2 = 0; e2 < e; ++e2) pos += fl[e2];
                for (int t = 0; t < elen; ++t) dst[(int64_t)pos * L + t] = E[(int64_t)e * L + t];
            }
        }
        tm.sync();
        return m0 + total;
    }

    // copy a set into the rows or cols table
    ENG void load_set(uint8_t *dst, int kind, int b, int len)
    {
        const int n = setn(kind, b);
        const uint8_t *src = setp(kind, b);
        for (int64_t q = tm.tid(); q < (int64_t)n * len; q += tm.nt()) {
            int64_t e = q / len, t = q - e * len;
            dst[e * L + t] = src[e * L + t];
        }
    }

    // set[kind][b] = table[perm[0..p-1]] (pivot order kept, tensorci2.jl:599-600)
    ENG_INLINE void select_into(int kind, int b, const uint8_t *table, const int32_t *perm, int p, int len)
    {
        uint8_t *dst = setp(kind, b);
        for (int64_t q = tm.tid(); q < (int64_t)p * len; q += tm.nt()) {
            int64_t k = q / len, t = q - k * len;
            dst[k * L + t] = table[(int64_t)perm[k] * L + t];
        }
        if (tm.tid() == 0) setn(kind, b) = p;
    }

    // the memo table is compiled into the generic kernel only (FK == 0): a
    // solve with one runs there
    ENG_INLINE bool cache_on() const { return FK == 0 && D.cache_log2 != 0; }

    // ------------------------------------------------------- memo table of f
    // f(x) through the solve's memo table, as CachedFunction (cachedfunction.jl:
    // 62-110) memoizes it: the key Σ_t coef[t]·(σ_t - 1) (177-184, 256 bits),
    // open addressing with linear probing.  A thread that finds a free entry
    // claims it (empty → busy), computes f and publishes key and value
    // (busy → ready); a thread that meets a busy entry or a full probe run
    // computes f itself without waiting.  f is deterministic, so every path
    // returns f's own bits.  *hit counts the answers from the table.
    ENG_INLINE double cached_f(const tci_point &x, uint64_t &hit)
    {
        uint64_t k[CACHE_LIMBS] = {0, 0, 0, 0};
        for (int t = 0; t < L; ++t) {
            const uint64_t s = (uint64_t)(tci_pt(&x, t) - 1);
            if (s) key_muladd(k, D.cache_coef + CACHE_LIMBS * t, s);
        }
        const uint64_t mask = ((uint64_t)1 << D.cache_log2) - 1;
        uint64_t h = key_hash(k) & mask;
        for (int probe = 0; probe < 64; ++probe, h = (h + 1) & mask) {
            uint32_t st = __atomic_load_n(D.cache_state + h, __ATOMIC_ACQUIRE);
            if (st == CACHE_EMPTY) {
                uint32_t expect = CACHE_EMPTY;
                if (__atomic_compare_exchange_n(D.cache_state + h, &expect, (uint32_t)CACHE_BUSY, false,
                                                __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
                    const double v = tci_functor_eval(FK ? FK : D.fkind, D.params, D.iparams, &x);
                    for (int q = 0; q < CACHE_LIMBS; ++q) D.cache_keys[CACHE_LIMBS * h + q] = k[q];
                    D.cache_vals[h] = v;
                    __atomic_store_n(D.cache_state + h, (uint32_t)CACHE_READY, __ATOMIC_RELEASE);
                    return v;
                }
                st = expect;
            }
            if (st != CACHE_READY) break;   // being written: do not wait
            const uint64_t *kk = D.cache_keys + CACHE_LIMBS * h;
            if (kk[0] == k[0] && kk[1] == k[1] && kk[2] == k[2] && kk[3] == k[3]) {
                ++hit;
                return D.cache_vals[h];
            }
        }
        return tci_functor_eval(FK ? FK : D.fkind, D.params, D.iparams, &x);
    }

    // ---------------------------------------------------------- evaluation
    // out[i + ld·j] = f(rows[i] ⊕ cols[j]) (batcheval.jl:50-57); returns the
    // NaN-propagating max of |out| folded into S.maxsample when fold != 0.
    // transpose != 0 writes out[j + ld·i].
    ENG void eval(const uint8_t *rows, int nl, int m, const uint8_t *cols, int n, double *out,
                  int64_t ld, bool transpose, bool fold)
    {
        const uint64_t t0 = tm.clock();
        double lmax = 0.0;
        const int64_t cnt = (int64_t)m * n;
        if (lut && m <= D.mrcap && n <= D.mrcap) {
            // the Lorentzian through its row and column sums and the table
            // (the same bits as the functor's division)
            lorentz_sums(rows, nl, m, cols, n);
            if (tm.stage_lds) {
                tm.assume_lds(tm.xs);
                tm.assume_lds(tm.ys);
            }
            tm.assume_lds(lut);
            for (int j = tm.wave(); j < n; j += tm.nwaves()) {
                const double cs = tm.ys[j];
                for (int i = tm.lane(); i < m; i += tm.wsize()) {
                    const double v = lut[(int)(tm.xs[i] + cs) + 1];
                    if (out) out[transpose ? (j + ld * i) : (i + ld * j)] = v;
                    lmax = jl_max(lmax, fabs(v));
                }
            }
        } else if (cache_on()) {
            uint64_t hits = 0;
            for (int j = tm.wave(); j < n; j += tm.nwaves())
                for (int i = tm.lane(); i < m; i += tm.wsize()) {
                    tci_point x;
                    x.a = rows + (int64_t)i * L;
                    x.b = cols + (int64_t)j * L;
                    x.na = nl;
                    x.L = L;
                    const double v = cached_f(x, hits);
                    if (out) out[transpose ? (j + ld * i) : (i + ld * j)] = v;
                    lmax = jl_max(lmax, fabs(v));
                }
            if (hits) __atomic_fetch_add(&S.cache_hits, hits, __ATOMIC_RELAXED);
            if (tm.tid() == 0) S.cache_lookups += (uint64_t)cnt;
        } else
        for (int j = tm.wave(); j < n; j += tm.nwaves())
            for (int i = tm.lane(); i < m; i += tm.wsize()) {
                tci_point x;
                x.a = rows + (int64_t)i * L;
                x.b = cols + (int64_t)j * L;
                x.na = nl;
                x.L = L;
                double v = tci_functor_eval(FK ? FK : D.fkind, D.params, D.iparams, &x);
                if (out) out[transpose ? (j + ld * i) : (i + ld * j)] = v;
                lmax = jl_max(lmax, fabs(v));
            }
        if (fold) {
            double g = tm.max_nan(lmax);
            if (tm.tid() == 0) {
                S.maxsample = jl_max(fabs(S.maxsample), g);
                S.ws[0] += (double)cnt;
                S.ws[5] += 1.0;
            }
        }
        tm.sync();
        if (tm.tid() == 0) S.tph[PH_EVAL] += tm.clock() - t0;
    }

    // --------------------------------------------------------------- rrLU
    // the register-tile factorization of A (lda = m) with tile shape T
    template <class T>
    ENG_INLINE int tile_factor(const double *A, int m, int n, int64_t maxrank, double reltol,
                                 double abstol, bool leftorth, double *perr, double *st, double *Lorig)
    {
        T R;
        R.setup(tm, m, n);
        R.load(tm, A, m);
        return R.factor(tm, maxrank, reltol, abstol, leftorth, pvals, perr, st, Lorig, rperm, cperm);
    }

    // rrlu!(A) on the m×n matrix A (lda = m).  With need_left, A's first p
    // columns hold the factor L in position order afterwards (used by
    // colstimespivotinv); otherwise only the permutations, pivots and error
    // are produced.  Small matrices are factored in registers.
    // (not inlined: one copy of the register tile code per kernel)
    ENG_NOINLINE int rrlu(double *A, int m, int n, int64_t maxrank, double reltol, double abstol,
                          bool leftorth, double *perr, bool need_left = false,
                          double *lscratch = nullptr)
    {
        double st[4];
        const uint64_t t0 = tm.clock();
        int p;
        if (Team::RA > 0 && Tile::fits(m, n)) {
            // L recorded by original row in lscratch, then permuted into A
            double *Lorig = need_left ? lscratch : nullptr;
            if (TileRC::fits(m, n)) p = tile_factor<TileRC>(A, m, n, maxrank, reltol, abstol, leftorth, perr, st, Lorig);
            else if (TileR::fits(m, n)) p = tile_factor<TileR>(A, m, n, maxrank, reltol, abstol, leftorth, perr, st, Lorig);
            else if (TileC::fits(m, n)) p = tile_factor<TileC>(A, m, n, maxrank, reltol, abstol, leftorth, perr, st, Lorig);
            else p = tile_factor<Tile>(A, m, n, maxrank, reltol, abstol, leftorth, perr, st, Lorig);
            if (need_left && p > 0) {
                tm.sync();
                for (int t = tm.wave(); t < p; t += tm.nwaves())
                    for (int pos = t + 1 + tm.lane(); pos < m; pos += tm.wsize())
                        A[pos + (int64_t)m * t] = Lorig[rperm[pos] + (int64_t)m * t];
            }
            tm.sync();
        } else {
            p = rrlu_team(tm, A, m, n, maxrank, reltol, abstol, leftorth, rperm, cperm,
                          pvals, perr, st);
        }
        if (tm.tid() == 0) {
            S.tph[PH_RRLU] += tm.clock() - t0;
            S.ws[1] += 8.0 * st[0];
            S.ws[2] += 8.0 * (double)m * (double)n + 16.0 * st[1] +
                       8.0 * (double)(p > 0 ? p : 0) * (double)(m + n);
            S.ws[3] += 2.0 * st[1];
            S.ws[4] += 1.0;
        }
        return p;
    }

    // updateerrors! (tensorci2.jl:161-169): pivoterrors[i] = max(pivoterrors[i],
    // |pivot i| or the last error), zero-padded (util.jl:12-14), element-wise
    // in parallel; ends with a barrier
    ENG_INLINE void updateerrors(int b, int p, double err)
    {
        const int npe = S.npe;
        const int n = npe > p + 1 ? npe : p + 1;
        double *pe = piverrs();
        for (int i = tm.tid(); i < n; i += tm.nt()) {
            const double a = i < npe ? pe[i] : 0.0;
            const double e = i < p ? fabs(pvals[i]) : (i == p ? err : 0.0);
            pe[i] = jl_max(a, e);
        }
        tm.sync();
        if (tm.tid() == 0) {
            D.bonderrors[b] = err;
            S.npe = n;
        }
    }

    // pivoterrors: in the team's shared block while the solve runs when
    // setcap + 1 entries fit there (copied to the arena by finish), else in
    // the arena
    ENG double *piverrs() const { return D.setcap + 1 <= kPivErrCap ? S.pe : D.pivoterrors; }

    ENG_INLINE int fail(int st)
    {
        if (tm.tid() == 0 && S.status == ST_OK) S.status = st;
        tm.sync();
        return S.status;
    }

    // Pi evaluation fused into the register-tile rrLU for the Lorentzian
    // functor: the matrix is built in registers from integer row and column
    // sums of v², never stored.  Returns p or -status like rrlu().
    ENG_INLINE bool lorentz_tile(int m, int n) const
    {
        return (FK == TCI_F_LORENTZ_INT || (FK == 0 && D.fkind == TCI_F_LORENTZ_INT)) && !cache_on() &&
               Team::RA > 0 && Tile::fits(m, n) &&
               m <= D.mrcap && n <= D.mrcap;
    }

    // integer sums of squares of the row parts (first nl entries of each
    // row of r) into xs and of the column parts (L - nl entries of each row
    // of c) into ys (exact in double: at most 250·255² < 2^53); ends with a
    // barrier
    ENG_INLINE void lorentz_sums(const uint8_t *r, int nl, int m, const uint8_t *c, int n)
    {
        for (int q = tm.tid(); q < m; q += tm.nt()) {
            int64_t acc = 0;
            for (int t = 0; t < nl; ++t) { int64_t v = r[(int64_t)q * L + t]; acc += v * v; }
            tm.xs[q] = (double)acc;
        }
        for (int q = tm.tid(); q < n; q += tm.nt()) {
            int64_t acc = 0;
            for (int t = 0; t < L - nl; ++t) { int64_t v = c[(int64_t)q * L + t]; acc += v * v; }
            tm.ys[q] = (double)acc;
        }
        tm.sync();
    }

    ENG_INLINE int eval_rrlu_lorentz(int m, int nl, int n, int64_t maxrank, double reltol,
                                       double abstol, bool leftorth, double *perr)
    {
        const uint64_t t0 = tm.clock();
        lorentz_sums(rows, nl, m, cols, n);
        if (TileRC::fits(m, n)) return lorentz_factor<TileRC>(m, n, maxrank, reltol, abstol, leftorth, perr, t0);
        if (TileR::fits(m, n)) return lorentz_factor<TileR>(m, n, maxrank, reltol, abstol, leftorth, perr, t0);
        if (TileC::fits(m, n)) return lorentz_factor<TileC>(m, n, maxrank, reltol, abstol, leftorth, perr, t0);
        return lorentz_factor<Tile>(m, n, maxrank, reltol, abstol, leftorth, perr, t0);
    }

    template <class T>
    ENG_NOINLINE int lorentz_factor(int m, int n, int64_t maxrank, double reltol, double abstol,
                                    bool leftorth, double *perr, uint64_t t0)
    {
        T R;
        R.setup(tm, m, n);
        const double lmax = R.fill_lorentz(tm, D.params[0], tm.xs, tm.ys, lut);
        const double g = tm.max_nan(lmax);
        if (tm.tid() == 0) {
            S.maxsample = jl_max(fabs(S.maxsample), g);
            S.ws[0] += (double)m * (double)n;
            S.ws[5] += 1.0;
            S.tph[PH_EVAL] += tm.clock() - t0;
        }
        const uint64_t t1 = tm.clock();
        double st[4];
        int p = R.factor(tm, maxrank, reltol, abstol, leftorth, pvals, perr, st, nullptr,
                         rperm, cperm);
        tm.sync();
        if (tm.tid() == 0) {
            S.tph[PH_RRLU] += tm.clock() - t1;
            S.ws[1] += 8.0 * st[0];
            S.ws[2] += 8.0 * (double)m * (double)n + 16.0 * st[1] +
                       8.0 * (double)(p > 0 ? p : 0) * (double)(m + n);
            S.ws[3] += 2.0 * st[1];
            S.ws[4] += 1.0;
        }
        return p;
    }

    // ------------------------------------------------------ updatepivots!
    // tensorci2.jl:510-551, 599-607 (pivotsearch = :full).  ekind < 0: no extras.
    ENG int updatepivots(int b, bool leftorth, double abstol, int ekind_i, int ekind_j)
    {
        const uint64_t t0 = tm.clock();
        int m = kron_rows(b);
        int n = kron_cols(b);
        if (m > D.mrcap || n > D.mrcap) return fail(ST_CAPACITY);
        tm.sync();
        if (ekind_i >= 0) {
            m = union_extras(true, b, m, ekind_i);
            if (S.status) return S.status;
            n = union_extras(false, b, n, ekind_j);
            if (S.status) return S.status;
        }
        double err;
        int p;
        const uint64_t t1 = tm.clock();
        if (lorentz_tile(m, n)) {
            perm_area(0, m, n);
            p = eval_rrlu_lorentz(m, b + 1, n, D.maxbonddim, 1e-14, abstol, leftorth, &err);
        } else {
            double *A = pibuf(m, n);
            perm_area(A == tm.lpi ? (((int64_t)m * n + 1) & ~(int64_t)1) : 0, m, n);
            eval(rows, b + 1, m, cols, n, A, m, false, true);
            p = rrlu(A, m, n, D.maxbonddim, 1e-14, abstol, leftorth, &err);
        }
        if (p < 0) return fail(-p);
        if (p > D.setcap) return fail(ST_CAPACITY);
        const uint64_t t2 = tm.clock();
        select_into(SET_I, b + 1, rows, rperm, p, b + 1);
        select_into(SET_J, b, cols, cperm, p, jlen(b));
        updateerrors(b, p, err);
        tm.sync();
        if (tm.tid() == 0) S.tph[PH_SETS] += (t1 - t0) + (tm.clock() - t2);
        return ST_OK;
    }

    // ---------------------------------------------- :rook pivot search
    // out[i + ld·j] = f(rows[ri[i]] ⊕ cols[ci[j]]) (SubMatrix, tensorci2.jl:
    // 488-503), |values| folded into maxsamplevalue as the reference folds
    // the SubMatrix maximum into it (581); out may be null
    ENG void eval_idx(int nl, const int32_t *ri, int ms, const int32_t *ci, int ns, double *out, int64_t ld)
    {
        const uint64_t t0 = tm.clock();
        double lmax = 0.0;
        uint64_t hits = 0;
        for (int j = tm.wave(); j < ns; j += tm.nwaves())
            for (int i = tm.lane(); i < ms; i += tm.wsize()) {
                tci_point x;
                x.a = rows + (int64_t)ri[i] * L;
                x.b = cols + (int64_t)ci[j] * L;
                x.na = nl;
                x.L = L;
                const double v = cache_on() ? cached_f(x, hits) : tci_functor_eval(FK ? FK : D.fkind, D.params,
                                                                                     D.iparams, &x);
                if (out) out[i + ld * j] = v;
                lmax = jl_max(lmax, fabs(v));
            }
        if (hits) __atomic_fetch_add(&S.cache_hits, hits, __ATOMIC_RELAXED);
        const double g = tm.max_nan(lmax);
        if (tm.tid() == 0) {
            S.maxsample = jl_max(fabs(S.maxsample), g);
            S.ws[0] += (double)ms * (double)ns;
            S.ws[5] += 1.0;
            if (cache_on()) S.cache_lookups += (uint64_t)ms * (uint64_t)ns;
            S.tph[PH_EVAL] += tm.clock() - t0;
        }
        tm.sync();
    }

    // out = the q in [0, N) with fl[q] == 0, ascending (wave 0, a ballot per
    // chunk); returns their count; ends with a barrier
    ENG_INLINE int compact_free(const uint8_t *fl, int N, int32_t *out)
    {
        if (tm.wave() == 0) {
            int base = 0;
            for (int q0 = 0; q0 < N; q0 += tm.wsize()) {
                const int q = q0 + tm.lane();
                const bool keep = q < N && !fl[q];
                const uint64_t mask = tm.wave_ballot(keep);
                if (keep) out[base + __builtin_popcountll(mask & ((((uint64_t)1) << tm.lane()) - 1))] = q;
                base += __builtin_popcountll(mask);
            }
            if (tm.lane() == 0) S.rk[3] = base;
        }
        tm.sync();
        return S.rk[3];
    }

    // fl[q] = (q ∈ set[0..ns)) for q < N; ends with a barrier
    ENG_INLINE void mark(uint8_t *fl, int N, const int32_t *set, int ns)
    {
        for (int q = tm.tid(); q < N; q += tm.nt()) fl[q] = 0;
        tm.sync();
        for (int e = tm.tid(); e < ns; e += tm.nt()) fl[set[e]] = 1;
        tm.sync();
    }

    // pushrandomsubset!(S, 1:N, max(1, |S|)) (util.jl:36-58): k draws from
    // setdiff(1:N, S) in ascending order, each rand(1:length(c)) (the
    // solve's PRNG: 1 + (32-bit draw mod length)) followed by deleteat!.
    // Every thread draws the same numbers from its own copy of the state.
    ENG int push_random_subset(int32_t *set, int ns, int N, int32_t *ca, int32_t *cb)
    {
        uint8_t *fl = D.flags;
        mark(fl, N, set, ns);
        int c = compact_free(fl, N, ca);
        int k = ns > 1 ? ns : 1;
        if (k > c) k = c;
        uint64_t rng[4] = {S.rng[0], S.rng[1], S.rng[2], S.rng[3]};
        int32_t *cur = ca, *nxt = cb;
        for (int i = 0; i < k; ++i) {
            const uint32_t r = (uint32_t)(eng_rng_next(rng) >> 32);
            const int idx = (int)(r % (uint32_t)c);
            if (tm.tid() == 0) set[ns + i] = cur[idx];
            for (int q = tm.tid(); q < c - 1; q += tm.nt()) nxt[q] = cur[q < idx ? q : q + 1];
            tm.sync();
            int32_t *t = cur;
            cur = nxt;
            nxt = t;
            --c;
        }
        if (tm.tid() == 0)
            for (int w = 0; w < 4; ++w) S.rng[w] = rng[w];
        tm.sync();
        return k;
    }

    // arrlu (matrixlu.jl:227-293) on the lazily evaluated Pi of bond b
    // (tensorci2.jl:552-583), numrookiter = 5.  I0/J0 start as the
    // positions of Iset[b+1]/Jset[b] in Icombined/Jcombined (findfirst).
    // The factorization's permutations RP/CP persist between the
    // submatrices as the reference's do (its swaps act on their first
    // |submatrix| entries).  Returns p (the pivots are I0/J0[0..p), their
    // values pvals, *err the last error), 0 for none, or -status.
    ENG int rook(int b, int m, int n, bool leftorth, double abstol, double *err)
    {
        const int mr = D.mrcap;
        int32_t *RP = D.rook, *CP = RP + mr, *I0 = CP + mr, *J0 = I0 + mr, *ta = J0 + mr, *tb = ta + mr;
        const int nl = b + 1;
        int64_t maxrank = D.maxbonddim;
        if (maxrank > m) maxrank = m;
        if (maxrank > n) maxrank = n;
        // positions of Iset[b+1] in Icombined and of Jset[b] in Jcombined
        const int nIs = setn(SET_I, b + 1), nJs = setn(SET_J, b);
        const uint8_t *Is = setp(SET_I, b + 1), *Js = setp(SET_J, b);
        for (int e = tm.tid(); e < nIs + nJs; e += tm.nt()) {
            // one thread per set entry scans the table in order (findfirst)
            const bool isI = e < nIs;
            const uint8_t *x = isI ? Is + (int64_t)e * L : Js + (int64_t)(e - nIs) * L;
            const uint8_t *tab = isI ? rows : cols;
            const int cnt = isI ? m : n, len = isI ? nl : jlen(b);
            int at = INT32_MAX;
            for (int i = 0; i < cnt; ++i)
                if (eq(tab + (int64_t)i * L, x, len)) {
                    at = i;
                    break;
                }
            ta[e] = at;
        }
        tm.sync();
        if (tm.tid() == 0) {
            int a = 0, c = 0;
            for (int e = 0; e < nIs; ++e)
                if (ta[e] != INT32_MAX) I0[a++] = ta[e];
            for (int e = 0; e < nJs; ++e)
                if (ta[nIs + e] != INT32_MAX) J0[c++] = ta[nIs + e];
            S.rk[0] = a;
            S.rk[1] = c;
        }
        for (int i = tm.tid(); i < m; i += tm.nt()) RP[i] = i;
        for (int j = tm.tid(); j < n; j += tm.nt()) CP[j] = j;
        tm.sync();
        bool islowrank = false;
        int p = 0, ms = 0, ns = 0;
        double e = 0.0;
        for (;;) {
            const int w = leftorth ? 1 : 0;
            const int k = leftorth ? push_random_subset(J0, S.rk[1], n, ta, tb)
                                   : push_random_subset(I0, S.rk[0], m, ta, tb);
            if (tm.tid() == 0) S.rk[w] += k;
            tm.sync();
            for (int it = 1; it <= 5; ++it) {
                const bool colmove = ((it % 2 == 0) == leftorth);
                const int nI0 = S.rk[0], nJ0 = S.rk[1];
                const int32_t *ri = colmove ? I0 : RP, *ci = colmove ? CP : J0;
                ms = colmove ? nI0 : m;
                ns = colmove ? n : nJ0;
                double *A = pibuf(ms, ns);
                perm_area(A == tm.lpi ? (((int64_t)ms * ns + 1) & ~(int64_t)1) : 0, ms, ns);
                eval_idx(nl, ri, ms, ci, ns, A, ms);
                p = rrlu(A, ms, ns, maxrank, 1e-14, abstol, leftorth, &e);
                if (p < 0) return p;
                // the swaps of this factorization act on RP/CP's first ms/ns entries
                for (int i = tm.tid(); i < ms; i += tm.nt()) ta[i] = RP[rperm[i]];
                for (int j = tm.tid(); j < ns; j += tm.nt()) tb[j] = CP[cperm[j]];
                tm.sync();
                for (int i = tm.tid(); i < ms; i += tm.nt()) RP[i] = ta[i];
                for (int j = tm.tid(); j < ns; j += tm.nt()) CP[j] = tb[j];
                if (tm.tid() == 0) S.rk[2] = 1;
                tm.sync();
                islowrank |= p < (ms < ns ? ms : ns);
                if (p == nI0 && p == nJ0) {
                    for (int k = tm.tid(); k < p; k += tm.nt())
                        if (RP[k] != I0[k] || CP[k] != J0[k]) S.rk[2] = 0;
                    tm.sync();
                    if (S.rk[2]) break;
                }
                for (int k = tm.tid(); k < p; k += tm.nt()) {
                    I0[k] = RP[k];
                    J0[k] = CP[k];
                }
                tm.sync();
                if (tm.tid() == 0) S.rk[0] = S.rk[1] = p;
                tm.sync();
            }
            if (islowrank || S.rk[0] >= maxrank) break;
        }
        // the finish (matrixlu.jl:270-290): rows outside I0 against J0 and
        // columns outside J0 against I0 are evaluated (the values enter
        // maxsamplevalue); the factors they complete are not needed here
        if (ms < m) {
            mark(D.flags, m, I0, S.rk[0]);
            const int c = compact_free(D.flags, m, ta);
            eval_idx(nl, ta, c, J0, S.rk[1], nullptr, 0);
        }
        if (ns < n) {
            mark(D.flags, n, J0, S.rk[1]);
            const int c = compact_free(D.flags, n, ta);
            eval_idx(nl, I0, S.rk[0], ta, c, nullptr, 0);
        }
        *err = e;
        return p;
    }

    // updatepivots! with pivotsearch = :rook (tensorci2.jl:552-595); the
    // full search when the rook search finds no pivot (584-595).  Kept out
    // of line, apart from the :full path.
    ENG_NOINLINE int updatepivots_rook(int b, bool leftorth, double abstol, int ekind_i, int ekind_j)
    {
        const uint64_t t0 = tm.clock();
        int m = kron_rows(b);
        int n = kron_cols(b);
        if (m > D.mrcap || n > D.mrcap) return fail(ST_CAPACITY);
        tm.sync();
        if (ekind_i >= 0) {
            m = union_extras(true, b, m, ekind_i);
            if (S.status) return S.status;
            n = union_extras(false, b, n, ekind_j);
            if (S.status) return S.status;
        }
        double err;
        const uint64_t t1 = tm.clock();
        const int p = rook(b, m, n, leftorth, abstol, &err);
        if (p < 0) return fail(-p);
        if (p == 0) return updatepivots(b, leftorth, abstol, ekind_i, ekind_j);
        if (p > D.setcap) return fail(ST_CAPACITY);
        const uint64_t t2 = tm.clock();
        int32_t *I0 = D.rook + 2 * D.mrcap, *J0 = I0 + D.mrcap;
        select_into(SET_I, b + 1, rows, I0, p, b + 1);
        select_into(SET_J, b, cols, J0, p, jlen(b));
        updateerrors(b, p, err);
        tm.sync();
        if (tm.tid() == 0) S.tph[PH_SETS] += (t1 - t0) + (tm.clock() - t2);
        return ST_OK;
    }

    // LDS scratch carved from the team's Pi area; null when it does not fit
    ENG double *lds_take(int64_t n, int64_t &off) const
    {
        if (!tm.lpi || off + n > tm.lpi_cap) return nullptr;
        double *p = tm.lpi + off;
        off += (n + 1) & ~(int64_t)1;
        return p;
    }

    ENG double *Tslot(int b) const { return D.T + D.tstride * b; }
    ENG void setTdims(int b, int64_t a, int64_t d, int64_t c)
    {
        if (tm.tid() == 0) { D.Tdims[3 * b] = a; D.Tdims[3 * b + 1] = d; D.Tdims[3 * b + 2] = c; }
    }

    // ------------------------------------------------- Pi1·P⁻¹ (site tensor)
    // T_b = transpose(transpose(P) \ transpose(Pi1)) (tensorci2.jl:391): LU
    // with partial pivoting of A = Pᵀ (c×c, first maximum |.| per column) by
    // one wave, then forward and back substitution of each of the r rows of
    // Pi1 by one lane.
    //
    // lu_wave: lanes stride over rows for the pivot search, the scaling and
    // the update (each element sees A[i,j] - A[i,k]·A[k,j] once per k, as in
    // the serial elimination) and over columns for the row swap.  ipiv: 2c
    // ints, the pivots and then the permutation they compose.  Returns 1 when
    // a pivot is zero (LAPACK's SingularException).
    ENG int lu_wave(double *A, int c, int32_t *ipiv, bool onchip)
    {
        int singular = 0;
        for (int k = 0; k < c; ++k) {
            // first maximum |A[i,k]| over i >= k (strictly greater wins, a
            // NaN never does; a NaN diagonal keeps row k)
            const double akk = fabs(A[k + (int64_t)c * k]);
            int pr = k;
            if (akk == akk) {
                double lv = -1.0;
                uint32_t li = 0xffffffffu;
                for (int i = k + tm.lane(); i < c; i += tm.wsize()) {
                    const double v = fabs(A[i + (int64_t)c * k]);
                    if (v > lv) { lv = v; li = (uint32_t)i; }
                }
                const double g = tm.wave_max(lv);
                pr = (int)tm.wave_min(lv == g ? li : 0xffffffffu);
            }
            if (tm.lane() == 0) ipiv[k] = pr;
            if (pr != k)
                for (int j = tm.lane(); j < c; j += tm.wsize()) {
                    const double x = A[k + (int64_t)c * j];
                    A[k + (int64_t)c * j] = A[pr + (int64_t)c * j];
                    A[pr + (int64_t)c * j] = x;
                }
            if (onchip) tm.wave_sync_lds(); else tm.wave_sync();
            const double d = A[k + (int64_t)c * k];
            if (d == 0.0) singular = 1;
            for (int i = k + 1 + tm.lane(); i < c; i += tm.wsize())
                A[i + (int64_t)c * k] = A[i + (int64_t)c * k] / d;
            if (onchip) tm.wave_sync_lds(); else tm.wave_sync();
            for (int i = k + 1 + tm.lane(); i < c; i += tm.wsize()) {
                const double l = A[i + (int64_t)c * k];
                for (int j = k + 1; j < c; ++j)
                    A[i + (int64_t)c * j] = A[i + (int64_t)c * j] - l * A[k + (int64_t)c * j];
            }
            if (onchip) tm.wave_sync_lds(); else tm.wave_sync();
        }
        if (tm.lane() == 0) {
            // the row interchanges as one permutation: column i of the
            // permuted right-hand side is column perm[i] of Pi1
            int32_t *perm = ipiv + c;
            for (int i = 0; i < c; ++i) perm[i] = i;
            for (int k = 0; k < c; ++k) {
                const int pk = ipiv[k], t = perm[k];
                perm[k] = perm[pk];
                perm[pk] = t;
            }
        }
        if (onchip) tm.wave_sync_lds(); else tm.wave_sync();
        return singular;
    }

    // Row q of X = Pi1·A⁻¹ in place over Pi1 (r rows, c columns), with the
    // LU and pivots of lu_wave.  x: scratch for the row (stride sx) or null,
    // in which case the row interchanges are applied to the row itself in
    // pivot order (the same permutation as perm).
    ENG void subst_row(double *Pi1, int r, int q, const double *A, int c, const int32_t *ipiv,
                       double *x, int64_t sx) const
    {
        const bool scratch = x != nullptr;
        if (scratch) {
            const int32_t *perm = ipiv + c;
            for (int i = 0; i < c; ++i) x[sx * i] = Pi1[q + (int64_t)r * perm[i]];
        } else {
            x = Pi1 + q;
            sx = r;
            for (int k = 0; k < c; ++k) {
                const int pk = ipiv[k];
                if (pk != k) {
                    const double t = x[sx * k];
                    x[sx * k] = x[sx * pk];
                    x[sx * pk] = t;
                }
            }
        }
        for (int i = 0; i < c; ++i) {
            double s = x[sx * i];
            for (int t = 0; t < i; ++t) s = s - A[i + (int64_t)c * t] * x[sx * t];
            x[sx * i] = s;
        }
        for (int i = c - 1; i >= 0; --i) {
            double s = x[sx * i];
            for (int t = i + 1; t < c; ++t) s = s - A[i + (int64_t)c * t] * x[sx * t];
            x[sx * i] = s / A[i + (int64_t)c * i];
        }
        if (scratch)
            for (int i = 0; i < c; ++i) Pi1[q + (int64_t)r * i] = x[sx * i];
    }

    // lu_wave and subst_row for c <= CR <= the wave width, with the rows in
    // registers: lane i holds row i of A during the elimination (the same
    // operations per element and the same pivot choice as lu_wave), then the
    // factors go to the wave's LDS block Ap, padded to CR×CR with the
    // identity, and lane q substitutes row q of Pi1 (r rows, in place) with
    // its c values in registers; a padded term subtracts 0·0, which leaves
    // every value unchanged.  ipiv: 2c ints as in lu_wave.  Returns 1 when a
    // pivot is zero.
    template <int CR>
    ENG_NOINLINE int site_solve_reg(const double *At, int c, int32_t *ipiv, double *Pi1, int r, double *Ap)
    {
        const int ln = tm.lane();
        const bool mine = ln < c;
        const int lr = mine ? ln : 0;
        double a[CR];
#pragma unroll
        for (int j = 0; j < CR; ++j) {
            const double v = At[lr + (int64_t)c * (j < c ? j : 0)];
            a[j] = (mine && j < c) ? v : 0.0;
        }
        int singular = 0;
#pragma unroll
        for (int k = 0; k < CR; ++k) {
            if (k >= c) continue;
            // first maximum |A[i,k]| over i >= k (a NaN never wins; a NaN
            // diagonal keeps row k)
            const double akk = tm.wave_bcast_f64(a[k], k);
            int pr = k;
            if (akk == akk) {
                const double v = fabs(a[k]);
                const double lv = (ln >= k && mine && v == v) ? v : -1.0;
                const double g = tm.wave_max(lv);
                pr = (int)tm.wave_min(lv == g ? (uint32_t)ln : 0xffffffffu);
            }
            pr = tm.uniform(pr);
            if (ln == 0) ipiv[k] = pr;
            if (pr != k) {
#pragma unroll
                for (int j = 0; j < CR; ++j) {
                    const double xk = tm.wave_bcast_f64(a[j], k), xp = tm.wave_bcast_f64(a[j], pr);
                    a[j] = ln == k ? xp : (ln == pr ? xk : a[j]);
                }
            }
            const double d = tm.wave_bcast_f64(a[k], k);
            if (d == 0.0) singular = 1;
            const bool below = ln > k && mine;
            if (below) a[k] = a[k] / d;
#pragma unroll
            for (int j = k + 1; j < CR; ++j) {
                if (j >= c) continue;
                const double pj = tm.wave_bcast_f64(a[j], k);
                if (below) a[j] = a[j] - a[k] * pj;
            }
        }
        // the factors, padded with the identity
        if (mine) {
#pragma unroll
            for (int j = 0; j < CR; ++j) Ap[ln + CR * j] = j < c ? a[j] : 0.0;
        }
        for (int i = c + ln; i < CR; i += tm.wsize())
            for (int j = 0; j < CR; ++j) Ap[i + CR * j] = i == j ? 1.0 : 0.0;
        if (ln == 0) {
            int32_t *perm = ipiv + c;
            for (int i = 0; i < c; ++i) perm[i] = i;
            for (int k = 0; k < c; ++k) {
                const int pk = ipiv[k], t = perm[k];
                perm[k] = perm[pk];
                perm[pk] = t;
            }
        }
        tm.wave_sync();
        const int32_t *perm = ipiv + c;
        for (int q0 = 0; q0 < r; q0 += tm.wsize()) {
            const int q = q0 + ln;
            if (q < r) {
                double x[CR];
#pragma unroll
                for (int i = 0; i < CR; ++i) {
                    const double v = Pi1[q + (int64_t)r * perm[i < c ? i : 0]];
                    x[i] = i < c ? v : 0.0;
                }
#pragma unroll
                for (int i = 0; i < CR; ++i) {
                    double s = x[i];
#pragma unroll
                    for (int t = 0; t < i; ++t) s = s - Ap[i + CR * t] * x[t];
                    x[i] = s;
                }
#pragma unroll
                for (int i = CR - 1; i >= 0; --i) {
                    double s = x[i];
#pragma unroll
                    for (int t = i + 1; t < CR; ++t) s = s - Ap[i + CR * t] * x[t];
                    x[i] = s / Ap[i + CR * i];
                }
#pragma unroll
                for (int i = 0; i < CR; ++i)
                    if (i < c) Pi1[q + (int64_t)r * i] = x[i];
            }
        }
        return singular;
    }

    // fillsitetensors! (globalsearch.jl:97-104) → setsitetensor!(tci, f, b)
    // (tensorci2.jl:367-394).  When no global search follows (nsearch == 0)
    // the tensors are never read before sweep1site! rebuilds all of them, so
    // only max|Pi1| (which feeds maxsamplevalue) and the shape check are
    // computed.  Otherwise each
    // site's Pi1 is evaluated into its tensor slot, and the independent solves
    // T_b = Pi1·P⁻¹ run in groups of sites whose Pᵀ fit in LDS together, one
    // wave per site.  Errors keep the reference's order: a singular P at a
    // site before the first non-square one wins.
    ENG int fillsitetensors(bool need_tensors)
    {
        int nsq = -1;   // first site whose P is not square
        for (int b = 0; b < L; ++b) {
            const int nI = setn(SET_I, b), d = D.localdims[b], nJ = setn(SET_J, b);
            const int m = kron_rows(b);
            if (m > D.mrcap) return fail(ST_CAPACITY);
            load_set(cols, SET_J, b, jlen(b));
            tm.sync();
            if (!need_tensors) {
                eval(rows, b + 1, m, cols, nJ, nullptr, 0, false, true);
                if (b + 1 < L && setn(SET_I, b + 1) != nJ) { nsq = b; break; }
                continue;
            }
            if ((int64_t)m * nJ > D.tstride) return fail(ST_CAPACITY);
            eval(rows, b + 1, m, cols, nJ, Tslot(b), m, false, true);
            setTdims(b, nI, d, nJ);
            if (b + 1 < L && setn(SET_I, b + 1) != nJ) { nsq = b; break; }
        }
        if (!need_tensors) {
            // the reference's "not square" check still holds; its LU of each
            // P (SingularException) is not repeated when no tensor is needed
            if (nsq >= 0) {
                if (tm.tid() == 0) D.ints[INT_ERRSITE] = nsq + 1;
                return fail(ST_NONSQUARE);
            }
            return ST_OK;
        }
        const int nsolve = nsq >= 0 ? nsq : L - 1;
        // per-wave row scratch for the substitutions (null: in place in HBM)
        int cmax = 1;
        for (int b = 0; b < nsolve; ++b) cmax = setn(SET_I, b + 1) > cmax ? setn(SET_I, b + 1) : cmax;
        int64_t off0 = 0;
        // per-wave padded factor blocks for the register path (bond
        // dimensions up to 16), else row scratch for the substitutions
        // (null: in place in HBM)
        double *rsl = cmax <= 16 && cmax <= tm.wsize() ? lds_take((int64_t)tm.nwaves() * 256, off0) : nullptr;
        double *xsl = rsl ? nullptr : lds_take((int64_t)tm.nwaves() * tm.wsize() * cmax, off0);
        for (int g0 = 0; g0 < nsolve;) {
            // a group: consecutive sites whose Pᵀ and pivots fit in LDS
            // together; a site that does not fit alone uses the HBM buffers
            int64_t off = off0;
            int g1 = g0;
            while (g1 < nsolve) {
                const int64_t np = setn(SET_I, g1 + 1);
                if (!lds_take(np * np + np, off)) break;
                ++g1;
            }
            const bool onchip = g1 > g0;
            if (!onchip) g1 = g0 + 1;
            off = off0;
            for (int b = g0; b < g1; ++b) {
                const int np = setn(SET_I, b + 1);
                double *At = onchip ? lds_take((int64_t)np * np + np, off) : D.pmat;
                load_set(cols, SET_J, b, jlen(b));
                tm.sync();
                // Pᵀ[jj, ii] = f(Iset[b+1][ii] ⊕ Jset[b][jj])
                eval(setp(SET_I, b + 1), b + 1, np, cols, np, At, np, true, false);
            }
            tm.sync();
            const uint64_t ts = tm.clock();
            int singular_site = 0;   // lowest singular site + 1 of this wave
            off = off0;
            for (int b = g0; b < g1; ++b) {
                const int m = setn(SET_I, b) * D.localdims[b], np = setn(SET_I, b + 1);
                double *At = onchip ? lds_take((int64_t)np * np + np, off) : D.pmat;
                int32_t *ipv = onchip ? (int32_t *)(At + (int64_t)np * np) : D.ipiv;
                if ((b - g0) % tm.nwaves() != tm.wave()) continue;
                double *T = Tslot(b);
                if (rsl) {
                    double *Ap = rsl + 256 * tm.wave();
                    const int sg = np <= 8 ? site_solve_reg<8>(At, np, ipv, T, m, Ap)
                                           : site_solve_reg<16>(At, np, ipv, T, m, Ap);
                    if (sg && !singular_site) singular_site = b + 1;
                    continue;
                }
                if (lu_wave(At, np, ipv, onchip) && !singular_site) singular_site = b + 1;
                double *xw = xsl ? xsl + (int64_t)tm.wave() * tm.wsize() * cmax + tm.lane() : nullptr;
                for (int q = tm.lane(); q < m; q += tm.wsize())
                    subst_row(T, m, q, At, np, ipv, xw, tm.wsize());
            }
            const int any_sing = tm.any(tm.lane() == 0 && singular_site != 0);
            if (tm.tid() == 0) S.tph[PH_SOLVE] += tm.clock() - ts;
            if (any_sing) {
                // the lowest singular site, as the site-ordered loop raises it
                if (tm.tid() == 0) S.cnt = 0x7fffffff;
                tm.sync();
                for (int w = 0; w < tm.nwaves(); ++w) {
                    if (tm.wave() == w && tm.lane() == 0 && singular_site && singular_site < S.cnt)
                        S.cnt = singular_site;
                    tm.sync();
                }
                if (tm.tid() == 0) D.ints[INT_ERRSITE] = S.cnt;
                return fail(ST_SINGULAR);
            }
            g0 = g1;
        }
        if (nsq >= 0) {
            if (tm.tid() == 0) D.ints[INT_ERRSITE] = nsq + 1;
            return fail(ST_NONSQUARE);
        }
        return ST_OK;
    }

    // ---------------------------------------------------------- sweep2site!
    // tensorci2.jl:855-916 with niter = 2
    ENG int sweep2site(double abstol)
    {
        for (int iter = 1; iter <= 2; ++iter) {
            const int hist = D.ints[INT_HIST];
            const bool extras = !D.strictlynested && hist != 0;
            const int ek_i = hist == 1 ? SET_HA_I : SET_HB_I;
            const int ek_j = hist == 1 ? SET_HA_J : SET_HB_J;
            const int nh_i = hist == 1 ? SET_HB_I : SET_HA_I;   // new history slot
            const int nh_j = hist == 1 ? SET_HB_J : SET_HA_J;
            // push!(Iset_history, deepcopy(Iset)) — only history[end] is ever read
            for (int b = 0; b < L; ++b) {
                const int nI = setn(SET_I, b), nJ = setn(SET_J, b);
                const uint8_t *si = setp(SET_I, b), *sj = setp(SET_J, b);
                uint8_t *di = setp(nh_i, b), *dj = setp(nh_j, b);
                for (int64_t q = tm.tid(); q < (int64_t)nI * L; q += tm.nt()) di[q] = si[q];
                for (int64_t q = tm.tid(); q < (int64_t)nJ * L; q += tm.nt()) dj[q] = sj[q];
            }
            tm.sync();
            if (tm.tid() == 0) {
                for (int b = 0; b < L; ++b) {
                    setn(nh_i, b) = setn(SET_I, b);
                    setn(nh_j, b) = setn(SET_J, b);
                }
                D.ints[INT_HIST] = hist == 1 ? 2 : 1;
                S.npe = 0;   // flushpivoterror!
            }
            tm.sync();
            const bool fwd = D.sweepstrategy == 1 || (D.sweepstrategy == 0 && (iter & 1));
            for (int s = 0; s < L - 1; ++s) {
                const int b = fwd ? s : L - 2 - s;
                // (:rook is compiled into the generic kernel only, FK == 0)
                int st = (FK == 0 && D.pivotsearch)
                             ? updatepivots_rook(b, fwd, abstol, extras ? ek_i : -1, extras ? ek_j : -1)
                             : updatepivots(b, fwd, abstol, extras ? ek_i : -1, extras ? ek_j : -1);
                if (st != ST_OK) {
                    if (tm.tid() == 0) D.ints[INT_ERRSITE] = b + 1;
                    return st;
                }
            }
        }
        const uint64_t t0 = tm.clock();
        // optimize! always calls sweep2site! with fillsitetensors=true
        // (tensorci2.jl:777-787), so its errors (a singular or non-square
        // pivot matrix) are raised even when no global search reads the
        // tensors
        int st = fillsitetensors(true);
        if (tm.tid() == 0) S.tph[PH_FILL] += tm.clock() - t0;
        return st;
    }

    // TT values of the candidates (abstracttensortrain.jl:131: the running
    // row vector times each site's slice, each product summed in ascending
    // bond order) with each thread's (at most kChainCand) running vectors in
    // registers, CR >= every bond dimension.  Each site tensor is staged in
    // LDS zero-padded to CR×d×CR, so the products run over a fixed CR×CR
    // block: a padded term adds 0·0 to a sum that is never -0, which leaves
    // it unchanged, and the entries past the bond dimension are reset to 0
    // after each site.  errs[c] = |f(x_c) - TT(x_c)|.  Returns false (and
    // does nothing) when the padded tensor does not fit in LDS.
    template <int CR>
    ENG_NOINLINE bool chain_reg(const uint8_t *cand, int ncand, double *errs, int64_t loff)
    {
        int dmax = 1;
        for (int t = 0; t < L; ++t) dmax = D.localdims[t] > dmax ? D.localdims[t] : dmax;
        int64_t off = loff;
        double *Tp = lds_take((int64_t)CR * dmax * CR, off);
        if (!Tp) return false;
        double v[kChainCand][CR];
#pragma unroll
        for (int q = 0; q < kChainCand; ++q)
#pragma unroll
            for (int i = 0; i < CR; ++i) v[q][i] = i == 0 ? 1.0 : 0.0;
        for (int t = 0; t < L; ++t) {
            const int d0 = tm.uniform((int)D.Tdims[3 * t]), d = tm.uniform((int)D.Tdims[3 * t + 1]);
            const int d2 = tm.uniform((int)D.Tdims[3 * t + 2]);
            const double *T = D.T + D.tstride * t;
            tm.sync();
            for (int e = tm.tid(); e < CR * d * CR; e += tm.nt()) {
                const int i = e % CR, r = e / CR, sl = r % d, j = r / d;
                Tp[e] = (i < d0 && j < d2) ? T[i + (int64_t)d0 * (sl + (int64_t)d * j)] : 0.0;
            }
            tm.sync();
#pragma unroll
            for (int q = 0; q < kChainCand; ++q) {
                const int c = tm.tid() + q * tm.nt();
                if (c < ncand) {
                    const double *Tc = Tp + CR * (cand[(int64_t)c * L + t] - 1);
                    double w[CR];
#pragma unroll
                    for (int j = 0; j < CR; ++j) {
                        const double *Tj = Tc + (int64_t)CR * d * j;
                        double acc = 0.0;
#pragma unroll
                        for (int i = 0; i < CR; ++i) acc = acc + v[q][i] * Tj[i];
                        w[j] = j < d2 ? acc : 0.0;
                    }
#pragma unroll
                    for (int j = 0; j < CR; ++j) v[q][j] = w[j];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kChainCand; ++q) {
            const int c = tm.tid() + q * tm.nt();
            if (c < ncand) {
                tci_point pt;
                pt.a = cand + (int64_t)c * L; pt.b = pt.a; pt.na = L; pt.L = L;
                const double fv = tci_functor_eval(FK ? FK : D.fkind, D.params, D.iparams, &pt);
                errs[c] = fabs(fv - v[q][0]);
            }
        }
        tm.sync();
        return true;
    }

    // ------------------------------------------------------ global search
    // DefaultGlobalPivotFinder (globalpivotfinder.jl:143-195), project PRNG.
    static constexpr int kChainReg = 16, kChainCand = 2;
    ENG int globalsearch(double abstol)
    {
        const int ns = D.nsearchglobalpivot;
        const int sumd = D.sumd;
        const int ncand = ns * sumd;
        uint8_t *found = D.gpts + (int64_t)ns * L;      // ns × L
        // the start points (ns × L), the candidates (ns·Σd × L) and their
        // errors (ncand, then ns best errors and ns indices) in LDS when
        // they fit, else in the arena
        int64_t loff = 0;
        uint8_t *pts = (uint8_t *)lds_take(((int64_t)ns * L + 7) / 8, loff);
        uint8_t *cand = (uint8_t *)lds_take(((int64_t)ncand * L + 7) / 8, loff);
        double *errs = lds_take((int64_t)ncand + 2 * ns, loff);
        if (!pts) pts = D.gpts;
        if (!cand) cand = found + (int64_t)ns * L;
        if (!errs) errs = D.work;
        if (tm.tid() == 0) {
            for (int q = 0; q < ns; ++q)
                for (int p = 0; p < L; ++p) {
                    // (a 32-bit draw, so a 32-bit remainder)
                    const uint32_t r = (uint32_t)(eng_rng_next(S.rng) >> 32);
                    pts[(int64_t)q * L + p] = (uint8_t)(1 + (int)(r % (uint32_t)D.localdims[p]));
                }
        }
        tm.sync();
        // candidate points: the point with site p set to v, p-major then v
        for (int c = tm.tid(); c < ncand; c += tm.nt()) {
            const int q = c / sumd;
            int r = c - q * sumd, p = 0;
            while (r >= D.localdims[p]) { r -= D.localdims[p]; ++p; }
            uint8_t *x = cand + (int64_t)c * L;
            for (int t = 0; t < L; ++t) x[t] = pts[(int64_t)q * L + t];
            x[p] = (uint8_t)(r + 1);
        }
        int chi = 1;
        for (int t = 0; t <= L; ++t) {
            const int64_t w = t < L ? D.Tdims[3 * t] : 1;
            chi = w > chi ? (int)w : chi;
        }
        int64_t tmax = 0;
        for (int t = 0; t < L; ++t) {
            const int64_t sz = D.Tdims[3 * t] * D.Tdims[3 * t + 1] * D.Tdims[3 * t + 2];
            tmax = sz > tmax ? sz : tmax;
        }
        bool done = false;
        if (chi <= kChainReg && ncand <= kChainCand * tm.nt()) {
            // bond dimensions up to 16: each thread carries its (at most two)
            // candidates' running vectors in registers
            if (chi <= 8) done = chain_reg<8>(cand, ncand, errs, loff);
            else if (chi <= 12) done = chain_reg<12>(cand, ncand, errs, loff);
            else done = chain_reg<16>(cand, ncand, errs, loff);
        }
        if (!done) {
            // TT values of all candidates, site by site (abstracttensortrain.jl:131,
            // each product summed in ascending bond order), chunks of candidates;
            // keep the running vectors in LDS, and each site tensor too when it fits
            int64_t off = loff;
            double *Ts = lds_take(tmax, off);
            const int64_t room = tm.lpi_cap > off ? tm.lpi_cap - off : 0;
            int chunk = (int)(room / (2 * (int64_t)chi));
            double *V = nullptr, *W = nullptr;
            if (chunk >= 1) {
                if (chunk > ncand) chunk = ncand;
                V = lds_take((int64_t)chunk * chi, off);
                W = lds_take((int64_t)chunk * chi, off);
            }
            if (!V || !W) {
                chunk = ncand;
                V = D.work + ncand;
                W = V + (int64_t)ncand * chi;
            }
            tm.sync();
            for (int c0 = 0; c0 < ncand; c0 += chunk) {
                const int cn = ncand - c0 < chunk ? ncand - c0 : chunk;
                for (int c = tm.tid(); c < cn; c += tm.nt()) V[(int64_t)c * chi] = 1.0;
                tm.sync();
                for (int t = 0; t < L; ++t) {
                    const int64_t d0 = D.Tdims[3 * t], d = D.Tdims[3 * t + 1], d2 = D.Tdims[3 * t + 2];
                    const double *T = D.T + D.tstride * t;
                    if (Ts) {
                        for (int64_t q = tm.tid(); q < d0 * d * d2; q += tm.nt()) Ts[q] = T[q];
                        tm.sync();
                        T = Ts;
                    }
                    for (int64_t o = tm.tid(); o < (int64_t)cn * d2; o += tm.nt()) {
                        const int64_t c = o / d2, j = o - c * d2;
                        const int64_t s = cand[(c0 + c) * L + t] - 1;
                        const double *Tc = T + d0 * (s + d * j);
                        const double *v = V + c * chi;
                        double acc = 0.0;
                        for (int64_t i = 0; i < d0; ++i) acc = acc + v[i] * Tc[i];
                        W[c * chi + j] = acc;
                    }
                    tm.sync();
                    double *tmp = V; V = W; W = tmp;
                }
                for (int c = tm.tid(); c < cn; c += tm.nt()) {
                    tci_point pt;
                    pt.a = cand + (int64_t)(c0 + c) * L; pt.b = pt.a; pt.na = L; pt.L = L;
                    const double fv = tci_functor_eval(FK ? FK : D.fkind, D.params, D.iparams, &pt);
                    errs[c0 + c] = fabs(fv - V[(int64_t)c * chi]);
                }
                tm.sync();
            }
        }
        // per point, the first candidate with the largest error (a strict
        // maximum over errors > 0, as in the reference's scan), one wave per
        // point; then thread 0 keeps the points above the threshold in order
        const double thr = abstol * D.tolmargin;
        double *bv = errs + ncand;                    // ns best errors
        int64_t *bix = (int64_t *)(bv + ns);          // ns candidate indices (-1: none)
        for (int q = tm.wave(); q < ns; q += tm.nwaves()) {
            double lb = 0.0;
            uint32_t li = 0xffffffffu;
            for (int r = tm.lane(); r < sumd; r += tm.wsize()) {
                const double e = errs[(int64_t)q * sumd + r];
                if (e > lb) { lb = e; li = (uint32_t)r; }
            }
            const double g = tm.wave_max(lb);
            const uint32_t ri = tm.wave_min((lb == g && g > 0.0) ? li : 0xffffffffu);
            if (tm.lane() == 0) {
                bv[q] = g;
                bix[q] = ri == 0xffffffffu ? -1 : (int64_t)q * sumd + ri;
            }
        }
        tm.sync();
        if (tm.tid() == 0) {
            int nf = 0;
            for (int q = 0; q < ns; ++q) {
                if (bv[q] > thr) {
                    const uint8_t *src = bix[q] >= 0 ? cand + bix[q] * L : pts + (int64_t)q * L;
                    for (int t = 0; t < L; ++t) found[(int64_t)nf * L + t] = src[t];
                    ++nf;
                }
            }
            if (nf > D.maxnglobalpivot) nf = D.maxnglobalpivot;
            S.nf = nf;
        }
        tm.sync();
        const int nf = S.nf;
        return addglobalpivots(found, nf) == ST_OK ? nf : -1;
    }

    // convergencecriterion (tensorci2.jl:609-628); thread 0
    ENG bool converged(int n, double tol) const
    {
        const int nch = D.ncheckhistory;
        if (n < nch) return false;
        bool all_err = true, all_ng = true, all_max = true;
        int64_t minrank = D.ranks[n - nch];
        for (int i = n - nch; i < n; ++i) {
            if (!(D.errors[i] < tol)) all_err = false;
            if (D.nglob[i] != 0) all_ng = false;
            if (D.ranks[i] < minrank) minrank = D.ranks[i];
            if (!(D.ranks[i] >= D.maxbonddim)) all_max = false;
        }
        if (!D.checkconvglobalpivot) all_ng = true;
        return (all_err && all_ng && minrank == D.ranks[n - 1]) || all_max;
    }

    // ------------------------------------------------ final sweep1site!
    // tensorci2.jl:402-461, forward, updatetensors = true
    // (inlined, and finish kept out of line: the kernel body measured
    // fastest with these two call decisions)
    ENG_INLINE int sweep1site_forward(double abstol)
    {
        if (tm.tid() == 0) S.npe = 0;
        for (int b = 0; b < L - 1; ++b) {
            const int nI = setn(SET_I, b), d = D.localdims[b], nJ = setn(SET_J, b);
            const int m = kron_rows(b);
            if (m > D.mrcap || nJ > D.mrcap) return fail(ST_CAPACITY);
            load_set(cols, SET_J, b, jlen(b));
            tm.sync();
            double *A = pibuf(m, nJ);
            int64_t off = perm_area(A == tm.lpi ? (((int64_t)m * nJ + 1) & ~(int64_t)1) : 0, m, nJ);
            eval(rows, b + 1, m, cols, nJ, A, m, false, true);
            double err;
            int p = rrlu(A, m, nJ, D.maxbonddim, 1e-14, abstol, true, &err, true, Tslot(b));
            if (p < 0) {
                if (tm.tid() == 0) D.ints[INT_ERRSITE] = b + 1;
                return fail(-p);
            }
            if (p > D.setcap || (int64_t)m * p > D.tstride) return fail(ST_CAPACITY);
            select_into(SET_I, b + 1, rows, rperm, p, b + 1);
            select_into(SET_J, b, cols, cperm, p, jlen(b));
            // left(luci) = colstimespivotinv (matrixluci.jl:48-57): rows at
            // positions < p are unit vectors, the rest solve x·L11 = L[i,1:p]
            // (L11 has a unit diagonal: s / 1.0 == s); solved rows go to LDS
            // when they fit, then are scattered to T in original row order
            double *T = Tslot(b);
            double *X = lds_take((int64_t)m * p, off);
            for (int i = p + tm.tid(); i < m; i += tm.nt()) {
                double *row = X ? X + i : T + rperm[i];
                for (int j = p - 1; j >= 0; --j) {
                    double s = A[i + (int64_t)m * j];
                    for (int t = j + 1; t < p; ++t) s = s - row[(int64_t)m * t] * A[t + (int64_t)m * j];
                    row[(int64_t)m * j] = s;
                }
            }
            if (X) tm.sync();
            for (int64_t q = tm.tid(); q < (int64_t)m * p; q += tm.nt()) {
                const int64_t j = q / m, i = q - j * m;
                if (i < p) T[rperm[i] + (int64_t)m * j] = (i == j) ? 1.0 : 0.0;
                else if (X) T[rperm[i] + (int64_t)m * j] = X[q];
            }
            setTdims(b, nI, d, p);
            tm.sync();
            int nan = 0;
            for (int64_t q = tm.tid(); q < (int64_t)m * p; q += tm.nt()) {
                double v = T[q];
                if (v != v) nan = 1;
            }
            if (tm.any(nan)) {
                if (tm.tid() == 0) D.ints[INT_ERRSITE] = b + 1;
                return fail(ST_NAN_TENSOR);
            }
            updateerrors(b, p, err);
            tm.sync();
        }
        // last site: T_L = filltensor(Iset[L], Jset[L], Val(1)) (tensorci2.jl:449-459)
        const int b = L - 1;
        const int nI = setn(SET_I, b), d = D.localdims[b], nJ = setn(SET_J, b);
        const int m = kron_rows(b);
        if (m > D.mrcap) return fail(ST_CAPACITY);
        load_set(cols, SET_J, b, jlen(b));
        tm.sync();
        if ((int64_t)m * nJ > D.tstride) return fail(ST_CAPACITY);
        eval(rows, b + 1, m, cols, nJ, Tslot(b), m, false, false);
        setTdims(b, nI, d, nJ);
        tm.sync();
        return ST_OK;
    }

    ENG int64_t rank() const
    {
        int64_t r = 0;
        for (int b = 1; b < L; ++b) if (setn(SET_I, b) > r) r = setn(SET_I, b);
        return r;
    }

    // ------------------------------------------------ crossinterpolate2
    // tensorci2.jl:943-953 → TensorCI2{T}(f, localdims, initialpivots)
    // (42-53) → optimize! (700-850)
    ENG int run()
    {
        if (tm.tid() == 0) {
            for (int i = 0; i < 8; ++i) S.tph[i] = 0;
            S.tph[PH_TOTAL] = tm.clock();
            S.status = ST_OK;
            for (int i = 0; i < 6; ++i) S.ws[i] = 0.0;
            S.cache_lookups = 0;
            S.cache_hits = 0;
            for (int q = 0; q < NSETKIND * L; ++q) setcnt[q] = 0;
            D.ints[INT_NITER] = 0;
            D.ints[INT_NPE] = 0;
            S.npe = 0;
            D.ints[INT_HIST] = 0;
            D.ints[INT_ERRSITE] = 0;
            for (int b = 0; b < L - 1; ++b) D.bonderrors[b] = 0.0;
            for (int b = 0; b < L; ++b) { D.Tdims[3 * b] = 0; D.Tdims[3 * b + 1] = 0; D.Tdims[3 * b + 2] = 0; }
        }
        tm.sync();
        if (cache_on()) {
            // an empty memo table (a new CachedFunction per solve)
            const int64_t ne = (int64_t)1 << D.cache_log2;
            for (int64_t q = tm.tid(); q < ne; q += tm.nt()) D.cache_state[q] = CACHE_EMPTY;
            tm.sync();
        }
        build_lut();
        int st = D.initsets ? load_state() : start_from_pivots();
        if (st != ST_OK) return finish(st);
        // optimize! argument checks (730-754)
        if (D.nsearchglobalpivot > 0 && D.nsearchglobalpivot < D.maxnglobalpivot) return finish(ST_GENERIC);
        const double tol = D.tolerance;
        if (D.maxbonddim == INT64_MAX && tol <= 0) return finish(ST_ARG);
        if (tm.tid() == 0) eng_rng_seed(S.rng, D.seed);
        tm.sync();
        return finish(optimize(tol));
    }

    // TensorCI2{T}(f, localdims, initialpivots) (tensorci2.jl:42-53)
    ENG_INLINE int start_from_pivots()
    {
        int st = addglobalpivots(D.initpivots, D.npivinit);
        if (st != ST_OK) return st;
        // maxsamplevalue = maximum(abs, f(x) for x in initialpivots) (49-50)
        if (tm.tid() == 0) {
            double mx = 0.0;
            for (int q = 0; q < D.npivinit; ++q) {
                tci_point pt;
                pt.a = D.initpivots + (int64_t)q * L; pt.b = pt.a; pt.na = L; pt.L = L;
                double v = fabs(tci_functor_eval(FK ? FK : D.fkind, D.params, D.iparams, &pt));
                mx = q == 0 ? v : jl_max(mx, v);
            }
            S.maxsample = mx;
            if (!(fabs(mx) > 0.0)) S.status = ST_ZERO_MAXSAMPLE;
        }
        tm.sync();
        return S.status;
    }

    // TensorCI2{T}(f, localdims, Iset, Jset) (tensorci2.jl:58-72), or the
    // state of an existing TensorCI2 that optimize! continues from: the sets
    // (and the last history entry) from the input image; maxsamplevalue as
    // given, else the maximum of |f| over every Iset[b] ⊕ σ ⊕ Jset[b], the
    // points reconstractglobalpivotsfromijset enumerates (171-190)
    ENG_INLINE int load_state()
    {
        const int64_t per = (int64_t)D.setcap * L;
        const int32_t *cn = (const int32_t *)(D.initsets + (int64_t)NSETKIND * L * per);
        for (int q = tm.tid(); q < NSETKIND * L; q += tm.nt()) setcnt[q] = cn[q];
        for (int64_t x = tm.tid(); x < (int64_t)NSETKIND * L * per; x += tm.nt()) sets[x] = D.initsets[x];
        if (tm.tid() == 0) {
            D.ints[INT_HIST] = D.init_hist ? 1 : 0;
            S.maxsample = D.init_maxsample;
        }
        tm.sync();
        if (D.init_maxsample > 0.0) return ST_OK;
        for (int b = 0; b < L; ++b) {
            const int m = kron_rows(b);
            if (m > D.mrcap) return fail(ST_CAPACITY);
            load_set(cols, SET_J, b, jlen(b));
            tm.sync();
            eval(rows, b + 1, m, cols, setn(SET_J, b), nullptr, 0, false, true);
        }
        if (tm.tid() == 0 && !(fabs(S.maxsample) > 0.0)) S.status = ST_ZERO_MAXSAMPLE;
        tm.sync();
        return S.status;
    }

    // optimize! (tensorci2.jl:700-850) from the current state: the 2-site
    // sweeps, the global search and the convergence test, then the final
    // sweep1site! and _sanitycheck
    ENG_INLINE int optimize(double tol)
    {
        int st = ST_OK;
        int niter = 0;
        for (int iter = 1; iter <= D.maxiter; ++iter) {
            const double norm = D.normalizeerror ? S.maxsample : 1.0;
            const double abstol = tol * norm;
            st = sweep2site(abstol);
            if (st != ST_OK) return st;
            double perr = D.bonderrors[0];
            for (int b = 1; b < L - 1; ++b) perr = jl_max(perr, D.bonderrors[b]);
            int ng = 0;
            if (D.nsearchglobalpivot > 0) {
                const uint64_t t0 = tm.clock();
                ng = globalsearch(abstol);
                if (tm.tid() == 0) S.tph[PH_GSEARCH] += tm.clock() - t0;
                if (ng < 0) return S.status;
            }
            if (tm.tid() == 0) {
                if (niter < D.maxiter_cap) {
                    D.ranks[niter] = rank();
                    D.errors[niter] = perr;
                    D.nglob[niter] = ng;
                }
                D.ints[INT_NITER] = niter + 1;
            }
            ++niter;
            tm.sync();
            bool conv = false;
            if (tm.tid() == 0) S.cnt = converged(niter, abstol) ? 1 : 0;
            tm.sync();
            conv = S.cnt != 0;
            tm.sync();
            if (conv) break;
        }
        const double norm = D.normalizeerror ? S.maxsample : 1.0;
        if (!(D.optflags & OPT_NO_FINAL_SWEEP)) {
            const uint64_t t1 = tm.clock();
            st = sweep1site_forward(tol * norm);
            if (tm.tid() == 0) S.tph[PH_SWEEP1] += tm.clock() - t1;
            if (st != ST_OK) return st;
            // _sanitycheck (globalsearch.jl:106-111)
            for (int b = 0; b < L - 1; ++b)
                if (setn(SET_I, b + 1) != setn(SET_J, b)) {
                    if (tm.tid() == 0) D.ints[INT_ERRSITE] = b + 1;
                    return ST_NONSQUARE;
                }
        }
        if (tm.tid() == 0)
            for (int i = 0; i < niter; ++i) D.errors[i] = D.errors[i] / norm;
        return ST_OK;
    }

    ENG_NOINLINE int finish(int st)
    {
        tm.sync();
        if (sets != D.sets) {   // sets held in LDS: write them to the solve's arena
            for (int q = tm.tid(); q < NSETKIND * L; q += tm.nt()) D.setn[q] = setcnt[q];
            const int64_t per = (int64_t)D.setcap * L;
            for (int64_t x = tm.tid(); x < (int64_t)NSETKIND * L * per; x += tm.nt()) {
                const int64_t q = x / per;
                if ((x - q * per) < (int64_t)setcnt[q] * L) D.sets[x] = sets[x];
            }
            tm.sync();
        }
        if (piverrs() != D.pivoterrors)
            for (int i = tm.tid(); i < S.npe; i += tm.nt()) D.pivoterrors[i] = S.pe[i];
        if (tm.tid() == 0) {
            D.ints[INT_NPE] = S.npe;
            D.ints[INT_STATUS] = st;
            S.tph[PH_TOTAL] = tm.clock() - S.tph[PH_TOTAL];
            for (int i = 0; i < 8; ++i) D.scal[SCAL_TPH + i] = (double)S.tph[i];
            D.scal[SCAL_MAXSAMPLE] = S.maxsample;
            for (int i = 0; i < 6; ++i) D.scal[SCAL_WS_PI + i] = S.ws[i];
            D.scal[SCAL_CACHE_LOOKUPS] = (double)S.cache_lookups;
            D.scal[SCAL_CACHE_HITS] = (double)S.cache_hits;
            D.ints[INT_DONE] = 1;
        }
        tm.sync();
        return st;
    }
};

}  // namespace tci
