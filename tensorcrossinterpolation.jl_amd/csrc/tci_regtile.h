// tci_regtile.h — full-pivot rrLU with the whole matrix in registers.
//
// Layout.  The team is NW waves of WS lanes.  Lane l of wave w keeps
// A[r, c] for r = l + WS·i (i < RA) and c = w + NW·j (j < CB), so every wave
// holds complete columns and every row is present in every wave.  A matrix
// of up to (WS·RA)×(NW·CB) is factored without touching memory.
//
// Rows and columns are never moved.  Each thread tracks the current position
// of its rows (rp); the positions of a wave's CB register columns are spread
// over its lanes (lane l holds that of register column l, in cpl), so the
// swap of step k (positions k ↔ pivot) costs a few vector operations.  The argmax key
// of an entry is its current position, colpos·256 + rowpos, so the smallest
// key among equal abs2 values is the first maximum of the reference's
// column-major scan of the physically swapped matrix (matrixlu.jl:1-32).
// Every entry sees the operation sequence of addpivot! (matrixlu.jl:114-136):
// a true division of the pivot column (or row), then a - x*y with two
// roundings.  Rows and columns that have been pivoted get NaN multipliers, so
// their entries turn into NaN and drop out of every later search (a NaN is
// never the maximum, as in submatrixargmax); a dead column's entries being
// NaN, its staged multiplier is NaN without any per-column test.
//
// One pivot step costs two team barriers:
//   A  the team maximum g of |entry| over the trailing block; the maximum
//      abs2 is then G = g·g (squaring is monotone).
//   B  the team minimum key among entries with x·x == G.  Before B each wave
//      that holds a candidate writes the column of its best candidate to its
//      LDS slot, so after B the winning column is already staged.
// The pivot row lies in every wave (one lane holds it) and is broadcast
// through a wave-private LDS slot without a team barrier.
#pragma once

#include "tci_common.h"

namespace tci {

template <class Team, int RA, int CB>
struct TileLU {
    static constexpr int WS = Team::WS, NW = Team::NW;
    static constexpr int MR = WS * RA, MC = NW * CB;   // capacity: MR rows, MC columns
    static_assert(MR <= 256, "row positions must fit the 8-bit part of the key");
    static constexpr int CL = (CB + WS - 1) / WS;   // column-position slots per lane
    double a[RA][CB];
    int rp[RA];     // current position of my rows (large for rows ≥ m)
    int cpl[CL];    // slot c: position of register column lane + WS·c (large if none / ≥ n)
    int jw;         // register column this wave staged in its last search
    int m, n;

    // position of register column j of this wave (wave-uniform)
    ENG_INLINE int colpos(Team &tm, int j) const { return tm.wave_bcast(cpl[j / WS], j % WS); }

    ENG_INLINE static bool fits(int m_, int n_) { return m_ <= MR && n_ <= MC; }

    ENG_INLINE void setup(const Team &tm, int m_, int n_)
    {
        m = m_;
        n = n_;
#pragma unroll
        for (int i = 0; i < RA; ++i) {
            const int r = tm.lane() + WS * i;
            rp[i] = r < m ? r : 0x3fffffff;
        }
#pragma unroll
        for (int c = 0; c < CL; ++c) {
            const int j = tm.lane() + WS * c;
            const int col = tm.wave() + NW * j;
            cpl[c] = (j < CB && col < n) ? col : 0x3fffffff;
        }
        jw = 0;
    }

    // the trailing block (positions >= p in both directions) as rrlu! leaves
    // it in A after factor returned p: A[rowpos + lda·colpos] = entry
    ENG_INLINE void store_trailing(Team &tm, double *A, int64_t lda, int p) const
    {
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const int c = colpos(tm, j);
            if (c < p || c >= n) continue;
#pragma unroll
            for (int i = 0; i < RA; ++i)
                if (rp[i] >= p && rp[i] < m) A[rp[i] + lda * c] = a[i][j];
        }
    }

    // entries outside the matrix hold NaN
    ENG_INLINE void load(const Team &tm, const double *A, int64_t lda)
    {
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int i = 0; i < RA; ++i) {
                const int r = tm.lane() + WS * i, c = tm.wave() + NW * j;
                a[i][j] = (r < m && c < n) ? A[r + lda * c] : 0.0 / 0.0;
            }
    }

    // f = c / (1 + rs[r] + cs[c]) for TCI_F_LORENTZ_INT, whose value is
    // c / (double)(Σ v² + 1) with an exact integer sum split into the row
    // part rs and the column part cs.  With lut (lut[s] = c / (double)s for
    // every sum that occurs) the division is a lookup with the same bits.
    // Returns this thread's NaN-propagating max|f| (the maxsamplevalue
    // contribution).
    ENG_INLINE double fill_lorentz(const Team &tm, double c, const double *rs, const double *cs,
                                   const double *lut)
    {
        // with the sums in LDS (the table always is) the loads are LDS
        // loads, not generic ones
        if (tm.stage_lds) return fill_lorentz_t<true>(tm, c, rs, cs, lut);
        return fill_lorentz_t<false>(tm, c, rs, cs, lut);
    }

    template <bool LDS>
    ENG_INLINE double fill_lorentz_t(const Team &tm, double c, const double *rs, const double *cs,
                                     const double *lut)
    {
        if (LDS) {
            tm.assume_lds(rs);
            tm.assume_lds(cs);
        }
        // (indices clamped into the arrays so every load is unconditional)
        double rsv[RA], csv[CB];
#pragma unroll
        for (int i = 0; i < RA; ++i) {
            const int r = tm.lane() + WS * i;
            const double v = rs[r < m ? r : 0];
            rsv[i] = r < m ? v : 0.0;
        }
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const int cc = tm.wave() + NW * j;
            const double v = cs[cc < n ? cc : 0];
            csv[j] = cc < n ? v : 0.0;
        }
        double lmax = 0.0;
        if (lut) {
            tm.assume_lds(lut);
            bool any = false;
#pragma unroll
            for (int j = 0; j < CB; ++j)
#pragma unroll
                for (int i = 0; i < RA; ++i) {
                    const bool ok = rp[i] < m && tm.wave() + NW * j < n;
                    const double v = lut[(int)(rsv[i] + csv[j]) + 1];
                    a[i][j] = ok ? v : 0.0 / 0.0;
                    lmax = ok ? fmax(lmax, fabs(v)) : lmax;
                    any = any || ok;
                }
            if (any && c != c) lmax = c;   // every value is NaN: max propagates it
            return lmax;
        }
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int i = 0; i < RA; ++i) {
                const bool ok = rp[i] < m && tm.wave() + NW * j < n;
                const double v = c / (rsv[i] + csv[j] + 1.0);
                a[i][j] = ok ? v : 0.0 / 0.0;
                lmax = ok ? jl_max(lmax, fabs(v)) : lmax;
            }
        return lmax;
    }

    // Barriers A and B of a pivot step (see the header).  lmax is this
    // thread's max |entry| (-1: none).  Returns the winning key (UINT32_MAX
    // when every trailing entry is NaN) and the wave whose LDS slot holds the
    // winning column.
    ENG_INLINE uint32_t search(Team &tm, double lmax, int &wstar)
    {
        const double wm = tm.wave_max(lmax);
        if (tm.lane() == 0) tm.tA(tm.wave()) = wm;
        tm.sync();                                            // A
        double g = tm.tA(0);
#pragma unroll
        for (int q = 1; q < NW; ++q) g = fmax(g, tm.tA(q));
        g = tm.uniform_f64(g);   // the same in every lane: keep it scalar
        uint32_t kw = 0xffffffffu;
        if (g >= 0.0) {
            const double G = g * g;
            // {x >= 0 : x·x == G} is an interval of doubles [lo, g] (rounding
            // keeps squaring monotone): find lo by stepping down from g, so
            // the per-entry test is |v| >= lo.  G == 0 means every entry is 0
            // in magnitude or squares to 0; G subnormal may span many doubles,
            // then the exact test v·v == G is used.
            double lo = 0.0;
            bool exact = false;
            if (G > 0.0) {
                lo = g;
                int t = 0;
                for (; t < 8; ++t) {
                    const double q = eng_as_f64(eng_as_i64(lo) - 1);
                    if (q * q == G) lo = q;
                    else break;
                }
                exact = t == 8;
            }
            const bool mine = exact ? lmax * lmax == G : lmax >= lo;
            if (tm.wave_any(mine)) {
                // column parts of the keys, read by the whole wave
                uint32_t ck[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) ck[j] = (uint32_t)colpos(tm, j) << 8;
                uint32_t kl = 0xffffffffu;
                if (mine) {
                    if (!exact) {
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
#pragma unroll
                            for (int i = 0; i < RA; ++i) {
                                const uint32_t key = ck[j] | (uint32_t)rp[i];
                                kl = eng_umin(kl, fabs(a[i][j]) >= lo ? key : 0xffffffffu);
                            }
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < CB; ++j) {
#pragma unroll
                            for (int i = 0; i < RA; ++i) {
                                const double v = a[i][j];
                                const uint32_t key = ck[j] | (uint32_t)rp[i];
                                kl = eng_umin(kl, v * v == G ? key : 0xffffffffu);
                            }
                        }
                    }
                }
                kw = tm.wave_min(kl);
                // stage the column of this wave's candidate (raw values)
                const int cpos = (int)(kw >> 8);
#pragma unroll
                for (int c = 0; c < CL; ++c) {
                    const int l = tm.wave_first(tm.lane() + WS * c < CB && cpl[c] == cpos);
                    if (l >= 0) jw = l + WS * c;
                }
                // (selected by position, not by index, so the tile stays in
                // registers rather than becoming an indexed array)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    if ((ck[j] >> 8) == (uint32_t)cpos) {
#pragma unroll
                        for (int i = 0; i < RA; ++i) tm.tX(tm.wave(), tm.lane() + WS * i) = a[i][j];
                    }
            }
        }
        if (tm.lane() == 0) tm.tB(tm.wave()) = kw;
        tm.sync();                                            // B
        uint32_t K = tm.uniform_u32(tm.tB(0));
        wstar = 0;
#pragma unroll
        for (int q = 1; q < NW; ++q) {
            const uint32_t v = tm.uniform_u32(tm.tB(q));
            if (v < K) { K = v; wstar = q; }
        }
        return K;
    }

    // Full-pivot rrLU (matrixlu.jl:141-181): same pivots, pivot values, error
    // and NaN verdicts as rrlu_team.  Requires entries outside the matrix to
    // hold NaN.  Outputs: rowperm/colperm (original index at each position,
    // written when non-null), pivval[k], *perr, stats[0..1] (argmax and Schur
    // element counts), with Lorig non-null the L column k by original row
    // (Lorig[orig + m·k]) and with Uorig non-null the U row k by original
    // column (Uorig[k + ldu·orig], the pivot row as addpivot! leaves it:
    // divided by the pivot when !leftorth).  Returns p, or -status for the
    // NaN errors.
    ENG_INLINE int factor(Team &tm, int64_t maxrank, double reltol, double abstol, bool leftorth,
                          double *pivval, double *perr, double *stats, double *Lorig,
                          int32_t *rowperm, int32_t *colperm, double *Uorig = nullptr,
                          int64_t ldu = 0)
    {
        const double NaN = 0.0 / 0.0;
        double lmax = -1.0;
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int i = 0; i < RA; ++i) lmax = fmax(lmax, fabs(a[i][j]));
        int wstar;
        uint32_t K = search(tm, lmax, wstar);
        int64_t mr = maxrank;
        if (m < mr) mr = m;
        if (n < mr) mr = n;
        double maxerr = 0.0, err = NaN;
        int64_t sarg = 0, sschur = 0;   // element counts (integers stay in scalar registers)
        int p = 0, nanx = 0, nany = 0;
        while (p < mr) {
            const int k = p;
            const bool cand = K != 0xffffffffu;
            const int pr = cand ? (int)(K & 255u) : k;
            const int pc = cand ? (int)(K >> 8) : k;
            // original row of position pr: held in every wave by one lane
            int orow = 0;
#pragma unroll
            for (int i = 0; i < RA; ++i) {
                const int l = tm.wave_first(rp[i] == pr);
                if (l >= 0) orow = l + WS * i;
            }
            const double av = cand ? tm.uniform_f64(tm.tX(wstar, orow)) : NaN;
            sarg += (int64_t)(m - k) * (n - k);
            const double e = fabs(av);
            err = e;
            if ((e < reltol * maxerr || e < abstol) && p > 0) break;
            maxerr = jl_max(maxerr, e);
#pragma unroll
            for (int i = 0; i < RA; ++i) rp[i] = rp[i] == k ? pr : (rp[i] == pr ? k : rp[i]);
#pragma unroll
            for (int c = 0; c < CL; ++c) cpl[c] = cpl[c] == k ? pc : (cpl[c] == pc ? k : cpl[c]);
            if (tm.tid() == 0 && pivval) pivval[k] = av;
            if (av != av) { nanx = 1; nany = 1; }
            // x: the staged pivot column, divided by the pivot when leftorth
            double x[RA];
#pragma unroll
            for (int i = 0; i < RA; ++i) {
                const bool live = rp[i] > k && rp[i] < m;
                double v = NaN;
                if (cand && live) {
                    v = tm.tX(wstar, tm.lane() + WS * i);
                    if (leftorth) v = v / av;
                }
                if (live && v != v) nanx = 1;
                if (Lorig && live && tm.wave() == 0) Lorig[(tm.lane() + WS * i) + (int64_t)m * k] = v;
                x[i] = live ? v : NaN;
            }
            // y: the pivot row, broadcast from its holder lane in this wave
            // through the wave's LDS slot.  Dead columns hold NaN already; the
            // pivot column (staged by wave wstar as its register column jw)
            // dies now.
            if (cand) {
                const int lo = orow % WS, io = orow / WS;
                if (tm.lane() == lo) {
#pragma unroll
                    for (int i = 0; i < RA; ++i)
                        if (i == io) {
#pragma unroll
                            for (int j = 0; j < CB; ++j) tm.tY(tm.wave(), j) = a[i][j];
                        }
                    if (tm.wave() == wstar) tm.tY(tm.wave(), jw) = NaN;
                }
                tm.wave_sync_lds();
                if (!leftorth) {
                    for (int j = tm.lane(); j < CB; j += WS) tm.tY(tm.wave(), j) = tm.tY(tm.wave(), j) / av;
                    tm.wave_sync_lds();
                }
            } else {
                for (int j = tm.lane(); j < CB; j += WS) tm.tY(tm.wave(), j) = NaN;
                tm.wave_sync_lds();
            }
            // U row: NaN on a live column (lane l checks register column l)
#pragma unroll
            for (int c = 0; c < CL; ++c) {
                const int j = tm.lane() + WS * c;
                if (j < CB && cpl[c] > k && cpl[c] < n) {
                    const double v = tm.tY(tm.wave(), j);
                    if (v != v) nany = 1;
                    if (Uorig) Uorig[k + ldu * (tm.wave() + NW * j)] = v;
                }
            }
            // trailing update and the next search.  With a full register
            // budget the multipliers are read from LDS up front so their
            // latency overlaps; with half of it (two workgroups per CU) each
            // is read as its column is updated, to keep the tile in registers
            lmax = -1.0;
            if (Team::WPE <= 2 && CB <= 32) {
                double yv[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) yv[j] = tm.tY(tm.wave(), j);
#pragma unroll
                for (int j = 0; j < CB; ++j)
#pragma unroll
                    for (int i = 0; i < RA; ++i) {
                        const double t = x[i] * yv[j];
                        const double u = a[i][j] - t;
                        a[i][j] = u;
                        lmax = fmax(lmax, fabs(u));
                    }
            } else {
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    const double yj = tm.tY(tm.wave(), j);
#pragma unroll
                    for (int i = 0; i < RA; ++i) {
                        const double t = x[i] * yj;
                        const double u = a[i][j] - t;
                        a[i][j] = u;
                        lmax = fmax(lmax, fabs(u));
                    }
                }
            }
            p = k + 1;
            sschur += (int64_t)(m - p) * (n - p);
            K = search(tm, lmax, wstar);
        }
        // position → original index
        if (rowperm && tm.wave() == 0) {
#pragma unroll
            for (int i = 0; i < RA; ++i)
                if (rp[i] < m) rowperm[rp[i]] = tm.lane() + WS * i;
        }
        if (colperm) {
#pragma unroll
            for (int c = 0; c < CL; ++c) {
                const int j = tm.lane() + WS * c;
                if (j < CB && cpl[c] < n) colperm[cpl[c]] = tm.wave() + NW * j;
            }
        }
        const int nanL = tm.any(nanx);
        const int nanU = tm.any(nany);
        stats[0] = (double)sarg;
        stats[1] = (double)sschur;
        if (p >= (m < n ? m : n)) err = 0.0;
        *perr = err;
        if (nanL) return -ST_NAN_L;
        if (nanU) return -ST_NAN_U;
        return p;
    }
};

}  // namespace tci
