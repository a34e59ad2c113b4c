// engine_host.cpp — TEST-ONLY CPU build of the device solve engine.
//
// Compiles tci_engine.h (the code the gfx950 kernel runs) with a one-thread
// "team" so the tests can check the engine's control logic and arithmetic
// against the oracle on the CPU.  The product library never contains or
// calls this build: it exists only under tests/.
#include <stdlib.h>
#include <string.h>

#include <barrier>
#include <string>
#include <thread>
#include <vector>

#include "tci_engine.h"
#include "tci_problem.h"

using namespace tci;

struct HostTeam {
    // register tile of the one-thread team: one wave of one lane holding up to 160×160
    static constexpr int NT = 1, WS = 1, NW = 1, RA = 160, CB = 160, WPE = 1;
    double lA[1], lX[160], lY[160];
    uint32_t lB[1];
    double &tA(int) { return lA[0]; }
    uint32_t &tB(int) { return lB[0]; }
    double &tX(int, int r) { return lX[r]; }
    double &tY(int, int j) { return lY[j]; }
    int tid() const { return 0; }
    static constexpr int nt() { return 1; }
    int lane() const { return 0; }
    static constexpr int wsize() { return 1; }
    int wave() const { return 0; }
    static constexpr int nwaves() { return 1; }
    double *xs = nullptr, *ys = nullptr;
    double *lpi = nullptr;
    int64_t lpi_cap = 0;
    uint8_t *lrows = nullptr, *lcols = nullptr;
    uint8_t *lsets = nullptr;
    int32_t *lsetn = nullptr;
    double *llut = nullptr;
    int64_t llut_cap = 0;
    void sync() {}
    void wave_sync() {}
    void wave_sync_lds() {}
    ArgBest argmax(ArgBest b) { return b; }
    int uniform(int x) const { return x; }
    uint32_t uniform_u32(uint32_t x) const { return x; }
    double uniform_f64(double x) const { return x; }
    double max_nan(double v) { return v; }
    double wave_max(double v) const { return v; }
    uint32_t wave_min(uint32_t v) const { return v; }
    bool wave_any(bool p) const { return p; }
    int wave_first(bool p) const { return p ? 0 : -1; }
    uint64_t wave_ballot(bool p) const { return p ? 1u : 0u; }
    int wave_bcast(int v, int) const { return v; }
    double wave_bcast_f64(double v, int) const { return v; }
    int any(int v) { return v; }
    uint64_t clock() const { return 0; }
    static constexpr bool stage_lds = false;
    static void assume_lds(const void *) {}
};

// Multi-threaded host team: NT std::threads emulate one workgroup of NT/64
// waves (one thread per lane), with std::barriers for __syncthreads and for
// the wave-level operations and shared slots for the reductions, so the
// device's cross-lane logic runs on the CPU.
struct MTShared {
    std::barrier<> *bar;
    std::vector<std::barrier<> *> wbar;
    std::vector<ArgBest> abuf;
    std::vector<double> dbuf;
    std::vector<int> ibuf;
    std::vector<uint32_t> ubuf;
};

template <int NT_, int RA_, int CB_>
struct MTTile {
    double A[NT_ / 64], X[(NT_ / 64) * 64 * RA_], Y[(NT_ / 64) * CB_];
    uint32_t B[NT_ / 64];
};

template <int NT_, int RA_, int CB_>
struct HostTeamMT {
    static constexpr int NT = NT_, WS = 64, NW = NT_ / 64, RA = RA_, CB = CB_, WPE = 1;
    MTTile<NT_, RA_, CB_> *tile = nullptr;
    double &tA(int w) { return tile->A[w]; }
    uint32_t &tB(int w) { return tile->B[w]; }
    double &tX(int w, int r) { return tile->X[w * 64 * RA + r]; }
    double &tY(int w, int j) { return tile->Y[w * CB + j]; }
    int t_tid = 0;
    int tid() const { return t_tid; }
    static constexpr int nt() { return NT_; }
    int lane() const { return t_tid % 64; }
    static constexpr int wsize() { return 64; }
    int wave() const { return t_tid / 64; }
    static constexpr int nwaves() { return NW; }
    double *xs = nullptr, *ys = nullptr, *lpi = nullptr;
    int64_t lpi_cap = 0;
    uint8_t *lrows = nullptr, *lcols = nullptr;
    uint8_t *lsets = nullptr;
    int32_t *lsetn = nullptr;
    double *llut = nullptr;
    int64_t llut_cap = 0;
    MTShared *sh = nullptr;
    void sync() { sh->bar->arrive_and_wait(); }
    void wsync() { sh->wbar[wave()]->arrive_and_wait(); }
    void wave_sync() { wsync(); }
    void wave_sync_lds() { wsync(); }
    ArgBest argmax(ArgBest b)
    {
        sh->abuf[t_tid] = b;
        sync();
        ArgBest r = sh->abuf[0];
        for (int t = 1; t < NT_; ++t) if (arg_better(sh->abuf[t], r)) r = sh->abuf[t];
        sync();
        return r;
    }
    double max_nan(double v)
    {
        sh->dbuf[t_tid] = v;
        sync();
        double r = sh->dbuf[0];
        for (int t = 1; t < NT_; ++t) r = jl_max(r, sh->dbuf[t]);
        sync();
        return r;
    }
    double wave_max(double v)
    {
        sh->dbuf[t_tid] = v;
        wsync();
        double r = sh->dbuf[wave() * 64];
        for (int t = 1; t < 64; ++t) r = fmax(r, sh->dbuf[wave() * 64 + t]);
        wsync();
        return r;
    }
    uint32_t wave_min(uint32_t v)
    {
        sh->ubuf[t_tid] = v;
        wsync();
        uint32_t r = sh->ubuf[wave() * 64];
        for (int t = 1; t < 64; ++t) r = std::min(r, sh->ubuf[wave() * 64 + t]);
        wsync();
        return r;
    }
    int wave_first(bool p)
    {
        sh->ibuf[t_tid] = p ? 1 : 0;
        wsync();
        int r = -1;
        for (int t = 0; t < 64 && r < 0; ++t) if (sh->ibuf[wave() * 64 + t]) r = t;
        wsync();
        return r;
    }
    bool wave_any(bool p) { return wave_first(p) >= 0; }
    uint64_t wave_ballot(bool p)
    {
        sh->ibuf[t_tid] = p ? 1 : 0;
        wsync();
        uint64_t r = 0;
        for (int t = 0; t < 64; ++t) if (sh->ibuf[wave() * 64 + t]) r |= (uint64_t)1 << t;
        wsync();
        return r;
    }
    int wave_bcast(int v, int l)
    {
        sh->ibuf[t_tid] = v;
        wsync();
        const int r = sh->ibuf[wave() * 64 + l];
        wsync();
        return r;
    }
    double wave_bcast_f64(double v, int l)
    {
        sh->dbuf[t_tid] = v;
        wsync();
        const double r = sh->dbuf[wave() * 64 + l];
        wsync();
        return r;
    }
    int any(int v)
    {
        sh->ibuf[t_tid] = v;
        sync();
        int r = 0;
        for (int t = 0; t < NT_; ++t) r |= sh->ibuf[t];
        sync();
        return r;
    }
    int uniform(int x) const { return x; }
    uint32_t uniform_u32(uint32_t x) const { return x; }
    double uniform_f64(double x) const { return x; }
    uint64_t clock() const { return 0; }
    static constexpr bool stage_lds = false;
    static void assume_lds(const void *) {}
};

// rrlu! through the register tile on NT host threads shaped like the
// 256-thread device team (4 waves of 64 lanes, 2×32 doubles per lane)
extern "C" int eh_rrlu_mt(int mode, double *A, int m, int n, int64_t maxrank, double reltol,
                          double abstol, int leftorth, int32_t *rowperm, int32_t *colperm,
                          double *err)
{
    constexpr int NT = 256, RA = 2, CB = 32;
    using Team = HostTeamMT<NT, RA, CB>;
    using Tile = TileLU<Team, RA, CB>;
    if (!Tile::fits(m, n)) return -100;
    std::barrier<> bar(NT);
    std::vector<std::barrier<> *> wb;
    for (int w = 0; w < NT / 64; ++w) wb.push_back(new std::barrier<>(64));
    MTShared sh;
    sh.bar = &bar;
    sh.wbar = wb;
    sh.abuf.resize(NT);
    sh.dbuf.resize(NT);
    sh.ibuf.resize(NT);
    sh.ubuf.resize(NT);
    MTTile<NT, RA, CB> tile;
    std::vector<int32_t> rp(m), cp(n);
    int pres = 0;
    double eres = 0;
    auto body = [&](int tid) {
        Team tm;
        tm.t_tid = tid;
        tm.sh = &sh;
        tm.tile = &tile;
        Tile R;
        R.setup(tm, m, n);
        R.load(tm, A, m);
        tm.sync();
        double e, st[2];
        int p = R.factor(tm, maxrank, reltol, abstol, leftorth != 0, nullptr, &e, st, nullptr,
                         rp.data(), cp.data());
        tm.sync();
        if (tid == 0) { pres = p; eres = e; }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < NT; ++t) th.emplace_back(body, t);
    for (auto &t : th) t.join();
    for (auto *b : wb) delete b;
    for (int i = 0; i < m; ++i) rowperm[i] = rp[i];
    for (int j = 0; j < n; ++j) colperm[j] = cp[j];
    *err = eres;
    return pres;
}

struct EhSolve {
    std::vector<uint8_t> inputs, arena, report;
    std::vector<double> lpi, llut;    // stand in for the workgroup's LDS areas
    std::vector<uint8_t> ltab, lsets;
    SolveDesc D;
    int status;
};

// give a host team the on-chip areas the device kernel has (Pi area, index
// tables, pivot sets), so the engine's LDS code paths run on the CPU too
template <class Team>
static void attach_lds(Team &tm, EhSolve *h, const SolveCaps &c)
{
    const int64_t tab = (int64_t)c.mrcap * c.L;
    h->ltab.assign((size_t)(2 * tab), 0);
    tm.lrows = h->ltab.data();
    tm.lcols = h->ltab.data() + tab;
    const int64_t sb = (int64_t)NSETKIND * c.L * c.setcap * c.L;
    const int64_t sba = (sb + 15) & ~(int64_t)15;
    h->lsets.assign((size_t)(sba + 4 * NSETKIND * c.L), 0);
    tm.lsets = h->lsets.data();
    tm.lsetn = (int32_t *)(h->lsets.data() + sba);
    h->lpi.assign(16384, 0.0);        // 128 KiB
    tm.lpi = h->lpi.data();
    tm.lpi_cap = (int64_t)h->lpi.size();
    h->llut.assign(8192, 0.0);
    tm.llut = h->llut.data();
    tm.llut_cap = (int64_t)h->llut.size();
}

static std::string g_msg;

extern "C" const char *eh_last_error(void) { return g_msg.c_str(); }

extern "C" void *eh_solve(const tci_problem *p, int *status)
{
    SolveCaps c;
    int st = check_problem(*p, c, &g_msg);
    if (st != TCI_OK) { *status = st; return nullptr; }
    EhSolve *h = new EhSolve();
    size_t ib, sb, rb;
    layout_solve(c, nullptr, nullptr, nullptr, nullptr, &ib, &sb, &rb);
    h->inputs.assign(ib, 0);
    h->arena.assign(sb, 0);
    h->report.assign(rb, 0);
    memset(&h->D, 0, sizeof(SolveDesc));
    layout_solve(c, &h->D, h->inputs.data(), h->arena.data(), h->report.data(), &ib, &sb, &rb);
    fill_desc_scalars(*p, h->D);
    pack_inputs(*p, c, h->D, h->inputs.data(), h->inputs.data());
    static HostTeam tm;   // large register tile: keep it off the stack
    tm = HostTeam();
    tm.xs = h->D.stage;
    tm.ys = h->D.stage + c.mrcap;
    if (getenv("EH_LDS")) attach_lds(tm, h, c);
    EngShared S;
    memset(&S, 0, sizeof S);
    Engine<HostTeam> *e = new Engine<HostTeam>(tm, h->D, S);
    h->status = e->run();
    delete e;
    *status = h->status;
    return h;
}

// The whole solve on 256 host threads shaped like a 4-wave device team, so
// the engine's cross-thread logic (wave reductions, staging, the parallel
// LU of Pᵀ) is exercised on the CPU
extern "C" void *eh_solve_mt(const tci_problem *p, int *status)
{
    constexpr int NT = 256, RA = 2, CB = 32;
    using Team = HostTeamMT<NT, RA, CB>;
    SolveCaps c;
    int st = check_problem(*p, c, &g_msg);
    if (st != TCI_OK) { *status = st; return nullptr; }
    EhSolve *h = new EhSolve();
    size_t ib, sb, rb;
    layout_solve(c, nullptr, nullptr, nullptr, nullptr, &ib, &sb, &rb);
    h->inputs.assign(ib, 0);
    h->arena.assign(sb, 0);
    h->report.assign(rb, 0);
    memset(&h->D, 0, sizeof(SolveDesc));
    layout_solve(c, &h->D, h->inputs.data(), h->arena.data(), h->report.data(), &ib, &sb, &rb);
    fill_desc_scalars(*p, h->D);
    pack_inputs(*p, c, h->D, h->inputs.data(), h->inputs.data());
    std::barrier<> bar(NT);
    std::vector<std::barrier<> *> wb;
    for (int w = 0; w < NT / 64; ++w) wb.push_back(new std::barrier<>(64));
    MTShared sh;
    sh.bar = &bar;
    sh.wbar = wb;
    sh.abuf.resize(NT);
    sh.dbuf.resize(NT);
    sh.ibuf.resize(NT);
    sh.ubuf.resize(NT);
    MTTile<NT, RA, CB> tile;
    EngShared S;
    memset(&S, 0, sizeof S);
    int res = 0;
    Team proto;
    const bool lds = getenv("EH_LDS") != nullptr;
    if (lds) attach_lds(proto, h, c);
    auto body = [&](int tid) {
        Team tm;
        tm.t_tid = tid;
        tm.sh = &sh;
        tm.tile = &tile;
        tm.xs = h->D.stage;
        tm.ys = h->D.stage + c.mrcap;
        if (lds) {
            tm.lrows = proto.lrows; tm.lcols = proto.lcols; tm.lsets = proto.lsets;
            tm.lsetn = proto.lsetn; tm.lpi = proto.lpi; tm.lpi_cap = proto.lpi_cap;
            tm.llut = proto.llut; tm.llut_cap = proto.llut_cap;
        }
        Engine<Team> e(tm, h->D, S);
        int r = e.run();
        if (tid == 0) res = r;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < NT; ++t) th.emplace_back(body, t);
    for (auto &t : th) t.join();
    for (auto *b : wb) delete b;
    h->status = res;
    *status = res;
    return h;
}

// rrlu! through the register tile (modes 0, 1) or the streaming path (mode 2)
// on the one-thread team; the streaming path leaves the packed factors in A.
extern "C" int eh_rrlu(int mode, double *A, int m, int n, int64_t maxrank, double reltol,
                       double abstol, int leftorth, int32_t *rowperm, int32_t *colperm,
                       double *pivval, double *err)
{
    static HostTeam tm;
    tm = HostTeam();
    std::vector<double> xs(m + n + 1), ys(m + n + 1);
    tm.xs = xs.data();
    tm.ys = ys.data();
    double st[2];
    int p;
    if (mode == 2) {
        p = rrlu_team(tm, A, m, n, maxrank, reltol, abstol, leftorth != 0, rowperm, colperm, pivval, err, st);
    } else {
        using Tile = TileLU<HostTeam, HostTeam::RA, HostTeam::CB>;
        if (!Tile::fits(m, n)) return -100;
        Tile *R = new Tile();
        R->setup(tm, m, n);
        R->load(tm, A, m);
        p = R->factor(tm, maxrank, reltol, abstol, leftorth != 0, pivval, err, st, nullptr,
                      rowperm, colperm);
        delete R;
    }
    return p;
}

extern "C" void eh_free(void *hp) { delete (EhSolve *)hp; }
extern "C" int eh_niter(void *hp) { return ((EhSolve *)hp)->D.ints[INT_NITER]; }
extern "C" int eh_errsite(void *hp) { return ((EhSolve *)hp)->D.ints[INT_ERRSITE]; }
extern "C" void eh_history(void *hp, int64_t *ranks, double *errors, int64_t *nglob)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    int n = D.ints[INT_NITER];
    if (n > D.maxiter_cap) n = D.maxiter_cap;
    for (int i = 0; i < n; ++i) {
        ranks[i] = D.ranks[i];
        errors[i] = D.errors[i];
        nglob[i] = D.nglob[i];
    }
}
// which 0/1: Iset/Jset, 2/3: the last history entry (as tci_batch_indexset)
static int eh_kind(const SolveDesc &D, int which)
{
    const int jside = which & 1, hist = D.ints[INT_HIST];
    if (which < 2) return jside ? SET_J : SET_I;
    return hist == 2 ? (jside ? SET_HB_J : SET_HB_I) : (jside ? SET_HA_J : SET_HA_I);
}
extern "C" int eh_setn(void *hp, int which, int b)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    if (which >= 2 && D.ints[INT_HIST] == 0) return 0;
    return D.setn[eh_kind(D, which) * D.L + b];
}
extern "C" void eh_set(void *hp, int which, int b, int32_t *out)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    const int L = D.L, kind = eh_kind(D, which), len = (which & 1) ? L - 1 - b : b;
    const int n = eh_setn(hp, which, b);
    const uint8_t *s = D.sets + ((int64_t)(kind * L + b) * D.setcap) * L;
    for (int i = 0; i < n; ++i)
        for (int t = 0; t < len; ++t) out[i * len + t] = s[(int64_t)i * L + t];
}
extern "C" void eh_Tdims(void *hp, int b, int64_t *d3)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    for (int i = 0; i < 3; ++i) d3[i] = D.Tdims[3 * b + i];
}
extern "C" void eh_T(void *hp, int b, double *out)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    int64_t n = D.Tdims[3 * b] * D.Tdims[3 * b + 1] * D.Tdims[3 * b + 2];
    memcpy(out, D.T + D.tstride * b, sizeof(double) * n);
}
extern "C" double eh_maxsample(void *hp) { return ((EhSolve *)hp)->D.scal[SCAL_MAXSAMPLE]; }
extern "C" int eh_npivoterrors(void *hp) { return ((EhSolve *)hp)->D.ints[INT_NPE]; }
extern "C" void eh_pivoterrors(void *hp, double *out)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    memcpy(out, D.pivoterrors, sizeof(double) * D.ints[INT_NPE]);
}
extern "C" void eh_bonderrors(void *hp, double *out)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    memcpy(out, D.bonderrors, sizeof(double) * (D.L - 1));
}
extern "C" void eh_workstats(void *hp, double *out6)
{
    const SolveDesc &D = ((EhSolve *)hp)->D;
    for (int i = 0; i < 6; ++i) out6[i] = D.scal[SCAL_WS_PI + i];
}
