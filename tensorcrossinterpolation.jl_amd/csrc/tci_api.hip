// tci_api.hip — the C ABI of libtci_hip.so (declared in include/tci.h).
//
// Host-side plumbing only: argument checks, device memory, copies and
// launches on the context's HIP stream.  All arithmetic of the TCI path runs
// in the gfx950 kernels of tci_kernels.hip; there is no CPU fallback.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <mutex>
#include <vector>

#include "../../include/tci.h"
#include "tci_kernels.h"
#include "tci_layout.h"
#include "tci_problem.h"

using namespace tci;

static thread_local std::string g_err;

// Base.max on Float64 (NaN-propagating)
static double nan_max(double a, double b)
{
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

static tci_status set_err(tci_status code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_err(TCI_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Device buffers of finished batches are kept for reuse, so a serving loop
// does not pay hipMalloc/hipFree per launch.
struct PoolBuf {
    void *ptr;
    size_t bytes;
};

struct tci_ctx {
    int device;
    hipStream_t stream;                 // synchronous calls and the first solve stream
    std::vector<int32_t> lu_order;      // host copy of the last batched rrLU job order
    std::vector<hipStream_t> extra;     // more solve streams (tci_ctx_set_streams)
    hipStream_t dl = nullptr;           // downloads of results (never behind a solve)
    size_t next = 0;
    uint64_t budget;
    uint64_t used;
    std::vector<PoolBuf> pool;          // cached device buffers of finished batches
    uint64_t pooled = 0;                // their bytes
    int lds_wg;         // LDS bytes (static + dynamic) a solve workgroup of the default team may use
    int team;           // threads per solve workgroup (the default team)
    int lds_cu = 0, lds_blk = 0, wg_env = 0;   // LDS per CU and per workgroup; TCI_WG_PER_CU (0: unset)
    // measured solve teams (tci_ctx_autotune_solve), ascending by
    // (smallest largest-mrcap, smallest batch)
    struct SolveTune {
        int32_t mrcap, count, team;
        bool operator<(const SolveTune &o) const { return mrcap != o.mrcap ? mrcap < o.mrcap : count < o.count; }
    };
    std::vector<SolveTune> solve_tune;
    int pi_lds_kb;      // LDS budget for Pi matrices (KiB)
    int rrlu_algo = TCI_RRLU_AUTO;   // factorization kernel of the standalone rrLU
    std::vector<tci_rrlu_tune> tune; // measured AUTO choices (tci_ctx_autotune_rrlu), by side
    int grid_team = 0;               // the autotuner's forced cooperative team (0: none)
};

// a device-to-host copy of a finished result on the context's download
// stream: not behind later solves, and not on the legacy default stream
static hipError_t ctx_download(tci_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->dl);
    return e == hipSuccess ? hipStreamSynchronize(ctx->dl) : e;
}

// LDS bytes (static + dynamic) a solve workgroup of `team` threads may use:
// the CU's LDS shared by the workgroups that should be resident together
// (TCI_WG_PER_CU; eight one-wave teams, two of 256 threads, else one), at most the per-workgroup
// limit; the kernel's static LDS is subtracted at launch
static int lds_for_team(const tci_ctx *c, int team)
{
    const int wgs = c->wg_env > 0 ? c->wg_env : (team == 256 ? 2 : team == 64 ? 8 : 1);
    const int per = (c->lds_cu > 0 ? c->lds_cu : c->lds_blk) / wgs;
    return per > c->lds_blk ? c->lds_blk : per;
}

// the team of a batch of `count` solves whose largest Pi-row capacity is
// max_mr: among the tuned entries at or below both, the one with the largest
// capacity and then the largest batch; the context's default team without one
static int solve_team_for(const tci_ctx *c, int max_mr, int count)
{
    int team = c->team;
    for (const auto &e : c->solve_tune)
        if (e.mrcap <= max_mr && e.count <= count) team = e.team;
    return team;
}

// Launch shape of the solve kernel: a team of `team` threads per solve; LDS
// holds the rrLU staging vectors, the index tables and as much of a Pi
// matrix as the budget allows (one resident workgroup per CU when large).
static LdsPlan make_plan(const tci_ctx *c, int64_t left, int max_mrcap, int max_tab, int64_t max_sets,
                         int64_t max_lut)
{
    LdsPlan p;
    memset(&p, 0, sizeof p);
    if (left < 0) left = 0;
    if ((int64_t)2 * max_mrcap * 8 <= left) {
        p.stage_cap = max_mrcap;
        left -= (int64_t)2 * max_mrcap * 8;
    }
    int64_t tab = ((int64_t)max_tab + 15) & ~(int64_t)15;
    if (2 * tab <= left) {
        p.tab_cap = (int32_t)tab;
        left -= 2 * tab;
    }
    // the pivot sets and counts (16-byte aligned parts) when they fit in a
    // quarter of what is left, so the Pi area keeps most of the LDS
    const int64_t sb = (max_sets + 15) & ~(int64_t)15;
    if (max_sets > 0 && sb <= left / 4) {
        p.sets_cap = (int32_t)sb;
        left -= sb;
    }
    // the Lorentzian value table (a few KiB) when it fits in a quarter
    if (max_lut > 0 && max_lut * 8 <= left / 4) {
        p.lut_cap = (int32_t)((max_lut + 1) & ~(int64_t)1);
        left -= (int64_t)p.lut_cap * 8;
    }
    int64_t pib = (int64_t)c->pi_lds_kb * 1024;
    if (pib > left) pib = left;
    p.pi_cap = pib > 0 ? pib / 8 : 0;
    return p;
}

// Device buffers of finished batches are cached and reused best-fit (up to
// twice the size asked for).  hipFree synchronizes the whole device, so a
// cached buffer is freed only when the cache and the buffers in use would
// exceed the context's HBM budget, never to keep the cache small.
static hipError_t pool_get(tci_ctx *c, size_t bytes, void **out, size_t *got)
{
    size_t best = (size_t)-1;
    for (size_t i = 0; i < c->pool.size(); ++i)
        if (c->pool[i].bytes >= bytes && c->pool[i].bytes <= 2 * bytes + (1 << 20) &&
            (best == (size_t)-1 || c->pool[i].bytes < c->pool[best].bytes))
            best = i;
    if (best != (size_t)-1) {
        *out = c->pool[best].ptr;
        *got = c->pool[best].bytes;
        c->pooled -= c->pool[best].bytes;
        c->pool.erase(c->pool.begin() + (long)best);
        return hipSuccess;
    }
    while (!c->pool.empty() && c->used + c->pooled > c->budget) {
        (void)hipFree(c->pool.front().ptr);
        c->pooled -= c->pool.front().bytes;
        c->pool.erase(c->pool.begin());
    }
    *got = bytes;
    return hipMalloc(out, bytes);
}

static void pool_put(tci_ctx *c, void *p, size_t bytes)
{
    if (!p) return;
    c->pool.push_back({p, bytes});
    c->pooled += bytes;
}

// a block of a slot-level server's memory: device memory, or pinned host
// memory with its device alias
struct ServeChunk {
    void *base = nullptr;
    size_t bytes = 0;
    bool device = true;
    int inuse = 0;           // its blocks held by requests
    bool orphan = false;     // the server stopped: freed with its last block
};
struct ServeBlock {
    uint8_t *ptr = nullptr;  // host view (pinned) or device pointer
    uint8_t *dev = nullptr;  // the device's view
    ServeChunk *chunk = nullptr;
};

struct tci_batch {
    tci_ctx *ctx;
    int count;
    std::vector<SolveDesc> desc;      // host copy (device pointers)
    std::vector<size_t> rep_off;
    uint8_t *d_in = nullptr, *d_state = nullptr, *d_rep = nullptr;
    SolveDesc *d_desc = nullptr;
    size_t bytes = 0, rep_bytes = 0, in_bytes = 0, state_bytes = 0, desc_bytes = 0;
    size_t pool_bytes[4] = {0, 0, 0, 0};   // sizes of the pool buffers held (inputs, state, report, descriptors)
    std::vector<uint8_t> h_in;        // upload image, alive until the copy ran
    std::vector<uint8_t> h_rep;
    bool fetched = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t stream = nullptr;
    // a request of a slot-level server (no events: its done mark is polled)
    bool served = false;
    bool polled = false;               // returned by tci_server_poll
    bool lost = false;                 // its server stopped on an error before the ticket ran
    tci_server *server = nullptr;      // null once the server stopped
    uint64_t ticket = 0;
    ServeBlock dblk{}, hblk{};         // its device block and pinned staging block
    int dcls = 0, hcls = 0;            // their size classes
    const uint8_t *h_rep_pin = nullptr;   // the report's pinned host copy
};

// A slot-level server: the persistent serving grid and its queue.
struct tci_server {
    tci_ctx *ctx = nullptr;
    hipStream_t kst = nullptr;                  // the grid's stream (a hardware queue of its own)
    ServeCtl *ctl = nullptr, *d_ctl = nullptr;  // coherent host memory and its device alias
    uint64_t *done = nullptr, *d_done = nullptr;
    ServeSlot *ring = nullptr, *d_ring = nullptr;
    ServeDev *d_g = nullptr;                    // the grid's words in device memory
    int cap = 0, team = 256, blocks = 0;
    uint64_t next = 0, idle_ticks = 0;
    uint32_t epoch = 0;                         // launches so far
    LdsPlan plan{};
    hipEvent_t kev = nullptr;
    bool running = false;
    std::vector<tci_batch *> slot;              // the request of each ring slot in flight
    std::vector<tci_batch *> pending;           // submitted and not yet returned by a poll
    // cached blocks per size class (device blocks, pinned staging blocks):
    // carved from chunks that live until the server stops and their last
    // block is released
    std::vector<std::vector<ServeBlock>> dfree, hfree;
    std::vector<ServeChunk *> chunks;
};

// ---- scratch of tci_rrlu_batched_noalloc_f64, the entry point that
// allocates nothing.  A file-scope device array: every GPU the library runs
// on holds its own copy, at its own address, until the process ends.  Host
// code never passes the array's name to a kernel or a copy (it is a host-side
// stand-in there); it uses the per-device address that hipGetSymbolAddress
// returns, looked up once per device at context creation.
static constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[kMaxDevices];

// with `device` current: this device's address of `scratch` (looked up once)
static tci_status scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return set_err(TCI_ERR_DEVICE, "device %d is beyond the %d devices the scratch table holds", device,
                       kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return TCI_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_err(TCI_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[device] = (uint8_t *)p;
    return TCI_OK;
}

static uint8_t *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

extern "C" {

const char *tci_last_error(void) { return g_err.c_str(); }
const char *tci_version(void) { return "tci-mi355x 0.1 (gfx950, FP64)"; }

tci_status tci_ctx_create(tci_ctx **out, int32_t device, uint64_t hbm_budget_bytes)
{
    if (!out) return set_err(TCI_ERR_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_err(TCI_ERR_DEVICE, "no HIP device available (this library has no CPU fallback)");
    if (device < 0 || device >= n) return set_err(TCI_ERR_DEVICE, "device %d out of range", device);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(TCI_ERR_DEVICE, "device %d is %s; this library is built for gfx950 only",
                       device, prop.gcnArchName);
    HIPCHK(hipSetDevice(device));
    const tci_status ss = scratch_setup(device);
    if (ss != TCI_OK) return ss;
    tci_ctx *c = new tci_ctx();
    c->device = device;
    c->used = 0;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->dl, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return set_err(TCI_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (hbm_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        hbm_budget_bytes = (uint64_t)(0.8 * (double)fr);
    }
    c->budget = hbm_budget_bytes;
    // LDS per solve workgroup: the CU's LDS shared by the workgroups that
    // should be resident together (TCI_WG_PER_CU; 2 for the default team),
    // at most the per-workgroup limit; the kernel's static LDS is subtracted
    // at launch
    int lds_blk = 0, lds_cu = 0;
    HIPCHK(hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HIPCHK(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
    const char *ev = getenv("TCI_TEAM");
    c->team = ev ? atoi(ev) : 256;
    if (c->team != 64 && c->team != 256 && c->team != 512 && c->team != 1024) c->team = 256;
    ev = getenv("TCI_WG_PER_CU");
    c->wg_env = ev && atoi(ev) > 0 ? atoi(ev) : 0;
    c->lds_cu = lds_cu;
    c->lds_blk = lds_blk;
    c->lds_wg = lds_for_team(c, c->team);
    ev = getenv("TCI_PI_LDS_KB");
    c->pi_lds_kb = ev ? atoi(ev) : 1 << 20;
    *out = c;
    return TCI_OK;
}

tci_status tci_ctx_destroy(tci_ctx *ctx)
{
    if (!ctx) return TCI_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t s : ctx->extra) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (PoolBuf &b : ctx->pool) (void)hipFree(b.ptr);
    (void)hipStreamDestroy(ctx->dl);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TCI_OK;
}

tci_status tci_ctx_synchronize(tci_ctx *ctx)
{
    if (!ctx) return set_err(TCI_ERR_ARG, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (hipStream_t s : ctx->extra) HIPCHK(hipStreamSynchronize(s));
    return TCI_OK;
}

tci_status tci_ctx_stream(tci_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return set_err(TCI_ERR_ARG, "bad arguments");
    *stream = (void *)ctx->stream;
    return TCI_OK;
}

tci_status tci_ctx_set_rrlu_algo(tci_ctx *ctx, int32_t algo)
{
    if (!ctx || algo < TCI_RRLU_AUTO || algo > TCI_RRLU_GRID) return set_err(TCI_ERR_ARG, "bad arguments");
    ctx->rrlu_algo = algo;
    return TCI_OK;
}

tci_status tci_ctx_set_streams(tci_ctx *ctx, int32_t n)
{
    if (!ctx || n < 1 || n > 64) return set_err(TCI_ERR_ARG, "need 1 <= n <= 64 streams");
    HIPCHK(hipSetDevice(ctx->device));
    while ((int)ctx->extra.size() > n - 1) {
        HIPCHK(hipStreamSynchronize(ctx->extra.back()));
        HIPCHK(hipStreamDestroy(ctx->extra.back()));
        ctx->extra.pop_back();
    }
    while ((int)ctx->extra.size() < n - 1) {
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        ctx->extra.push_back(s);
    }
    ctx->next = 0;
    return TCI_OK;
}

// the stream a new batch goes to: an idle one if any, else round robin
static hipStream_t pick_stream(tci_ctx *ctx)
{
    const size_t ns = 1 + ctx->extra.size();
    for (size_t t = 0; t < ns; ++t) {
        const size_t q = (ctx->next + t) % ns;
        hipStream_t s = q == 0 ? ctx->stream : ctx->extra[q - 1];
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) {
            ctx->next = (q + 1) % ns;
            return s;
        }
        (void)hipGetLastError();   // a busy stream is not an error: clear it
    }
    const size_t q = ctx->next % ns;
    ctx->next = (q + 1) % ns;
    return q == 0 ? ctx->stream : ctx->extra[q - 1];
}

// ------------------------------------------------------------------- rrLU
// The job shapes of a batched rrLU, read once, and their partition between
// the register-tile kernel (jobs up to 128×128, when the context's
// algorithm allows it) and the streaming one.  tile_budget: the most tile
// scratch (doubles) the tile jobs may take, -1 for no limit; a job past it
// goes to the streaming kernel (the results are the same bits).
struct LuPlan {
    std::vector<tci_lu_job> h;
    std::vector<int32_t> tile, stream, grid;
    std::vector<int64_t> soff;
    int64_t sc_tot = 0;   // tile scratch, doubles
    int64_t mx = 0;       // largest side of any job
};

static tci_status lu_partition(tci_ctx *ctx, int64_t tile_budget, LuPlan &P, bool allow_grid = true);

// the cooperative factorization takes a job of at least this many elements
// when a batch has at most kGridMaxJobs of them (AUTO)
constexpr int64_t kGridMinElems = (int64_t)384 * 384;
constexpr int64_t kGridMaxJobs = 8;

// reads and checks the job table (and the largest side), partitions it
static tci_status lu_read(tci_ctx *ctx, tci_lu_job *jobs, int64_t count, LuPlan &P)
{
    P.h.resize((size_t)count);
    HIPCHK(hipMemcpyAsync(P.h.data(), jobs, sizeof(tci_lu_job) * (size_t)count, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (const tci_lu_job &j : P.h) {
        if (j.m < 0 || j.n < 0 || j.m > INT32_MAX / 2 || j.n > INT32_MAX / 2 || j.m * j.n > ((int64_t)1 << 31))
            return set_err(TCI_ERR_ARG, "job shape out of range");
        if ((j.m > 0 && (!j.A || !j.rowperm)) || (j.n > 0 && (!j.A || !j.colperm)))
            return set_err(TCI_ERR_ARG, "job has NULL buffers");
        if (j.maxrank < 0) return set_err(TCI_ERR_ARG, "maxrank must be >= 0");
        int64_t s = j.m > j.n ? j.m : j.n;
        P.mx = s > P.mx ? s : P.mx;
    }
    return TCI_OK;
}

static tci_status lu_plan(tci_ctx *ctx, tci_lu_job *jobs, int64_t count, int64_t tile_budget, LuPlan &P)
{
    const tci_status s = lu_read(ctx, jobs, count, P);
    return s != TCI_OK ? s : lu_partition(ctx, tile_budget, P);
}

// the tuned entry for a job of largest side s: the first entry at least s,
// the last one beyond them; null when the context is not tuned
static const tci_rrlu_tune *tune_for(const tci_ctx *c, int64_t m, int64_t n)
{
    if (c->tune.empty()) return nullptr;
    const int64_t s = m > n ? m : n;
    for (const tci_rrlu_tune &t : c->tune)
        if (t.side >= s) return &t;
    return &c->tune.back();
}

static tci_status lu_partition(tci_ctx *ctx, int64_t tile_budget, LuPlan &P, bool allow_grid)
{
    const int64_t count = (int64_t)P.h.size();
    for (int64_t q = 0; q < count; ++q) {
        const tci_lu_job &j = P.h[(size_t)q];
        bool fits = j.m >= 1 && j.n >= 1 && j.m <= kTileSide && j.n <= kTileSide;
        if (fits && ctx->rrlu_algo == TCI_RRLU_AUTO) {
            const tci_rrlu_tune *t = tune_for(ctx, j.m, j.n);
            if (t && !t->tile) fits = false;
        }
        if (ctx->rrlu_algo == TCI_RRLU_TILE && !fits && j.m > 0 && j.n > 0)
            return set_err(TCI_ERR_ARG, "job %lld (%lld x %lld) does not fit the 128 x 128 register tile",
                           (long long)q, (long long)j.m, (long long)j.n);
        int64_t mr = j.maxrank < j.m ? j.maxrank : j.m;
        if (j.n < mr) mr = j.n;
        const int64_t need = mr * (j.m + j.n);
        if (fits && ctx->rrlu_algo != TCI_RRLU_STREAM && (tile_budget < 0 || P.sc_tot + need <= tile_budget)) {
            P.tile.push_back((int32_t)q);
            P.soff.push_back(P.sc_tot);
            P.sc_tot += need;
        } else {
            if (ctx->rrlu_algo == TCI_RRLU_TILE && fits)
                return set_err(TCI_ERR_ARG, "the register-tile scratch cannot hold job %lld", (long long)q);
            P.stream.push_back((int32_t)q);
        }
    }
    // jobs that go to the cooperative whole-chip factorization: with GRID
    // every job that fits; with AUTO the large ones of a batch too small to
    // fill the chip with one workgroup per job
    if (allow_grid && (ctx->rrlu_algo == TCI_RRLU_GRID || ctx->rrlu_algo == TCI_RRLU_AUTO)) {
        std::vector<int32_t> big, rest;
        // tuned: the fewest jobs any of the large ones' entries allows
        int64_t maxjobs = ctx->tune.empty() ? kGridMaxJobs : INT64_MAX;
        for (int32_t q : P.stream) {
            const tci_lu_job &j = P.h[(size_t)q];
            const bool fits = j.m >= 1 && j.n >= 1 && j.m <= kGridMaxSide && j.n <= kGridMaxSide;
            const tci_rrlu_tune *t = tune_for(ctx, j.m, j.n);
            const bool large = t ? t->grid_max_jobs > 0 : j.m * j.n >= kGridMinElems;
            if (fits && (ctx->rrlu_algo == TCI_RRLU_GRID || large)) {
                big.push_back(q);
                if (t && t->grid_max_jobs < maxjobs) maxjobs = t->grid_max_jobs;
            } else {
                rest.push_back(q);
            }
        }
        if (ctx->rrlu_algo == TCI_RRLU_GRID || (!big.empty() && (int64_t)big.size() <= maxjobs)) {
            P.grid = big;
            P.stream = rest;
        }
    }
    return TCI_OK;
}

// the cooperative factorization of the grid jobs, one after the other
static tci_status lu_run_grid(tci_ctx *ctx, tci_lu_job *jobs, const LuPlan &P)
{
    if (P.grid.empty()) return TCI_OK;
    hipStream_t st = ctx->stream;
    int64_t mx = 0;
    for (int32_t q : P.grid) {
        const tci_lu_job &j = P.h[(size_t)q];
        mx = j.m * j.n > mx ? j.m * j.n : mx;
    }
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const size_t b_out = lay_align(sizeof(double) * (size_t)mx), b_cand = lay_align(rrlu_grid_cand_bytes(ncu));
    uint8_t *d = nullptr;
    HIPCHK(hipMallocAsync((void **)&d, b_out + b_cand + 256, st));
    hipError_t e = hipSuccess;
    for (int32_t q : P.grid) {
        const tci_lu_job &j = P.h[(size_t)q];
        // the team: the autotuner's forced one, the tuned one, or the default
        const tci_rrlu_tune *t = tune_for(ctx, j.m, j.n);
        const int team = ctx->grid_team ? ctx->grid_team : t ? t->grid_team : rrlu_grid_default_team();
        const int G = rrlu_grid_blocks(ctx->device, rrlu_grid_lds(j.m, j.n), team);
        if (G < 1 || G > ncu) {
            e = hipErrorInvalidConfiguration;
            break;
        }
        int blocks = G < (int)j.n ? G : (int)j.n;
        int32_t *flags = (int32_t *)(d + b_out + b_cand);
        e = hipMemsetAsync(flags, 0, sizeof(int32_t), st);
        if (e == hipSuccess) e = launch_rrlu_grid(jobs + q, j.m, j.n, (double *)d, d + b_out, flags, blocks, team, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(j.A, d, sizeof(double) * (size_t)(j.m * j.n), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) break;
    }
    (void)hipFreeAsync(d, st);
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "cooperative rrlu: %s", hipGetErrorString(e));
    return TCI_OK;
}

// largest side of the streaming jobs
static int64_t stream_side(const LuPlan &P)
{
    int64_t mx = 0;
    for (int32_t q : P.stream) {
        const int64_t s = P.h[(size_t)q].m > P.h[(size_t)q].n ? P.h[(size_t)q].m : P.h[(size_t)q].n;
        mx = s > mx ? s : mx;
    }
    return mx;
}

// longest-processing-time order of the streaming jobs (largest first), so
// that the workgroups that start last are the short ones; results do not
// depend on the order
static void stream_order(const LuPlan &P, std::vector<int32_t> &order)
{
    order = P.stream;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return P.h[(size_t)a].m * P.h[(size_t)a].n > P.h[(size_t)b].m * P.h[(size_t)b].n;
    });
}

constexpr int64_t kLuLdsStageCap = 48 * 1024 / 16;   // doubles per staging vector in LDS

tci_status tci_rrlu_batched_f64(tci_ctx *ctx, tci_lu_job *jobs, int64_t count)
{
    if (!ctx || (!jobs && count > 0) || count < 0) return set_err(TCI_ERR_ARG, "bad arguments");
    if (count == 0) return TCI_OK;
    if (count > INT32_MAX) return set_err(TCI_ERR_ARG, "too many jobs");
    HIPCHK(hipSetDevice(ctx->device));
    LuPlan P;
    tci_status ps = lu_plan(ctx, jobs, count, -1, P);
    if (ps != TCI_OK) return ps;
    std::vector<int32_t> &tile = P.tile, &stream = P.stream;
    std::vector<int64_t> &soff = P.soff;
    const int64_t sc_tot = P.sc_tot;
    int64_t mx;
    ps = lu_run_grid(ctx, jobs, P);
    if (ps != TCI_OK) return ps;
    if (!tile.empty()) {
        uint8_t *d = nullptr;
        const size_t b_idx = lay_align(sizeof(int32_t) * tile.size());
        const size_t b_off = lay_align(sizeof(int64_t) * soff.size());
        HIPCHK(hipMallocAsync((void **)&d, b_idx + b_off + sizeof(double) * (size_t)(sc_tot + 1), ctx->stream));
        hipError_t e = hipMemcpyAsync(d, tile.data(), sizeof(int32_t) * tile.size(), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d + b_idx, soff.data(), sizeof(int64_t) * soff.size(), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = launch_rrlu_tile(jobs, (const int32_t *)d, (int64_t)tile.size(), (double *)(d + b_idx + b_off),
                                 (const int64_t *)(d + b_idx), ctx->stream);
        // the host vectors must outlive the copies
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFreeAsync(d, ctx->stream);
        if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "rrlu tile launch: %s", hipGetErrorString(e));
    }
    if (stream.empty()) return TCI_OK;
    mx = stream_side(P);
    double *gstage = nullptr;
    if (mx > kLuLdsStageCap) HIPCHK(hipMallocAsync((void **)&gstage, sizeof(double) * 2 * mx * count, ctx->stream));
    // (the host copy of the order lives in the context: the next call on
    // this stream synchronizes before it reuses it)
    std::vector<int32_t> &order = ctx->lu_order;
    stream_order(P, order);
    const int64_t ns = (int64_t)order.size();
    int32_t *dorder = nullptr;
    HIPCHK(hipMallocAsync((void **)&dorder, sizeof(int32_t) * (size_t)ns, ctx->stream));
    HIPCHK(hipMemcpyAsync(dorder, order.data(), sizeof(int32_t) * (size_t)ns, hipMemcpyHostToDevice,
                          ctx->stream));
    // (staging is indexed by job number, so it covers all `count` jobs)
    hipError_t e = launch_rrlu(jobs, ns, mx, gstage, 2 * mx, dorder, ctx->stream);
    (void)hipFreeAsync(dorder, ctx->stream);
    if (gstage) (void)hipFreeAsync(gstage, ctx->stream);
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "rrlu launch: %s", hipGetErrorString(e));
    return TCI_OK;
}

// ---- the autotuner of the standalone rrLU (tci_ctx_autotune_rrlu)
static void tune_lcg(uint64_t &x, double *v, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        x = x * 6364136223846793005ULL + 1442695040888963407ULL;
        v[i] = (double)(int64_t)(x >> 11) * 0x1.0p-52 - 1.0;   // in [-1, 1)
    }
}

tci_status tci_ctx_autotune_rrlu(tci_ctx *ctx, tci_rrlu_tune *table, int32_t cap, int32_t *count)
{
    if (!ctx || !count || cap < 0 || (cap > 0 && !table)) return set_err(TCI_ERR_ARG, "bad arguments");
    *count = 0;
    HIPCHK(hipSetDevice(ctx->device));
    static const int kSides[] = {32, 64, 128, 256, 512, 1024, 2048, 4096};
    static const int kTeams[] = {256, 512, 1024};
    const int algo0 = ctx->rrlu_algo;
    std::vector<tci_rrlu_tune> tuned0 = ctx->tune;
    ctx->tune.clear();
    std::vector<tci_rrlu_tune> out;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint8_t *mem = nullptr;
    tci_status st = TCI_OK;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    uint64_t seed = 0x243F6A8885A308D3ULL;
    for (int s : kSides) {
        if (he != hipSuccess || st != TCI_OK) break;
        const int r = s / 2 < 64 ? s / 2 : 64;
        const int kmax = s <= kTileSide ? 512 : 64;
        const size_t mat = (size_t)s * s;
        // pristine matrix, kmax working copies, X, Y, permutations, jobs
        const size_t b_mat = lay_align(sizeof(double) * mat);
        const size_t b_xy = lay_align(sizeof(double) * (size_t)s * r);
        const size_t b_perm = lay_align(sizeof(int32_t) * 2 * (size_t)s);
        const size_t b_jobs = lay_align(sizeof(tci_lu_job) * (size_t)kmax);
        const size_t total = b_mat * (1 + (size_t)kmax) + 2 * b_xy + b_perm * (size_t)kmax + b_jobs;
        if (ctx->used + total > ctx->budget) break;   // larger buckets do not fit the budget
        if ((he = hipMalloc((void **)&mem, total)) != hipSuccess) break;
        double *A0 = (double *)mem, *W = (double *)(mem + b_mat);
        double *X = (double *)(mem + b_mat * (1 + (size_t)kmax)), *Y = (double *)((uint8_t *)X + b_xy);
        int32_t *perm = (int32_t *)((uint8_t *)Y + b_xy);
        tci_lu_job *djobs = (tci_lu_job *)((uint8_t *)perm + b_perm * (size_t)kmax);
        // A0 = X·Y, rank r
        std::vector<double> hx((size_t)s * r), hy((size_t)s * r);
        tune_lcg(seed, hx.data(), hx.size());
        tune_lcg(seed, hy.data(), hy.size());
        he = hipMemcpy(X, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(Y, hy.data(), sizeof(double) * hy.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess) he = launch_gemm(s, s, r, X, s, Y, r, A0, s, ctx->stream);
        std::vector<tci_lu_job> hj((size_t)kmax);
        for (int q = 0; q < kmax; ++q) {
            tci_lu_job &j = hj[(size_t)q];
            memset(&j, 0, sizeof j);
            j.A = (double *)((uint8_t *)W + b_mat * (size_t)q);
            j.rowperm = (int32_t *)((uint8_t *)perm + b_perm * (size_t)q);
            j.colperm = j.rowperm + s;
            j.m = j.n = s;
            j.maxrank = s;
            j.reltol = 1e-10;
            j.abstol = 0.0;
            j.leftorth = 1;
        }
        if (he == hipSuccess) he = hipMemcpy(djobs, hj.data(), sizeof(tci_lu_job) * hj.size(), hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
        // k fresh copies of A0, then the factorization timed on the context's stream
        auto run = [&](int algo, int team, int k, float *ms) -> bool {
            for (int q = 0; q < k && he == hipSuccess; ++q)
                he = hipMemcpyAsync((uint8_t *)W + b_mat * (size_t)q, A0, sizeof(double) * mat,
                                    hipMemcpyDeviceToDevice, ctx->stream);
            if (he != hipSuccess) return false;
            ctx->rrlu_algo = algo;
            ctx->grid_team = team;
            he = hipEventRecord(e0, ctx->stream);
            if (he != hipSuccess) return false;
            st = tci_rrlu_batched_f64(ctx, djobs, k);
            if (st != TCI_OK) return false;
            if ((he = hipEventRecord(e1, ctx->stream)) != hipSuccess) return false;
            if ((he = hipEventSynchronize(e1)) != hipSuccess) return false;
            return (he = hipEventElapsedTime(ms, e0, e1)) == hipSuccess;
        };
        tci_rrlu_tune t;
        memset(&t, 0, sizeof t);
        t.side = s;
        float ms = 0.0f, warm = 0.0f;
        bool ok = run(TCI_RRLU_STREAM, 0, 1, &warm) && run(TCI_RRLU_STREAM, 0, 1, &t.ms_stream);
        if (ok && s <= kTileSide) {
            ok = run(TCI_RRLU_TILE, 0, 1, &warm) && run(TCI_RRLU_TILE, 0, 512, &t.ms_tile_batch) &&
                 run(TCI_RRLU_STREAM, 0, 512, &t.ms_stream_batch);
            t.tile = t.ms_tile_batch < t.ms_stream_batch;
        }
        t.ms_grid = -1.0f;
        for (int team : kTeams) {
            if (!ok) break;
            ok = run(TCI_RRLU_GRID, team, 1, &warm) && run(TCI_RRLU_GRID, team, 1, &ms);
            if (ok && (t.ms_grid < 0.0f || ms < t.ms_grid)) {
                t.ms_grid = ms;
                t.grid_team = team;
            }
        }
        // the most jobs k for which k cooperative runs beat one streaming
        // launch of k (a launch is never faster than its slowest job)
        for (int k = 1; ok && k <= 64; k *= 2) {
            if ((double)k * t.ms_grid < t.ms_stream) {
                t.grid_max_jobs = k;
                continue;
            }
            ok = run(TCI_RRLU_STREAM, 0, k, &ms);
            if (!ok || (double)k * t.ms_grid >= ms) break;
            t.grid_max_jobs = k;
        }
        (void)hipFree(mem);
        mem = nullptr;
        if (!ok) break;
        out.push_back(t);
    }
    if (mem) (void)hipFree(mem);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    ctx->rrlu_algo = algo0;
    ctx->grid_team = 0;
    if (st != TCI_OK) {
        ctx->tune = tuned0;
        return st;
    }
    if (he != hipSuccess) {
        ctx->tune = tuned0;
        return set_err(TCI_ERR_DEVICE, "rrlu autotune: %s", hipGetErrorString(he));
    }
    ctx->tune = out;
    for (size_t q = 0; q < out.size() && (int32_t)q < cap; ++q) table[q] = out[q];
    *count = (int32_t)out.size();
    return TCI_OK;
}

tci_status tci_ctx_set_rrlu_tuning(tci_ctx *ctx, const tci_rrlu_tune *table, int32_t n)
{
    if (!ctx || n < 0 || (n > 0 && !table)) return set_err(TCI_ERR_ARG, "bad arguments");
    for (int32_t q = 0; q < n; ++q) {
        const tci_rrlu_tune &t = table[q];
        if (t.side < 1 || (q > 0 && t.side <= table[q - 1].side) || t.grid_max_jobs < 0 ||
            (t.grid_team != 256 && t.grid_team != 512 && t.grid_team != 1024) || (t.tile && t.side > kTileSide))
            return set_err(TCI_ERR_ARG, "tuning entry %d: sides ascending, team 256/512/1024, tile only up to 128",
                           (int)q);
    }
    ctx->tune.assign(table, table + n);
    return TCI_OK;
}

// tci_rrlu_batched_f64 without any device allocation: the job indices, the
// streaming order, the staging vectors and the register-tile scratch live in
// this device's `scratch` (kNoAllocScratchBytes).  Layout (offsets from the
// scratch base, each part 256-byte aligned): tile job indices (int32 ×
// count), their tile-scratch offsets (int64 × count), the streaming order
// (int32 × count), the staging vectors (2·side doubles per job, only when a
// side exceeds the LDS staging), then the tile scratch.  Every part is
// written (copies or kernels) before it is read, so nothing is cleared.
tci_status tci_rrlu_batched_noalloc_f64(tci_ctx *ctx, tci_lu_job *jobs, int64_t count)
{
    if (!ctx || (!jobs && count > 0) || count < 0) return set_err(TCI_ERR_ARG, "bad arguments");
    if (count == 0) return TCI_OK;
    if (count > INT32_MAX) return set_err(TCI_ERR_ARG, "too many jobs");
    HIPCHK(hipSetDevice(ctx->device));
    uint8_t *base = scratch_of(ctx->device);
    if (!base) return set_err(TCI_ERR_DEVICE, "device %d has no scratch (create its context first)", ctx->device);
    LuPlan P;
    tci_status ps = lu_read(ctx, jobs, count, P);
    if (ps != TCI_OK) return ps;
    const size_t o_idx = 0;
    const size_t o_soff = o_idx + lay_align(sizeof(int32_t) * (size_t)count);
    const size_t o_order = o_soff + lay_align(sizeof(int64_t) * (size_t)count);
    const size_t o_stage = o_order + lay_align(sizeof(int32_t) * (size_t)count);
    const bool staged = P.mx > kLuLdsStageCap;
    const size_t stage_bytes = staged ? sizeof(double) * 2 * (size_t)P.mx * (size_t)count : 0;
    const size_t o_tile = o_stage + lay_align(stage_bytes);
    if (o_tile > kNoAllocScratchBytes)
        return set_err(TCI_ERR_ARG, "this batch needs %zu bytes of scratch; the limit is %zu bytes per device",
                       o_tile, kNoAllocScratchBytes);
    // tile jobs while their scratch fits, the others stream (same bits)
    ps = lu_partition(ctx, (int64_t)((kNoAllocScratchBytes - o_tile) / sizeof(double)), P, false);
    if (ps != TCI_OK) return ps;
    std::vector<int32_t> order;
    stream_order(P, order);
    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
    if (!P.tile.empty()) {
        e = hipMemcpyAsync(base + o_idx, P.tile.data(), sizeof(int32_t) * P.tile.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(base + o_soff, P.soff.data(), sizeof(int64_t) * P.soff.size(), hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess && !order.empty())
        e = hipMemcpyAsync(base + o_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice, st);
    // the host vectors must outlive the copies
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && !P.tile.empty())
        e = launch_rrlu_tile(jobs, (const int32_t *)(base + o_idx), (int64_t)P.tile.size(),
                             (double *)(base + o_tile), (const int64_t *)(base + o_soff), st);
    if (e == hipSuccess && !order.empty()) {
        const int64_t mx = stream_side(P);
        e = launch_rrlu(jobs, (int64_t)order.size(), mx, staged ? (double *)(base + o_stage) : nullptr, 2 * mx,
                        (const int32_t *)(base + o_order), st);
    }
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "rrlu (no allocation): %s", hipGetErrorString(e));
    return TCI_OK;
}

tci_status tci_rrlu_f64(tci_ctx *ctx, const double *A, int64_t lda, int64_t m, int64_t n,
                        int64_t maxrank, double reltol, double abstol, int32_t leftorth,
                        int64_t *rowperm, int64_t *colperm, double *Lh, int64_t ldl, double *Uh,
                        int64_t ldu, int64_t *npivot, double *error)
{
    if (!ctx || m < 0 || n < 0 || lda < (m > 0 ? m : 1) || maxrank < 0 || !npivot || !error)
        return set_err(TCI_ERR_ARG, "bad arguments");
    if ((m > 0 && n > 0 && !A) || (m > 0 && !rowperm) || (n > 0 && !colperm))
        return set_err(TCI_ERR_ARG, "NULL buffer");
    int64_t cap = m < n ? m : n;
    if (maxrank < cap) cap = maxrank;
    if ((cap > 0 && m > 0 && (!Lh || ldl < m)) || (cap > 0 && n > 0 && (!Uh || ldu < cap)))
        return set_err(TCI_ERR_ARG, "L/U buffers too small");
    if (m * n > ((int64_t)1 << 31)) return set_err(TCI_ERR_ARG, "matrix too large");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t mn = (size_t)(m * n);
    double *dA = nullptr, *dL = nullptr, *dU = nullptr;
    int32_t *dperm = nullptr;
    tci_lu_job *dj = nullptr;
    HIPCHK(hipMallocAsync((void **)&dA, sizeof(double) * (mn ? mn : 1), st));
    HIPCHK(hipMallocAsync((void **)&dperm, sizeof(int32_t) * (size_t)(m + n + 1), st));
    HIPCHK(hipMallocAsync((void **)&dj, sizeof(tci_lu_job), st));
    if (mn) HIPCHK(hipMemcpy2DAsync(dA, sizeof(double) * m, A, sizeof(double) * lda, sizeof(double) * m,
                                    n, hipMemcpyHostToDevice, st));
    tci_lu_job j;
    memset(&j, 0, sizeof j);
    j.A = dA;
    j.rowperm = dperm;
    j.colperm = dperm + m;
    j.m = m;
    j.n = n;
    j.maxrank = maxrank;
    j.reltol = reltol;
    j.abstol = abstol;
    j.leftorth = leftorth ? 1 : 0;
    HIPCHK(hipMemcpyAsync(dj, &j, sizeof j, hipMemcpyHostToDevice, st));
    tci_status s = tci_rrlu_batched_f64(ctx, dj, 1);
    if (s != TCI_OK) return s;
    HIPCHK(hipMemcpyAsync(&j, dj, sizeof j, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int64_t p = j.npivot;
    if (p > 0) {
        HIPCHK(hipMallocAsync((void **)&dL, sizeof(double) * (size_t)(m * p + 1), st));
        HIPCHK(hipMallocAsync((void **)&dU, sizeof(double) * (size_t)(p * n + 1), st));
        HIPCHK(launch_lu_unpack(dA, m, n, p, j.leftorth, dL, m, dU, p, st));
        HIPCHK(hipMemcpy2DAsync(Lh, sizeof(double) * ldl, dL, sizeof(double) * m, sizeof(double) * m, p,
                                hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpy2DAsync(Uh, sizeof(double) * ldu, dU, sizeof(double) * p, sizeof(double) * p, n,
                                hipMemcpyDeviceToHost, st));
    }
    std::vector<int32_t> perm((size_t)(m + n + 1));
    HIPCHK(hipMemcpyAsync(perm.data(), dperm, sizeof(int32_t) * (size_t)(m + n), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < m; ++i) rowperm[i] = perm[(size_t)i];
    for (int64_t q = 0; q < n; ++q) colperm[q] = perm[(size_t)(m + q)];
    (void)hipFreeAsync(dA, st);
    (void)hipFreeAsync(dperm, st);
    (void)hipFreeAsync(dj, st);
    if (dL) (void)hipFreeAsync(dL, st);
    if (dU) (void)hipFreeAsync(dU, st);
    *npivot = p;
    *error = j.error;
    if (j.status == TCI_ERR_NAN_L) return set_err(TCI_ERR_NAN_L, "lu.L contains NaNs");
    if (j.status == TCI_ERR_NAN_U) return set_err(TCI_ERR_NAN_U, "lu.U contains NaNs");
    return TCI_OK;
}

tci_status tci_luci_f64(tci_ctx *ctx, const double *A, int64_t lda, int64_t m, int64_t n, int64_t maxrank,
                        double reltol, double abstol, int32_t leftorth, int64_t *rowperm, int64_t *colperm,
                        double *left, int64_t ldleft, double *right, int64_t ldright, int64_t *npivot,
                        double *error)
{
    if (!ctx || m < 0 || n < 0 || lda < (m > 0 ? m : 1) || maxrank < 0 || !npivot || !error)
        return set_err(TCI_ERR_ARG, "bad arguments");
    if ((m > 0 && n > 0 && !A) || (m > 0 && !rowperm) || (n > 0 && !colperm))
        return set_err(TCI_ERR_ARG, "NULL buffer");
    int64_t cap = m < n ? m : n;
    if (maxrank < cap) cap = maxrank;
    if ((cap > 0 && m > 0 && (!left || ldleft < m)) || (cap > 0 && n > 0 && (!right || ldright < cap)))
        return set_err(TCI_ERR_ARG, "left/right buffers too small");
    if (m * n > ((int64_t)1 << 31)) return set_err(TCI_ERR_ARG, "matrix too large");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t mn = (size_t)(m * n);
    double *dA = nullptr, *dLU = nullptr;
    int32_t *dperm = nullptr;
    tci_lu_job *dj = nullptr;
    HIPCHK(hipMallocAsync((void **)&dA, sizeof(double) * (mn ? mn : 1), st));
    HIPCHK(hipMallocAsync((void **)&dperm, sizeof(int32_t) * (size_t)(m + n + 1), st));
    HIPCHK(hipMallocAsync((void **)&dj, sizeof(tci_lu_job), st));
    if (mn) HIPCHK(hipMemcpy2DAsync(dA, sizeof(double) * m, A, sizeof(double) * lda, sizeof(double) * m, n,
                                    hipMemcpyHostToDevice, st));
    tci_lu_job j;
    memset(&j, 0, sizeof j);
    j.A = dA;
    j.rowperm = dperm;
    j.colperm = dperm + m;
    j.m = m;
    j.n = n;
    j.maxrank = maxrank;
    j.reltol = reltol;
    j.abstol = abstol;
    j.leftorth = leftorth ? 1 : 0;
    HIPCHK(hipMemcpyAsync(dj, &j, sizeof j, hipMemcpyHostToDevice, st));
    tci_status s = tci_rrlu_batched_f64(ctx, dj, 1);
    if (s != TCI_OK) return s;
    HIPCHK(hipMemcpyAsync(&j, dj, sizeof j, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int64_t p = j.npivot;
    if (j.status == TCI_OK && p > 0) {
        // unpacked L, U, then left and right (device), then host copies
        const size_t nl = (size_t)(m * p), nu = (size_t)(p * n);
        HIPCHK(hipMallocAsync((void **)&dLU, sizeof(double) * (2 * (nl + nu) + 1), st));
        double *dL = dLU, *dU = dLU + nl, *dleft = dU + nu, *dright = dleft + nl;
        HIPCHK(launch_lu_unpack(dA, m, n, p, j.leftorth, dL, m, dU, p, st));
        HIPCHK(launch_luci(m, n, p, j.leftorth, dperm, dperm + m, dL, dU, dleft, m, dright, p, st));
        HIPCHK(hipMemcpy2DAsync(left, sizeof(double) * ldleft, dleft, sizeof(double) * m, sizeof(double) * m, p,
                                hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpy2DAsync(right, sizeof(double) * ldright, dright, sizeof(double) * p, sizeof(double) * p, n,
                                hipMemcpyDeviceToHost, st));
    }
    std::vector<int32_t> perm((size_t)(m + n + 1));
    HIPCHK(hipMemcpyAsync(perm.data(), dperm, sizeof(int32_t) * (size_t)(m + n), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < m; ++i) rowperm[i] = perm[(size_t)i];
    for (int64_t q = 0; q < n; ++q) colperm[q] = perm[(size_t)(m + q)];
    (void)hipFreeAsync(dA, st);
    (void)hipFreeAsync(dperm, st);
    (void)hipFreeAsync(dj, st);
    if (dLU) (void)hipFreeAsync(dLU, st);
    *npivot = p;
    *error = j.error;
    if (j.status == TCI_ERR_NAN_L) return set_err(TCI_ERR_NAN_L, "lu.L contains NaNs");
    if (j.status == TCI_ERR_NAN_U) return set_err(TCI_ERR_NAN_U, "lu.U contains NaNs");
    return TCI_OK;
}

tci_status tci_gemm_f64(tci_ctx *ctx, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                        const double *B, int64_t ldb, double *Cm, int64_t ldc)
{
    if (!ctx || M < 0 || N < 0 || K < 0 || lda < (M > 0 ? M : 1) || ldb < (K > 0 ? K : 1) || ldc < (M > 0 ? M : 1))
        return set_err(TCI_ERR_ARG, "bad arguments");
    if (M * N == 0) return TCI_OK;
    if ((K > 0 && (!A || !B)) || !Cm) return set_err(TCI_ERR_ARG, "NULL buffer");
    if (M * K > ((int64_t)1 << 31) || K * N > ((int64_t)1 << 31) || M * N > ((int64_t)1 << 31))
        return set_err(TCI_ERR_ARG, "matrix too large");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    double *d = nullptr;
    const size_t na = (size_t)(M * K), nb = (size_t)(K * N), nc = (size_t)(M * N);
    HIPCHK(hipMallocAsync((void **)&d, sizeof(double) * (na + nb + nc + 1), st));
    hipError_t e = hipSuccess;
    if (na) e = hipMemcpy2DAsync(d, sizeof(double) * M, A, sizeof(double) * lda, sizeof(double) * M, K,
                                 hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nb)
        e = hipMemcpy2DAsync(d + na, sizeof(double) * K, B, sizeof(double) * ldb, sizeof(double) * K, N,
                             hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_gemm(M, N, K, d, M, d + na, K > 0 ? K : 1, d + na + nb, M, st);
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(Cm, sizeof(double) * ldc, d + na + nb, sizeof(double) * M, sizeof(double) * M, N,
                             hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(d, st);
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "gemm: %s", hipGetErrorString(e));
    return TCI_OK;
}

// ---------------------------------------------------------- tensor trains
// A tensor train resident on the device for compress! and the TT →
// TensorCI2 sweeps: each site keeps the slot (and capacity) it had on
// upload, since no sweep step makes a site larger (a new bond is at most the
// old one).  Work buffers hold the unpacked L and U, the factor that is
// carried to the neighbour and the product, each sized for the largest site.
struct TtDevice {
    tci_ctx *ctx = nullptr;
    int L = 0;
    std::vector<int64_t> links;
    std::vector<int32_t> d;
    std::vector<int64_t> off;   // slot of each site in `cores` (doubles)
    int64_t maxsite = 0, maxside = 0;
    uint8_t *arena = nullptr;
    double *cores = nullptr, *wl = nullptr, *wu = nullptr, *wc = nullptr, *wp = nullptr, *diag = nullptr;
    int32_t *perm = nullptr;
    tci_lu_job *job = nullptr;
    ~TtDevice()
    {
        if (arena) (void)hipFreeAsync(arena, ctx->stream);
    }
};

static int64_t tt_site_size(const TtDevice &T, int b) { return T.links[(size_t)b] * T.d[(size_t)b] * T.links[(size_t)b + 1]; }

static tci_status tt_check(int32_t L, const int64_t *linkdims, const int32_t *localdims, const double *cores)
{
    if (L < 1 || !linkdims || !localdims || !cores) return set_err(TCI_ERR_ARG, "bad arguments");
    for (int32_t b = 0; b <= L; ++b)
        if (linkdims[b] < 1 || linkdims[b] > INT32_MAX / 2) return set_err(TCI_ERR_ARG, "bond %d out of range", b);
    for (int32_t b = 0; b < L; ++b) {
        if (localdims[b] < 1) return set_err(TCI_ERR_ARG, "local dimension %d out of range", b + 1);
        if (linkdims[b] * localdims[b] > ((int64_t)1 << 31) / linkdims[b + 1] ||
            localdims[b] * linkdims[b + 1] > ((int64_t)1 << 31) / linkdims[b])
            return set_err(TCI_ERR_ARG, "site %d too large", b + 1);
    }
    return TCI_OK;
}

static tci_status tt_upload(tci_ctx *ctx, TtDevice &T, int32_t L, const int64_t *linkdims, const int32_t *localdims,
                            const double *cores)
{
    T.ctx = ctx;
    T.L = L;
    T.links.assign(linkdims, linkdims + L + 1);
    T.d.assign(localdims, localdims + L);
    T.off.resize((size_t)L);
    int64_t tot = 0;
    for (int b = 0; b < L; ++b) {
        T.off[(size_t)b] = tot;
        const int64_t sz = tt_site_size(T, b);
        tot += sz;
        T.maxsite = sz > T.maxsite ? sz : T.maxsite;
        const int64_t s1 = T.links[(size_t)b] * T.d[(size_t)b], s2 = T.d[(size_t)b] * T.links[(size_t)b + 1];
        const int64_t side = (s1 > s2 ? s1 : s2) + T.links[(size_t)b] + T.links[(size_t)b + 1];
        T.maxside = side > T.maxside ? side : T.maxside;
    }
    // cores | L | U | carried factor | product | diagonal | perm | job
    const size_t b_cores = lay_align(sizeof(double) * (size_t)tot);
    const size_t b_w = lay_align(sizeof(double) * (size_t)T.maxsite);
    const size_t b_diag = lay_align(sizeof(double) * (size_t)T.maxside);
    const size_t b_perm = lay_align(sizeof(int32_t) * (size_t)T.maxside);
    const size_t bytes = b_cores + 4 * b_w + b_diag + b_perm + lay_align(sizeof(tci_lu_job));
    HIPCHK(hipMallocAsync((void **)&T.arena, bytes, ctx->stream));
    uint8_t *p = T.arena;
    T.cores = (double *)p;
    p += b_cores;
    T.wl = (double *)p;
    p += b_w;
    T.wu = (double *)p;
    p += b_w;
    T.wc = (double *)p;
    p += b_w;
    T.wp = (double *)p;
    p += b_w;
    T.diag = (double *)p;
    p += b_diag;
    T.perm = (int32_t *)p;
    p += b_perm;
    T.job = (tci_lu_job *)p;
    HIPCHK(hipMemcpyAsync(T.cores, cores, sizeof(double) * (size_t)tot, hipMemcpyHostToDevice, ctx->stream));
    return TCI_OK;
}

// the cores back to back with the current bonds
static tci_status tt_download(TtDevice &T, int64_t *linkdims, double *cores)
{
    int64_t o = 0;
    for (int b = 0; b < T.L; ++b) {
        const int64_t sz = tt_site_size(T, b);
        HIPCHK(hipMemcpyAsync(cores + o, T.cores + T.off[(size_t)b], sizeof(double) * (size_t)sz,
                              hipMemcpyDeviceToHost, T.ctx->stream));
        o += sz;
    }
    HIPCHK(hipStreamSynchronize(T.ctx->stream));
    for (int b = 0; b <= T.L; ++b) linkdims[b] = T.links[(size_t)b];
    return TCI_OK;
}

// One sweep (tensortrain.jl:147-183 in either direction, conversion.jl:95-147):
// per bond, the rrLU of the grouped site matrix, its left/right factors
// (method 0 rrLU with permutations, 1 MatrixLUCI), the factored site
// replaced by one of them and the other multiplied into the neighbour.
// The whole sweep in one workgroup (launch_tt_sweep) for trains whose every
// bond matrix holds at most kTtFusedElems entries: there the host-driven
// sweep spends its time in per-bond round trips (about 0.3 ms each), not in
// the factorizations.  Same operations per entry, so the same bits.
constexpr int64_t kTtFusedElems = (int64_t)256 * 256;
constexpr int64_t kTtStageCap = 48 * 1024 / 16;   // doubles per LDS staging vector
constexpr int64_t kTtMatCap = 96 * 1024 / 8;       // doubles of LDS for a bond matrix

static tci_status tt_sweep_fused(TtDevice &T, int32_t method, int32_t forward, int64_t maxrank, double reltol,
                                 double abstol, int64_t R, int32_t *pivots, double *pivoterrors)
{
    tci_ctx *ctx = T.ctx;
    hipStream_t st = ctx->stream;
    const int L = T.L;
    const size_t b_links = lay_align(sizeof(int64_t) * (size_t)(L + 1)), b_d = lay_align(sizeof(int32_t) * (size_t)L);
    const size_t b_off = lay_align(sizeof(int64_t) * (size_t)L);
    const size_t b_piv = lay_align(sizeof(int32_t) * (size_t)(L - 1) * 2 * (size_t)R);
    const size_t b_pe = lay_align(sizeof(double) * (size_t)(R + 1)), b_st = lay_align(sizeof(int32_t) * 2);
    const size_t b_stage = lay_align(sizeof(double) * 2 * (size_t)T.maxside);
    uint8_t *d = nullptr;
    HIPCHK(hipMallocAsync((void **)&d, b_links + b_d + b_off + b_piv + b_pe + b_st + b_stage, st));
    TtSweepArgs a;
    memset(&a, 0, sizeof a);
    a.L = L;
    a.method = method;
    a.forward = forward;
    a.stage_cap = (int32_t)(T.maxside < kTtStageCap ? T.maxside : kTtStageCap);
    a.mat_cap = (int32_t)(T.maxsite < kTtMatCap ? T.maxsite : kTtMatCap);
    a.maxrank = maxrank;
    a.R = R;
    a.reltol = reltol;
    a.abstol = abstol;
    a.links = (int64_t *)d;
    a.d = (const int32_t *)(d + b_links);
    a.off = (const int64_t *)(d + b_links + b_d);
    a.piv = (int32_t *)(d + b_links + b_d + b_off);
    a.pe = (double *)(d + b_links + b_d + b_off + b_piv);
    a.status = (int32_t *)(d + b_links + b_d + b_off + b_piv + b_pe);
    a.gstage = (double *)(d + b_links + b_d + b_off + b_piv + b_pe + b_st);
    a.cores = T.cores;
    a.wl = T.wl;
    a.wu = T.wu;
    a.wc = T.wc;
    a.wp = T.wp;
    a.perm = T.perm;
    std::vector<int32_t> hst(2, 0), hpiv((size_t)(L - 1) * 2 * (size_t)R);
    std::vector<double> hpe((size_t)(R + 1), 0.0);
    hipError_t e = hipMemcpyAsync(a.links, T.links.data(), sizeof(int64_t) * (size_t)(L + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync((void *)a.d, T.d.data(), sizeof(int32_t) * (size_t)L, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync((void *)a.off, T.off.data(), sizeof(int64_t) * (size_t)L, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(a.pe, hpe.data(), sizeof(double) * (size_t)(R + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(a.status, hst.data(), sizeof(int32_t) * 2, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_tt_sweep(a, st);
    std::vector<int64_t> links((size_t)(L + 1));
    if (e == hipSuccess) e = hipMemcpyAsync(links.data(), a.links, sizeof(int64_t) * (size_t)(L + 1), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hpiv.data(), a.piv, sizeof(int32_t) * hpiv.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hpe.data(), a.pe, sizeof(double) * hpe.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hst.data(), a.status, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(d, st);
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "fused tt sweep: %s", hipGetErrorString(e));
    if (hst[0] == TCI_ERR_NAN_L) return set_err(TCI_ERR_NAN_L, "lu.L contains NaNs");
    if (hst[0] == TCI_ERR_NAN_U) return set_err(TCI_ERR_NAN_U, "lu.U contains NaNs");
    if (hst[0] != TCI_OK) return set_err((tci_status)hst[0], "rrlu failed at bond %d", hst[1] + 1);
    T.links = links;
    if (pivoterrors)
        for (int64_t k = 0; k <= R; ++k) pivoterrors[k] = hpe[(size_t)k];
    if (pivots) {
        int64_t q = 0;
        for (int b = 0; b < L - 1; ++b) {
            const int64_t p = links[(size_t)b + 1];
            const int32_t *src = hpiv.data() + (size_t)b * 2 * (size_t)R;
            for (int64_t k = 0; k < 2 * p; ++k) pivots[q++] = src[k];
        }
    }
    return TCI_OK;
}

static tci_status tt_sweep(TtDevice &T, int32_t method, int32_t forward, int64_t maxrank, double reltol,
                           double abstol, int32_t *pivots, double *pivoterrors)
{
    tci_ctx *ctx = T.ctx;
    hipStream_t st = ctx->stream;
    const int L = T.L;
    int64_t R = 0;
    for (int b = 1; b < L; ++b) R = T.links[(size_t)b] > R ? T.links[(size_t)b] : R;
    {
        int64_t big = 0;
        for (int b = 0; b + 1 < L; ++b) {
            const int64_t mm = forward ? T.links[(size_t)b] * T.d[(size_t)b] : T.links[(size_t)b + 1];
            const int64_t nn = forward ? T.links[(size_t)b + 1] : T.d[(size_t)b + 1] * T.links[(size_t)b + 2];
            big = mm * nn > big ? mm * nn : big;
        }
        const char *ev = getenv("TCI_TT_FUSED");
        const bool fused = ev ? atoi(ev) != 0 : big <= kTtFusedElems;
        if (fused && L > 1 && R > 0) return tt_sweep_fused(T, method, forward, maxrank, reltol, abstol, R, pivots, pivoterrors);
    }
    if (pivoterrors)
        for (int64_t k = 0; k <= R; ++k) pivoterrors[k] = 0.0;
    std::vector<std::vector<int32_t>> piv((size_t)(L > 1 ? L - 1 : 0));
    std::vector<int32_t> hperm((size_t)T.maxside);
    std::vector<double> hdiag((size_t)T.maxside);
    for (int i = 0; i < L - 1; ++i) {
        const int b = forward ? i : L - 2 - i;
        const int s = forward ? b : b + 1;
        const int64_t m = forward ? T.links[(size_t)b] * T.d[(size_t)b] : T.links[(size_t)b + 1];
        const int64_t n = forward ? T.links[(size_t)b + 1] : T.d[(size_t)b + 1] * T.links[(size_t)b + 2];
        double *A = T.cores + T.off[(size_t)s];
        tci_lu_job j;
        memset(&j, 0, sizeof j);
        j.A = A;
        j.rowperm = T.perm;
        j.colperm = T.perm + m;
        j.m = m;
        j.n = n;
        j.maxrank = maxrank;
        j.reltol = reltol;
        j.abstol = abstol;
        j.leftorth = forward ? 1 : 0;
        HIPCHK(hipMemcpyAsync(T.job, &j, sizeof j, hipMemcpyHostToDevice, st));
        tci_status ls = tci_rrlu_batched_f64(ctx, T.job, 1);
        if (ls != TCI_OK) return ls;
        // the pivots are the diagonal of the factored matrix
        const int64_t mn = m < n ? m : n;
        HIPCHK(hipMemcpyAsync(&j, T.job, sizeof j, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpy2DAsync(hdiag.data(), sizeof(double), A, sizeof(double) * (size_t)(m + 1), sizeof(double),
                                (size_t)mn, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hperm.data(), T.perm, sizeof(int32_t) * (size_t)(m + n), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (j.status == TCI_ERR_NAN_L) return set_err(TCI_ERR_NAN_L, "lu.L contains NaNs");
        if (j.status == TCI_ERR_NAN_U) return set_err(TCI_ERR_NAN_U, "lu.U contains NaNs");
        if (j.status != TCI_OK) return set_err((tci_status)j.status, "rrlu failed at bond %d", b + 1);
        const int64_t p = j.npivot;
        HIPCHK(launch_lu_unpack(A, m, n, p, j.leftorth, T.wl, m, T.wu, p, st));
        // forward: the site takes left (m×p), right goes to wc; backward:
        // the site takes right (p×n), left goes to wc
        double *lft = forward ? A : T.wc, *rgt = forward ? T.wc : A;
        if (method == 0)
            HIPCHK(launch_lu_permute(m, n, p, T.perm, T.perm + m, T.wl, T.wu, lft, m, rgt, p, st));
        else
            HIPCHK(launch_luci(m, n, p, j.leftorth, T.perm, T.perm + m, T.wl, T.wu, lft, m, rgt, p, st));
        if (forward) {
            // site b+1 ← right · (χ_{b+1} × d_{b+1} χ_{b+2})
            const int64_t nn = T.d[(size_t)b + 1] * T.links[(size_t)b + 2];
            double *nx = T.cores + T.off[(size_t)b + 1];
            HIPCHK(launch_gemm(p, nn, n, T.wc, p, nx, n, T.wp, p, st));
            HIPCHK(hipMemcpyAsync(nx, T.wp, sizeof(double) * (size_t)(p * nn), hipMemcpyDeviceToDevice, st));
        } else {
            // site b ← (χ_b d_b × χ_{b+1}) · left
            const int64_t mm = T.links[(size_t)b] * T.d[(size_t)b];
            double *pv = T.cores + T.off[(size_t)b];
            HIPCHK(launch_gemm(mm, p, m, pv, mm, T.wc, m, T.wp, mm, st));
            HIPCHK(hipMemcpyAsync(pv, T.wp, sizeof(double) * (size_t)(mm * p), hipMemcpyDeviceToDevice, st));
        }
        T.links[(size_t)b + 1] = p;
        piv[(size_t)b].assign(hperm.data(), hperm.data() + p);
        piv[(size_t)b].insert(piv[(size_t)b].end(), hperm.data() + m, hperm.data() + m + p);
        if (pivoterrors) {
            for (int64_t k = 0; k < p; ++k) pivoterrors[k] = nan_max(pivoterrors[k], fabs(hdiag[(size_t)k]));
            pivoterrors[p] = nan_max(pivoterrors[p], j.error);
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    if (pivots) {
        int64_t q = 0;
        for (const auto &v : piv)
            for (int32_t x : v) pivots[q++] = x;
    }
    return TCI_OK;
}

tci_status tci_tt_sweep_f64(tci_ctx *ctx, int32_t L, int64_t *linkdims, const int32_t *localdims, double *cores,
                            int32_t method, int32_t forward, int64_t maxrank, double reltol, double abstol,
                            int32_t *pivots, double *pivoterrors)
{
    if (!ctx) return set_err(TCI_ERR_ARG, "bad arguments");
    tci_status s = tt_check(L, linkdims, localdims, cores);
    if (s != TCI_OK) return s;
    if (method != TCI_FACTOR_LU && method != TCI_FACTOR_CI)
        return set_err(TCI_ERR_ARG, "method must be TCI_FACTOR_LU or TCI_FACTOR_CI");
    if (maxrank < 1) return set_err(TCI_ERR_ARG, "maxrank must be >= 1");
    HIPCHK(hipSetDevice(ctx->device));
    TtDevice T;
    s = tt_upload(ctx, T, L, linkdims, localdims, cores);
    if (s == TCI_OK) s = tt_sweep(T, method, forward ? 1 : 0, maxrank, reltol, abstol, pivots, pivoterrors);
    if (s == TCI_OK) s = tt_download(T, linkdims, cores);
    return s;
}

tci_status tci_tt_compress_f64(tci_ctx *ctx, int32_t L, int64_t *linkdims, const int32_t *localdims, double *cores,
                               int32_t method, double tolerance, int64_t maxbonddim, int32_t normalizeerror)
{
    if (!ctx) return set_err(TCI_ERR_ARG, "bad arguments");
    tci_status s = tt_check(L, linkdims, localdims, cores);
    if (s != TCI_OK) return s;
    if (method == TCI_FACTOR_SVD)
        return set_err(TCI_ERR_ARG, "compress! with :SVD is not provided by this library (use :LU or :CI)");
    if (method != TCI_FACTOR_LU && method != TCI_FACTOR_CI) return set_err(TCI_ERR_GENERIC, "Not implemented yet.");
    if (maxbonddim < 1) return set_err(TCI_ERR_ARG, "maxbonddim must be >= 1");
    HIPCHK(hipSetDevice(ctx->device));
    TtDevice T;
    s = tt_upload(ctx, T, L, linkdims, localdims, cores);
    // left to right with no truncation (tolerance 0.0: reltol = abstol = 0),
    // then right to left with the caller's truncation (tensortrain.jl:98-104)
    if (s == TCI_OK) s = tt_sweep(T, method, 1, TCI_UNLIMITED, 0.0, 0.0, nullptr, nullptr);
    if (s == TCI_OK)
        s = tt_sweep(T, method, 0, maxbonddim, normalizeerror ? tolerance : 1e-14, normalizeerror ? 0.0 : tolerance,
                     nullptr, nullptr);
    if (s == TCI_OK) s = tt_download(T, linkdims, cores);
    return s;
}

// ------------------------------------------------------------ Pi tensors
tci_status tci_pi_eval_f64(tci_ctx *ctx, const tci_functor *f, int32_t L, const int32_t *localdims,
                           const uint8_t *left, int64_t nI, int32_t nl, const uint8_t *right,
                           int64_t nJ, int32_t nr, double *out, double *maxabs_inout)
{
    if (!ctx || !f || !localdims || L < 1 || nl < 0 || nr < 0 || nl + nr > L || nI < 0 || nJ < 0)
        return set_err(TCI_ERR_ARG, "bad arguments");
    const int M = L - nl - nr;
    const int64_t lim = (int64_t)1 << 34;   // elements of one evaluation (128 GiB of output)
    int64_t nc = 1;
    for (int t = 0; t < M; ++t) {
        const int64_t d = localdims[nl + t];
        if (d < 1 || d > 255) return set_err(TCI_ERR_ARG, "local dimensions must be in 1..255");
        nc *= d;
        if (nc > lim) return set_err(TCI_ERR_ARG, "Pi tensor too large");
    }
    if (nI > lim || nJ > lim || (nI > 0 && nc > lim / nI)) return set_err(TCI_ERR_ARG, "Pi tensor too large");
    const int64_t nrow = nI * nc;
    if (nJ > 0 && nrow > lim / nJ) return set_err(TCI_ERR_ARG, "Pi tensor too large");
    const int64_t tot = nrow * nJ;
    if (tot == 0) return TCI_OK;   // empty Pi (batcheval.jl:40-42)
    if (!out || (nI > 0 && nl > 0 && !left) || (nJ > 0 && nr > 0 && !right))
        return set_err(TCI_ERR_ARG, "NULL buffer");
    // the functor and local dims are checked for every L, so no parameter
    // read of the eval kernel goes past its block
    std::string msg;
    const int fst = check_functor(*f, L, localdims, &msg);
    if (fst != TCI_OK) return set_err((tci_status)fst, "%s", msg.c_str());
    for (int64_t i = 0; i < nI * nl; ++i)
        if (left[i] < 1 || left[i] > localdims[i % nl]) return set_err(TCI_ERR_ARG, "left index out of range");
    for (int64_t j = 0; j < nJ * nr; ++j)
        if (right[j] < 1 || right[j] > localdims[L - nr + j % nr]) return set_err(TCI_ERR_ARG, "right index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int blocks = eval_blocks(tot);
    size_t need = 0;
    auto a = [&](size_t b) { size_t at = need; need = lay_align(need + (b ? b : 1)); return at; };
    size_t o_out = a(sizeof(double) * (size_t)tot);
    size_t o_bm = a(sizeof(double) * (size_t)(blocks + 1));
    size_t o_left = a((size_t)(nI * nl));
    size_t o_right = a((size_t)(nJ * nr));
    size_t o_rows = a((size_t)(nrow * (nl + M)));
    size_t o_cd = a(sizeof(int32_t) * (size_t)(M + 1));
    size_t o_par = a(sizeof(double) * (size_t)f->nparams);
    size_t o_ipar = a(sizeof(int32_t) * (size_t)f->niparams);
    uint8_t *d = nullptr;
    HIPCHK(hipMallocAsync((void **)&d, need, st));
    if (nI * nl) HIPCHK(hipMemcpyAsync(d + o_left, left, (size_t)(nI * nl), hipMemcpyHostToDevice, st));
    if (nJ * nr) HIPCHK(hipMemcpyAsync(d + o_right, right, (size_t)(nJ * nr), hipMemcpyHostToDevice, st));
    if (M) HIPCHK(hipMemcpyAsync(d + o_cd, localdims + nl, sizeof(int32_t) * M, hipMemcpyHostToDevice, st));
    if (f->nparams) HIPCHK(hipMemcpyAsync(d + o_par, f->params, sizeof(double) * f->nparams, hipMemcpyHostToDevice, st));
    if (f->niparams) HIPCHK(hipMemcpyAsync(d + o_ipar, f->iparams, sizeof(int32_t) * f->niparams, hipMemcpyHostToDevice, st));
    const uint8_t *rows = d + o_left;
    int rstride = nl;
    if (M > 0) {
        HIPCHK(launch_kron_center(d + o_left, nI, nl, (const int32_t *)(d + o_cd), M, nc, d + o_rows,
                                  nl + M, st));
        rows = d + o_rows;
        rstride = nl + M;
    }
    HIPCHK(launch_eval(f->kind, (const double *)(d + o_par), (const int32_t *)(d + o_ipar), L, rows,
                       nrow, rstride, nl + M, d + o_right, nJ, nr, (double *)(d + o_out),
                       (double *)(d + o_bm), st));
    HIPCHK(hipMemcpyAsync(out, d + o_out, sizeof(double) * (size_t)tot, hipMemcpyDeviceToHost, st));
    std::vector<double> bm((size_t)blocks);
    HIPCHK(hipMemcpyAsync(bm.data(), d + o_bm, sizeof(double) * blocks, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFreeAsync(d, st);
    if (maxabs_inout) {
        double mx = fabs(*maxabs_inout);
        for (double v : bm) mx = nan_max(mx, v);
        *maxabs_inout = mx;
    }
    return TCI_OK;
}

// ------------------------------------------------------------- solves
void tci_options_default(tci_options *o)
{
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->tolerance = 1e-8;
    o->maxbonddim = TCI_UNLIMITED;
    o->maxiter = 20;
    o->sweepstrategy = 0;
    o->normalizeerror = 1;
    o->ncheckhistory = 3;
    o->maxnglobalpivot = 5;
    o->nsearchglobalpivot = 5;
    o->tolmarginglobalsearch = 10.0;
    o->strictlynested = 0;
    o->checkconvglobalpivot = 1;
    o->seed = 0;
    o->chi_cap = 0;
}

tci_status tci_capacity_plan(const tci_problem *p, int64_t *chi_used, int64_t *chi_limit)
{
    if (!p || !chi_used || !chi_limit) return set_err(TCI_ERR_ARG, "bad arguments");
    SolveCaps c;
    std::string msg;
    int s = check_problem(*p, c, &msg);
    if (s != TCI_OK) return set_err((tci_status)s, "%s", msg.c_str());
    int64_t used = p->opts.chi_cap > 0 ? p->opts.chi_cap : default_chi_cap(*p);
    if (p->state && largest_state_set(*p) > used) used = largest_state_set(*p);
    *chi_used = used < 1 ? 1 : used;
    int64_t lim = rank_bound(p->L, p->localdims);
    if (p->opts.maxbonddim < lim) lim = p->opts.maxbonddim;
    *chi_limit = lim < 1 ? 1 : lim;
    return TCI_OK;
}

static void serve_block_put(tci_server *s, tci_ctx *ctx, const ServeBlock &k, int cls, bool device);

static void batch_release(tci_batch *b)
{
    if (!b) return;
    if (b->served) {
        if (b->dblk.ptr) serve_block_put(b->server, b->ctx, b->dblk, b->dcls, true);
        if (b->hblk.ptr) serve_block_put(b->server, b->ctx, b->hblk, b->hcls, false);
    } else if (b->ctx) {
        pool_put(b->ctx, b->d_desc, b->pool_bytes[3]);
        pool_put(b->ctx, b->d_in, b->pool_bytes[0]);
        pool_put(b->ctx, b->d_state, b->pool_bytes[1]);
        pool_put(b->ctx, b->d_rep, b->pool_bytes[2]);
        b->ctx->used -= b->bytes;
    } else {
        if (b->d_desc) (void)hipFree(b->d_desc);
        if (b->d_in) (void)hipFree(b->d_in);
        if (b->d_state) (void)hipFree(b->d_state);
        if (b->d_rep) (void)hipFree(b->d_rep);
    }
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
}

// the launch-independent part of a submit: the problems checked and laid
// out, the regions taken from the pool and the inputs and descriptors
// uploaded on st; *sh gets what the LDS plan and the kernel choice need
struct BatchShape {
    int max_mr = 0, max_tab = 0, fkind = 0;
    int64_t max_sets = 0, max_lut = 0;
};

static tci_status batch_prepare(tci_ctx *ctx, int32_t count, const tci_problem *problems, hipStream_t st,
                                tci_batch **out, BatchShape *sh)
{
    if (!ctx || !out || count < 1 || !problems) return set_err(TCI_ERR_ARG, "bad arguments");
    *out = nullptr;
    std::vector<SolveCaps> caps((size_t)count);
    std::vector<size_t> ib((size_t)count), sb((size_t)count), rb((size_t)count);
    size_t tin = 0, tstate = 0, trep = 0;
    int max_mr = 0, max_tab = 0;
    int64_t max_sets = 0, max_lut = 0;
    // one functor-specialized kernel when every problem has that functor;
    // the generic kernel otherwise, and for any :rook search or memo table
    // (the specialized kernels carry only the :full path without a table)
    int fkind = problems[0].f.kind;
    for (int i = 0; i < count; ++i) {
        if (problems[i].f.kind != fkind || problems[i].opts.pivotsearch || problems[i].opts.cache_log2) fkind = 0;
        std::string msg;
        int s = check_problem(problems[i], caps[(size_t)i], &msg);
        if (s != TCI_OK) return set_err((tci_status)s, "problem %d: %s", i, msg.c_str());
        layout_solve(caps[(size_t)i], nullptr, nullptr, nullptr, nullptr, &ib[(size_t)i], &sb[(size_t)i], &rb[(size_t)i]);
        tin += ib[(size_t)i];
        tstate += sb[(size_t)i];
        trep += rb[(size_t)i];
        if (caps[(size_t)i].mrcap > max_mr) max_mr = caps[(size_t)i].mrcap;
        const int tab = caps[(size_t)i].mrcap * caps[(size_t)i].L;
        if (tab > max_tab) max_tab = tab;
        const SolveCaps &cp = caps[(size_t)i];
        const int64_t sbytes = ((((int64_t)NSETKIND * cp.L * cp.setcap * cp.L) + 15) & ~(int64_t)15) +
                               4 * (int64_t)NSETKIND * cp.L;
        if (sbytes > max_sets) max_sets = sbytes;
        if (problems[i].f.kind == TCI_F_LORENTZ_INT) {
            int64_t smax = 2;
            for (int t = 0; t < problems[i].L; ++t)
                smax += (int64_t)problems[i].localdims[t] * problems[i].localdims[t];
            if (smax > max_lut) max_lut = smax;
        }
    }
    const size_t total = tin + tstate + trep + sizeof(SolveDesc) * (size_t)count;
    if (ctx->used + total > ctx->budget)
        return set_err(TCI_ERR_OOM, "batch needs %zu bytes, HBM budget has %llu left", total,
                       (unsigned long long)(ctx->budget - ctx->used));
    HIPCHK(hipSetDevice(ctx->device));
    tci_batch *b = new tci_batch();
    b->ctx = nullptr;
    b->count = count;
    auto fail = [&](hipError_t e, const char *what) {
        batch_release(b);
        return set_err(e == hipErrorOutOfMemory ? TCI_ERR_OOM : TCI_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    };
    hipError_t e;
    b->ctx = ctx;
    b->bytes = total;
    ctx->used += total;
    b->in_bytes = tin;
    b->state_bytes = tstate;
    b->desc_bytes = sizeof(SolveDesc) * (size_t)count;
    if ((e = pool_get(ctx, tin, (void **)&b->d_in, &b->pool_bytes[0])) != hipSuccess) return fail(e, "hipMalloc(inputs)");
    if ((e = pool_get(ctx, tstate, (void **)&b->d_state, &b->pool_bytes[1])) != hipSuccess)
        return fail(e, "hipMalloc(state)");
    if ((e = pool_get(ctx, trep, (void **)&b->d_rep, &b->pool_bytes[2])) != hipSuccess) return fail(e, "hipMalloc(report)");
    if ((e = pool_get(ctx, b->desc_bytes, (void **)&b->d_desc, &b->pool_bytes[3])) != hipSuccess)
        return fail(e, "hipMalloc(desc)");
    b->rep_bytes = trep;
    b->desc.resize((size_t)count);
    b->rep_off.resize((size_t)count);
    b->h_in.assign(tin, 0);
    uint8_t *himg = b->h_in.data();
    size_t oin = 0, ost = 0, orp = 0;
    for (int i = 0; i < count; ++i) {
        SolveDesc &D = b->desc[(size_t)i];
        memset(&D, 0, sizeof D);
        size_t x, y, z;
        layout_solve(caps[(size_t)i], &D, b->d_in + oin, b->d_state + ost, b->d_rep + orp, &x, &y, &z);
        fill_desc_scalars(problems[i], D);
        pack_inputs(problems[i], caps[(size_t)i], D, himg + oin, b->d_in + oin);
        b->rep_off[(size_t)i] = orp;
        oin += ib[(size_t)i];
        ost += sb[(size_t)i];
        orp += rb[(size_t)i];
    }
    b->stream = st;
    if ((e = hipMemcpyAsync(b->d_in, himg, tin, hipMemcpyHostToDevice, st)) != hipSuccess) return fail(e, "upload inputs");
    if ((e = hipMemcpyAsync(b->d_desc, b->desc.data(), sizeof(SolveDesc) * (size_t)count, hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail(e, "upload descriptors");
    sh->max_mr = max_mr;
    sh->max_tab = max_tab;
    sh->max_sets = max_sets;
    sh->max_lut = max_lut;
    sh->fkind = fkind;
    *out = b;
    return TCI_OK;
}

tci_status tci_solve_submit(tci_ctx *ctx, int32_t count, const tci_problem *problems, tci_batch **out)
{
    if (!ctx || !out || count < 1 || !problems) return set_err(TCI_ERR_ARG, "bad arguments");
    *out = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx);
    BatchShape sh;
    tci_batch *b = nullptr;
    tci_status s = batch_prepare(ctx, count, problems, st, &b, &sh);
    if (s != TCI_OK) return s;
    auto fail = [&](hipError_t e, const char *what) {
        batch_release(b);
        return set_err(e == hipErrorOutOfMemory ? TCI_ERR_OOM : TCI_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return fail(e, "hipEventCreate");
    if ((e = hipEventRecord(b->ev0, st)) != hipSuccess) return fail(e, "hipEventRecord");
    const int team = solve_team_for(ctx, sh.max_mr, count);
    const int stat = solve_static_lds(team, sh.fkind);
    if (stat < 0) return fail(hipErrorInvalidValue, "solve kernel attributes");
    const LdsPlan plan = make_plan(ctx, (int64_t)lds_for_team(ctx, team) - stat, sh.max_mr, sh.max_tab, sh.max_sets,
                                   sh.max_lut);
    if ((e = launch_solve(b->d_desc, count, team, sh.fkind, plan, st)) != hipSuccess) return fail(e, "solve launch");
    if ((e = hipEventRecord(b->ev1, st)) != hipSuccess) return fail(e, "hipEventRecord");
    *out = b;
    return TCI_OK;
}

// ---- the solve team's autotuner (tci_ctx_autotune_solve)
static void solve_tune_put(tci_ctx *ctx, int32_t key, int32_t count, int32_t team)
{
    auto &v = ctx->solve_tune;
    for (auto &e : v)
        if (e.mrcap == key && e.count == count) {
            e.team = team;
            return;
        }
    v.push_back({key, count, team});
    std::sort(v.begin(), v.end());
}

tci_status tci_ctx_autotune_solve(tci_ctx *ctx, const tci_problem *sample, int32_t batch, float *ms4, int32_t *team_out)
{
    if (!ctx || !sample || batch < 1 || batch > 65536) return set_err(TCI_ERR_ARG, "bad arguments");
    SolveCaps c;
    std::string msg;
    const int chk = check_problem(*sample, c, &msg);
    if (chk != TCI_OK) return set_err((tci_status)chk, "sample: %s", msg.c_str());
    static const int kTeams[4] = {64, 256, 512, 1024};
    std::vector<tci_problem> ps((size_t)batch, *sample);
    const auto saved = ctx->solve_tune;
    float best = -1.0f;
    int bt = ctx->team;
    tci_status st = TCI_OK;
    for (int q = 0; q < 4 && st == TCI_OK; ++q) {
        ctx->solve_tune.assign(1, tci_ctx::SolveTune{0, 0, kTeams[q]});   // every batch on this team
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && st == TCI_OK; ++rep) {   // a warm-up launch, then the timed one
            tci_batch *b = nullptr;
            st = tci_solve_submit(ctx, batch, ps.data(), &b);
            if (st == TCI_OK) st = tci_batch_wait(b);
            if (st == TCI_OK) st = tci_batch_elapsed_ms(b, &ms);
            if (b) (void)tci_batch_free(b);
        }
        if (st != TCI_OK) break;
        if (ms4) ms4[q] = ms;
        if (best < 0.0f || ms < best) {
            best = ms;
            bt = kTeams[q];
        }
    }
    ctx->solve_tune = saved;
    if (st != TCI_OK) return st;
    solve_tune_put(ctx, c.mrcap, batch, bt);
    if (team_out) *team_out = bt;
    return TCI_OK;
}

tci_status tci_ctx_set_solve_tuning(tci_ctx *ctx, const int32_t *min_mrcap, const int32_t *min_count,
                                    const int32_t *team, int32_t n)
{
    if (!ctx || n < 0 || (n > 0 && (!min_mrcap || !min_count || !team))) return set_err(TCI_ERR_ARG, "bad arguments");
    for (int32_t q = 0; q < n; ++q)
        if (min_mrcap[q] < 0 || min_count[q] < 0 ||
            (team[q] != 64 && team[q] != 256 && team[q] != 512 && team[q] != 1024))
            return set_err(TCI_ERR_ARG, "entry %d: mrcap >= 0, count >= 0, team 64, 256, 512 or 1024", (int)q);
    ctx->solve_tune.clear();
    for (int32_t q = 0; q < n; ++q) solve_tune_put(ctx, min_mrcap[q], min_count[q], team[q]);
    return TCI_OK;
}

static tci_status serve_ensure(tci_server *s);
static tci_status served_query(tci_batch *b);

static tci_status fetch_reports(tci_batch *b)
{
    if (b->fetched) return TCI_OK;
    HIPCHK(hipSetDevice(b->ctx->device));
    b->h_rep.resize(b->rep_bytes);
    if (b->served) {
        // poll the done mark (relaunching the grid if it went idle); the
        // workgroup copied the report to pinned host memory before it
        for (;;) {
            const tci_status q = served_query(b);
            if (q == TCI_OK) break;
            if (q != TCI_ERR_PENDING) return q;
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
        std::atomic_thread_fence(std::memory_order_acquire);
        memcpy(b->h_rep.data(), b->h_rep_pin, b->rep_bytes);
        b->fetched = true;
        return TCI_OK;
    }
    HIPCHK(hipEventSynchronize(b->ev1));
    HIPCHK(ctx_download(b->ctx, b->h_rep.data(), b->d_rep, b->rep_bytes));
    b->fetched = true;
    return TCI_OK;
}

tci_status tci_batch_query(tci_batch *b)
{
    if (!b) return set_err(TCI_ERR_ARG, "batch is NULL");
    if (b->served) return served_query(b);
    hipError_t e = hipEventQuery(b->ev1);
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();   // not finished is not an error: clear it
        return TCI_ERR_PENDING;
    }
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "hipEventQuery: %s", hipGetErrorString(e));
    return TCI_OK;
}

tci_status tci_batch_wait(tci_batch *b)
{
    if (!b) return set_err(TCI_ERR_ARG, "batch is NULL");
    return fetch_reports(b);
}

// host view of solve i's report block
static bool report_view(tci_batch *b, int32_t i, SolveDesc &R)
{
    if (!b || i < 0 || i >= b->count) return false;
    if (fetch_reports(b) != TCI_OK) return false;
    R = b->desc[(size_t)i];
    const uint8_t *dev = b->d_rep + b->rep_off[(size_t)i];
    uint8_t *host = b->h_rep.data() + b->rep_off[(size_t)i];
    auto h = [&](const void *p) { return host + ((const uint8_t *)p - dev); };
    R.ints = (int32_t *)h(R.ints);
    R.scal = (double *)h(R.scal);
    R.ranks = (int64_t *)h(R.ranks);
    R.nglob = (int64_t *)h(R.nglob);
    R.errors = (double *)h(R.errors);
    R.pivoterrors = (double *)h(R.pivoterrors);
    R.bonderrors = (double *)h(R.bonderrors);
    R.setn = (int32_t *)h(R.setn);
    return true;
}

tci_status tci_batch_result(tci_batch *b, int32_t i, tci_result *out)
{
    SolveDesc R;
    if (!out || !report_view(b, i, R)) return set_err(TCI_ERR_ARG, "bad batch/index");
    out->status = R.ints[INT_DONE] ? R.ints[INT_STATUS] : TCI_ERR_DEVICE;
    out->niter = R.ints[INT_NITER] < R.maxiter_cap ? R.ints[INT_NITER] : R.maxiter_cap;
    out->L = R.L;
    out->npivoterrors = R.ints[INT_NPE];
    out->maxsamplevalue = R.scal[SCAL_MAXSAMPLE];
    for (int k = 0; k < 6; ++k) out->workstats[k] = R.scal[SCAL_WS_PI + k];
    return TCI_OK;
}

tci_status tci_batch_history(tci_batch *b, int32_t i, int64_t *ranks, double *errors, int64_t *ng)
{
    SolveDesc R;
    if (!report_view(b, i, R)) return set_err(TCI_ERR_ARG, "bad batch/index");
    int n = R.ints[INT_NITER] < R.maxiter_cap ? R.ints[INT_NITER] : R.maxiter_cap;
    for (int k = 0; k < n; ++k) {
        if (ranks) ranks[k] = R.ranks[k];
        if (errors) errors[k] = R.errors[k];
        if (ng) ng[k] = R.nglob[k];
    }
    return TCI_OK;
}

tci_status tci_batch_linkdims(tci_batch *b, int32_t i, int64_t *linkdims)
{
    SolveDesc R;
    if (!linkdims || !report_view(b, i, R)) return set_err(TCI_ERR_ARG, "bad batch/index");
    for (int s = 1; s < R.L; ++s) linkdims[s - 1] = R.setn[SET_I * R.L + s];
    return TCI_OK;
}

tci_status tci_batch_indexset(tci_batch *b, int32_t i, int32_t which, int32_t site, int32_t *count,
                              uint8_t *out, int64_t cap_bytes)
{
    SolveDesc R;
    if (!count || !report_view(b, i, R) || site < 1 || site > R.L || which < 0 || which > 3)
        return set_err(TCI_ERR_ARG, "bad arguments");
    const int L = R.L, b0 = site - 1, jside = which & 1;
    // which 2/3: the last history entry, in the slot the solve wrote last
    const int hist = R.ints[INT_HIST];
    const int kind = which < 2 ? (jside ? SET_J : SET_I)
                               : (hist == 2 ? (jside ? SET_HB_J : SET_HB_I) : (jside ? SET_HA_J : SET_HA_I));
    const int len = jside ? L - 1 - b0 : b0;
    const int n = (which >= 2 && hist == 0) ? 0 : R.setn[kind * L + b0];
    *count = n;
    if ((int64_t)n * len > cap_bytes || (!out && n * len > 0)) return set_err(TCI_ERR_ARG, "output buffer too small");
    if (n == 0 || len == 0) return TCI_OK;
    std::vector<uint8_t> raw((size_t)n * L);
    const uint8_t *src = R.sets + ((int64_t)(kind * L + b0) * R.setcap) * L;
    HIPCHK(ctx_download(b->ctx, raw.data(), src, raw.size()));
    for (int e = 0; e < n; ++e) memcpy(out + (size_t)e * len, raw.data() + (size_t)e * L, (size_t)len);
    return TCI_OK;
}

tci_status tci_batch_sitetensor(tci_batch *b, int32_t i, int32_t site, int64_t *dims3, double *out, int64_t cap)
{
    SolveDesc R;
    if (!dims3 || !report_view(b, i, R) || site < 1 || site > R.L) return set_err(TCI_ERR_ARG, "bad arguments");
    HIPCHK(ctx_download(b->ctx, dims3, R.Tdims + 3 * (site - 1), sizeof(int64_t) * 3));
    const int64_t n = dims3[0] * dims3[1] * dims3[2];
    if (n > cap || (!out && n > 0)) return set_err(TCI_ERR_ARG, "output buffer too small");
    if (n > 0) HIPCHK(ctx_download(b->ctx, out, R.T + R.tstride * (site - 1), sizeof(double) * n));
    return TCI_OK;
}

tci_status tci_batch_pivoterrors(tci_batch *b, int32_t i, double *out, int64_t cap)
{
    SolveDesc R;
    if (!report_view(b, i, R)) return set_err(TCI_ERR_ARG, "bad batch/index");
    const int n = R.ints[INT_NPE];
    if (n > cap || (!out && n > 0)) return set_err(TCI_ERR_ARG, "output buffer too small");
    for (int k = 0; k < n; ++k) out[k] = R.pivoterrors[k];
    return TCI_OK;
}

tci_status tci_batch_bonderrors(tci_batch *b, int32_t i, double *out)
{
    SolveDesc R;
    if (!out || !report_view(b, i, R)) return set_err(TCI_ERR_ARG, "bad batch/index");
    for (int k = 0; k < R.L - 1; ++k) out[k] = R.bonderrors[k];
    return TCI_OK;
}

tci_status tci_batch_elapsed_ms(tci_batch *b, float *ms)
{
    if (!b || !ms) return set_err(TCI_ERR_ARG, "bad arguments");
    if (b->served) return set_err(TCI_ERR_ARG, "a served request has no launch of its own to time");
    HIPCHK(hipEventSynchronize(b->ev1));
    HIPCHK(hipEventElapsedTime(ms, b->ev0, b->ev1));
    return TCI_OK;
}

tci_status tci_batch_profile(tci_batch *b, int32_t i, double *us8)
{
    SolveDesc R;
    if (!us8 || !report_view(b, i, R)) return set_err(TCI_ERR_ARG, "bad batch/index");
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->ctx->device));
    const double us_per_tick = khz > 0 ? 1e3 / (double)khz : 0.0;
    for (int k = 0; k < 8; ++k) us8[k] = R.scal[SCAL_TPH + k] * us_per_tick;
    return TCI_OK;
}

tci_status tci_batch_cachestats(tci_batch *b, int32_t i, int64_t *entries, int64_t *lookups, int64_t *hits)
{
    SolveDesc R;
    if (!report_view(b, i, R)) return set_err(TCI_ERR_ARG, "bad batch/index");
    if (entries) *entries = R.cache_log2 ? (int64_t)1 << R.cache_log2 : 0;
    if (lookups) *lookups = (int64_t)R.scal[SCAL_CACHE_LOOKUPS];
    if (hits) *hits = (int64_t)R.scal[SCAL_CACHE_HITS];
    return TCI_OK;
}

tci_status tci_batch_cachedata(tci_batch *b, int32_t i, int64_t cap, uint8_t *points, double *values,
                               int64_t *count)
{
    SolveDesc R;
    if (!count || !report_view(b, i, R) || cap < 0) return set_err(TCI_ERR_ARG, "bad arguments");
    *count = 0;
    if (!R.cache_log2) return TCI_OK;
    const int64_t ne = (int64_t)1 << R.cache_log2;
    const int L = R.L;
    std::vector<uint32_t> st((size_t)ne);
    std::vector<uint64_t> keys((size_t)ne * CACHE_LIMBS);
    std::vector<double> vals((size_t)ne);
    std::vector<int32_t> ld((size_t)L);
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(ctx_download(b->ctx, st.data(), R.cache_state, sizeof(uint32_t) * st.size()));
    HIPCHK(ctx_download(b->ctx, keys.data(), R.cache_keys, sizeof(uint64_t) * keys.size()));
    HIPCHK(ctx_download(b->ctx, vals.data(), R.cache_vals, sizeof(double) * vals.size()));
    HIPCHK(ctx_download(b->ctx, ld.data(), R.localdims, sizeof(int32_t) * ld.size()));
    int64_t n = 0;
    for (int64_t e = 0; e < ne; ++e) {
        if (st[(size_t)e] != CACHE_READY) continue;
        if (n < cap && points && values) {
            // decodecachekey (cachedfunction.jl:188-196): divrem by d_1, d_2, …
            uint64_t k[CACHE_LIMBS];
            for (int q = 0; q < CACHE_LIMBS; ++q) k[q] = keys[(size_t)(e * CACHE_LIMBS + q)];
            for (int s = 0; s < L; ++s) {
                const uint64_t d = (uint64_t)ld[(size_t)s];
                uint64_t rem = 0;
                for (int q = CACHE_LIMBS - 1; q >= 0; --q) {
                    // (rem·2^64 + k[q]) / d in 32-bit halves (rem < d < 2^8)
                    const uint64_t hi = (rem << 32) | (k[q] >> 32);
                    const uint64_t qh = hi / d, rh = hi % d;
                    const uint64_t lo = (rh << 32) | (k[q] & 0xffffffffULL);
                    const uint64_t ql = lo / d;
                    rem = lo % d;
                    k[q] = (qh << 32) | ql;
                }
                points[(size_t)(n * L + s)] = (uint8_t)(rem + 1);
            }
            values[(size_t)n] = vals[(size_t)e];
        }
        ++n;
    }
    *count = n;
    if (n > cap && (points || values)) return set_err(TCI_ERR_ARG, "output buffers hold %lld of %lld entries",
                                                      (long long)cap, (long long)n);
    return TCI_OK;
}

// every point coordinate in 1..d_b (the kernel indexes the cores with it)
static bool points_in_range(int L, const int32_t *localdims, int64_t npts, const uint8_t *pts)
{
    for (int64_t q = 0; q < npts; ++q)
        for (int b = 0; b < L; ++b) {
            const int v = pts[q * L + b];
            if (v < 1 || v > localdims[b]) return false;
        }
    return true;
}

// device part shared by the two TT evaluation entry points: points upload,
// launch, results back; dims3/coff/cores/params are device pointers
static tci_status tt_eval_device(tci_ctx *ctx, hipStream_t st, int L, const int64_t *d_dims3,
                                 const int64_t *d_coff, const double *d_cores, int chimax, int64_t npts,
                                 const uint8_t *pts, int fkind, const double *d_params,
                                 const int32_t *d_iparams, double *tt_out, double *f_out, double *maxerr,
                                 double *maxabsf)
{
    const int np = tt_eval_points_per_block(chimax);
    const int64_t blocks = (npts + np - 1) / np;
    size_t need = 0;
    auto a = [&](size_t bytes) { size_t at = need; need = lay_align(need + (bytes ? bytes : 1)); return at; };
    const size_t o_pts = a((size_t)(npts * L));
    const size_t o_tt = a(sizeof(double) * (size_t)npts);
    const size_t o_f = a(sizeof(double) * (size_t)npts);
    const size_t o_bm = a(sizeof(double) * (size_t)(2 * blocks));
    uint8_t *d = nullptr;
    HIPCHK(hipMallocAsync((void **)&d, need, st));
    hipError_t e = hipMemcpyAsync(d + o_pts, pts, (size_t)(npts * L), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_tt_eval(L, d_dims3, d_coff, d_cores, d + o_pts, npts, chimax, fkind, d_params, d_iparams,
                           (double *)(d + o_tt), fkind ? (double *)(d + o_f) : nullptr,
                           (double *)(d + o_bm), st);
    std::vector<double> bm;
    if (e == hipSuccess && tt_out)
        e = hipMemcpyAsync(tt_out, d + o_tt, sizeof(double) * (size_t)npts, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && fkind && f_out)
        e = hipMemcpyAsync(f_out, d + o_f, sizeof(double) * (size_t)npts, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && fkind) {
        bm.resize((size_t)(2 * blocks));
        e = hipMemcpyAsync(bm.data(), d + o_bm, sizeof(double) * bm.size(), hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(d, st);
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "TT evaluation: %s", hipGetErrorString(e));
    if (fkind) {
        double me = 0.0, mf = 0.0;
        for (int64_t q = 0; q < blocks; ++q) {
            me = nan_max(me, bm[(size_t)(2 * q)]);
            mf = nan_max(mf, bm[(size_t)(2 * q + 1)]);
        }
        if (maxerr) *maxerr = me;
        if (maxabsf) *maxabsf = mf;
    }
    return TCI_OK;
}

tci_status tci_tt_evaluate_f64(tci_ctx *ctx, int32_t L, const int64_t *linkdims, const int32_t *localdims,
                               const double *cores, int64_t npts, const uint8_t *pts, double *out)
{
    if (!ctx || L < 1 || !linkdims || !localdims || npts < 0 || (npts > 0 && (!pts || !out)))
        return set_err(TCI_ERR_ARG, "bad arguments");
    if (linkdims[0] != 1 || linkdims[L] != 1)
        return set_err(TCI_ERR_ARG, "a tensor train has bond dimension 1 at both ends");
    int64_t tot = 0, chimax = 1;
    std::vector<int64_t> dims3((size_t)(3 * L)), coff((size_t)L);
    for (int b = 0; b < L; ++b) {
        if (localdims[b] < 1 || localdims[b] > 255) return set_err(TCI_ERR_ARG, "local dimensions must be in 1..255");
        if (linkdims[b + 1] < 1 || linkdims[b + 1] > 8192) return set_err(TCI_ERR_ARG, "bond dimensions must be in 1..8192");
        dims3[(size_t)(3 * b)] = linkdims[b];
        dims3[(size_t)(3 * b + 1)] = localdims[b];
        dims3[(size_t)(3 * b + 2)] = linkdims[b + 1];
        coff[(size_t)b] = tot;
        tot += linkdims[b] * localdims[b] * linkdims[b + 1];
        chimax = linkdims[b + 1] > chimax ? linkdims[b + 1] : chimax;
    }
    if (npts == 0) return TCI_OK;
    if (!cores) return set_err(TCI_ERR_ARG, "cores is NULL");
    if (!points_in_range(L, localdims, npts, pts)) return set_err(TCI_ERR_ARG, "point index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    uint8_t *d = nullptr;
    size_t need = 0;
    auto a = [&](size_t bytes) { size_t at = need; need = lay_align(need + (bytes ? bytes : 1)); return at; };
    const size_t o_c = a(sizeof(double) * (size_t)tot);
    const size_t o_d = a(sizeof(int64_t) * dims3.size());
    const size_t o_o = a(sizeof(int64_t) * coff.size());
    HIPCHK(hipMallocAsync((void **)&d, need, st));
    hipError_t e = hipMemcpyAsync(d + o_c, cores, sizeof(double) * (size_t)tot, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d + o_d, dims3.data(), sizeof(int64_t) * dims3.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d + o_o, coff.data(), sizeof(int64_t) * coff.size(), hipMemcpyHostToDevice, st);
    tci_status s = TCI_OK;
    if (e != hipSuccess) s = set_err(TCI_ERR_DEVICE, "TT upload: %s", hipGetErrorString(e));
    else s = tt_eval_device(ctx, st, L, (const int64_t *)(d + o_d), (const int64_t *)(d + o_o),
                            (const double *)(d + o_c), (int)chimax, npts, pts, 0, nullptr, nullptr, out,
                            nullptr, nullptr, nullptr);
    (void)hipFreeAsync(d, st);
    (void)hipStreamSynchronize(st);
    return s;
}

tci_status tci_tt_floatingzone_f64(tci_ctx *ctx, const tci_functor *f, int32_t L, const int64_t *linkdims,
                                   const int32_t *localdims, const double *cores, int64_t npts, uint8_t *points,
                                   double *errors, int64_t nsweeps, double earlystoptol)
{
    if (!ctx || !f || L < 1 || L > 255 || !linkdims || !localdims || npts < 0 || (npts > 0 && (!points || !errors)))
        return set_err(TCI_ERR_ARG, "bad arguments");
    if (nsweeps < 1) return set_err(TCI_ERR_GENERIC, "nsweeps should be positive!");
    if (linkdims[0] != 1 || linkdims[L] != 1)
        return set_err(TCI_ERR_ARG, "a tensor train has bond dimension 1 at both ends");
    std::string msg;
    const int fst = check_functor(*f, L, localdims, &msg);
    if (fst != TCI_OK) return set_err((tci_status)fst, "%s", msg.c_str());
    int64_t tot = 0, chimax = 1, dmax = 1;
    std::vector<int64_t> dims3((size_t)(3 * L)), coff((size_t)L), roff((size_t)L + 1);
    for (int b = 0; b < L; ++b) {
        if (localdims[b] < 1 || localdims[b] > 255) return set_err(TCI_ERR_ARG, "local dimensions must be in 1..255");
        if (linkdims[b + 1] < 1 || linkdims[b + 1] > 8192) return set_err(TCI_ERR_ARG, "bond dimensions must be in 1..8192");
        dims3[(size_t)(3 * b)] = linkdims[b];
        dims3[(size_t)(3 * b + 1)] = localdims[b];
        dims3[(size_t)(3 * b + 2)] = linkdims[b + 1];
        coff[(size_t)b] = tot;
        tot += linkdims[b] * localdims[b] * linkdims[b + 1];
        chimax = linkdims[b + 1] > chimax ? linkdims[b + 1] : chimax;
        chimax = linkdims[b] > chimax ? linkdims[b] : chimax;
        dmax = localdims[b] > dmax ? localdims[b] : dmax;
        roff[(size_t)b + 1] = roff[(size_t)b] + linkdims[b + 1];   // E_{b+1} at roff[b]
    }
    if (npts == 0) return TCI_OK;
    if (!cores) return set_err(TCI_ERR_ARG, "cores is NULL");
    if (!points_in_range(L, localdims, npts, points)) return set_err(TCI_ERR_ARG, "point index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int blocks = (int)(npts < 1024 ? npts : 1024);
    const int64_t wstride = roff[(size_t)L] + chimax + dmax * chimax + dmax;
    size_t need = 0;
    auto a = [&](size_t bytes) { size_t at = need; need = lay_align(need + (bytes ? bytes : 1)); return at; };
    const size_t o_c = a(sizeof(double) * (size_t)tot);
    const size_t o_d = a(sizeof(int64_t) * dims3.size());
    const size_t o_o = a(sizeof(int64_t) * coff.size());
    const size_t o_r = a(sizeof(int64_t) * roff.size());
    const size_t o_par = a(sizeof(double) * (size_t)f->nparams);
    const size_t o_ipar = a(sizeof(int32_t) * (size_t)f->niparams);
    const size_t o_pts = a((size_t)(npts * L));
    const size_t o_err = a(sizeof(double) * (size_t)npts);
    const size_t o_w = a(sizeof(double) * (size_t)(wstride * blocks));
    const size_t o_dig = a((size_t)(dmax * L * blocks));
    uint8_t *d = nullptr;
    HIPCHK(hipMallocAsync((void **)&d, need, st));
    hipError_t e = hipMemcpyAsync(d + o_c, cores, sizeof(double) * (size_t)tot, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d + o_d, dims3.data(), sizeof(int64_t) * dims3.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d + o_o, coff.data(), sizeof(int64_t) * coff.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d + o_r, roff.data(), sizeof(int64_t) * roff.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && f->nparams)
        e = hipMemcpyAsync(d + o_par, f->params, sizeof(double) * f->nparams, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && f->niparams)
        e = hipMemcpyAsync(d + o_ipar, f->iparams, sizeof(int32_t) * f->niparams, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d + o_pts, points, (size_t)(npts * L), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_floatingzone(L, (const int64_t *)(d + o_d), (const int64_t *)(d + o_o), (const double *)(d + o_c),
                                (const int64_t *)(d + o_r), (int)chimax, (int)dmax, f->kind,
                                (const double *)(d + o_par), (const int32_t *)(d + o_ipar), d + o_pts,
                                (double *)(d + o_err), npts, nsweeps, earlystoptol, (double *)(d + o_w), wstride,
                                d + o_dig, blocks, st);
    if (e == hipSuccess) e = hipMemcpyAsync(points, d + o_pts, (size_t)(npts * L), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(errors, d + o_err, sizeof(double) * (size_t)npts, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(d, st);
    (void)hipStreamSynchronize(st);
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "floating zone search: %s", hipGetErrorString(e));
    return TCI_OK;
}

tci_status tci_batch_evaluate(tci_batch *b, int32_t i, int64_t npts, const uint8_t *pts, double *tt_out,
                              double *f_out, double *maxerr, double *maxabsf)
{
    SolveDesc R;
    if (!report_view(b, i, R) || npts < 0 || (npts > 0 && !pts)) return set_err(TCI_ERR_ARG, "bad arguments");
    if (R.ints[INT_STATUS] != TCI_OK || !R.ints[INT_DONE])
        return set_err(TCI_ERR_ARG, "solve %d did not finish successfully", i);
    if (npts == 0) {
        if (maxerr) *maxerr = 0.0;
        if (maxabsf) *maxabsf = 0.0;
        return TCI_OK;
    }
    tci_ctx *ctx = b->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    const int L = R.L;
    std::vector<int32_t> ld((size_t)L);
    std::vector<int64_t> dims3((size_t)(3 * L)), coff((size_t)L);
    HIPCHK(ctx_download(ctx, ld.data(), R.localdims, sizeof(int32_t) * (size_t)L));
    HIPCHK(ctx_download(ctx, dims3.data(), R.Tdims, sizeof(int64_t) * dims3.size()));
    if (!points_in_range(L, ld.data(), npts, pts)) return set_err(TCI_ERR_ARG, "point index out of range");
    int64_t chimax = 1;
    for (int s = 0; s < L; ++s) {
        coff[(size_t)s] = R.tstride * s;
        chimax = dims3[(size_t)(3 * s + 2)] > chimax ? dims3[(size_t)(3 * s + 2)] : chimax;
        if (dims3[(size_t)(3 * s + 1)] != ld[(size_t)s]) return set_err(TCI_ERR_GENERIC, "site tensor %d is missing", s + 1);
    }
    hipStream_t st = ctx->stream;
    int64_t *d_off = nullptr;
    HIPCHK(hipMallocAsync((void **)&d_off, sizeof(int64_t) * coff.size(), st));
    hipError_t e = hipMemcpyAsync(d_off, coff.data(), sizeof(int64_t) * coff.size(), hipMemcpyHostToDevice, st);
    tci_status s = TCI_OK;
    if (e != hipSuccess) s = set_err(TCI_ERR_DEVICE, "offsets upload: %s", hipGetErrorString(e));
    else s = tt_eval_device(ctx, st, L, R.Tdims, d_off, R.T, (int)chimax, npts, pts, R.fkind, R.params,
                            R.iparams, tt_out, f_out, maxerr, maxabsf);
    (void)hipFreeAsync(d_off, st);
    (void)hipStreamSynchronize(st);
    return s;
}

tci_status tci_batch_free(tci_batch *b)
{
    if (!b) return TCI_OK;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        if (b->ev1) (void)hipEventSynchronize(b->ev1);
    }
    if (b->served) {
        // the grid may still be running it: its blocks are released once done
        // (a lost request has no grid left to wait for)
        const tci_status w = b->lost ? TCI_OK : fetch_reports(b);
        if (w != TCI_OK) return w;
        if (tci_server *s = b->server) {
            s->slot[(size_t)(b->ticket % (uint64_t)s->cap)] = nullptr;
            if (!b->polled) {
                auto it = std::find(s->pending.begin(), s->pending.end(), b);
                if (it != s->pending.end()) {
                    *it = s->pending.back();
                    s->pending.pop_back();
                }
            }
        }
    }
    batch_release(b);
    return TCI_OK;
}

// ---------------------------------------------------------------- serving
// Memory of served requests.  A request takes one device block (its
// descriptor, inputs, state and report) and one pinned staging block (the
// image the claiming workgroup copies in, and the report it copies out), each
// rounded up to a power-of-two size class and carved from chunks that stay
// allocated while the server runs, so a submit calls no allocator and no
// stream operation once the classes it needs are warm (a device-wide
// synchronizing free would wait for the resident grid to leave).  A block
// returns to its class when its request is freed; the grid finished with it
// before the done mark the host saw.
constexpr int kServeMinClass = 12;                    // 4 KiB
constexpr size_t kServeChunkBytes = (size_t)32 << 20; // blocks are carved from 32 MiB chunks

static int serve_class(size_t bytes)
{
    int c = kServeMinClass;
    while (((size_t)1 << c) < bytes) ++c;
    return c;
}

static void serve_chunk_free(tci_ctx *ctx, ServeChunk *k)
{
    if (k->device) {
        (void)hipFree(k->base);
        if (ctx) ctx->used -= k->bytes;
    } else {
        (void)hipHostFree(k->base);
    }
    delete k;
}

static void serve_block_put(tci_server *s, tci_ctx *ctx, const ServeBlock &k, int cls, bool device)
{
    ServeChunk *c = k.chunk;
    c->inuse -= 1;
    if (s) {
        (device ? s->dfree : s->hfree)[(size_t)cls].push_back(k);
    } else if (c->orphan && c->inuse == 0) {
        serve_chunk_free(ctx, c);
    }
}

// a block of class cls; TCI_ERR_OOM when a new device chunk would exceed the
// context's HBM budget
static tci_status serve_block_get(tci_server *s, int cls, bool device, ServeBlock *out)
{
    auto &fl = (device ? s->dfree : s->hfree)[(size_t)cls];
    if (fl.empty()) {
        const size_t bsz = (size_t)1 << cls;
        const size_t csz = bsz > kServeChunkBytes ? bsz : kServeChunkBytes;
        tci_ctx *ctx = s->ctx;
        if (device && ctx->used + csz > ctx->budget)
            return set_err(TCI_ERR_OOM, "served request needs a %zu-byte block, HBM budget has %llu left", bsz,
                           (unsigned long long)(ctx->budget - ctx->used));
        ServeChunk *c = new ServeChunk();
        c->bytes = csz;
        c->device = device;
        hipError_t e;
        uint8_t *dev = nullptr;
        if (device) {
            e = hipMalloc(&c->base, csz);
            dev = (uint8_t *)c->base;
        } else {
            e = hipHostMalloc(&c->base, csz, hipHostMallocCoherent | hipHostMallocMapped);
            if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&dev, c->base, 0);
        }
        if (e != hipSuccess) {
            if (c->base) (void)(device ? hipFree(c->base) : hipHostFree(c->base));
            delete c;
            return set_err(e == hipErrorOutOfMemory ? TCI_ERR_OOM : TCI_ERR_DEVICE, "served memory: %s",
                           hipGetErrorString(e));
        }
        if (device) ctx->used += csz;
        s->chunks.push_back(c);
        for (size_t o = 0; o + bsz <= csz; o += bsz)
            fl.push_back(ServeBlock{(uint8_t *)c->base + o, dev + o, c});
    }
    *out = fl.back();
    fl.pop_back();
    out->chunk->inuse += 1;
    return TCI_OK;
}

static tci_status serve_ensure(tci_server *s)
{
    if (s->running) {
        const hipError_t q = hipEventQuery(s->kev);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError();
            return TCI_OK;
        }
        if (q != hipSuccess) return set_err(TCI_ERR_DEVICE, "serving grid: %s", hipGetErrorString(q));
        s->running = false;
    }
    // nothing published beyond what ran: stay idle
    if (__atomic_load_n(&s->ctl->tail, __ATOMIC_ACQUIRE) == 0) return TCI_OK;
    s->epoch += 1;
    hipError_t e = launch_serve(s->d_ring, s->d_g, s->d_ctl, s->d_done, s->cap, s->plan, s->idle_ticks, s->epoch,
                                s->team, s->blocks, s->kst);
    if (e == hipSuccess) e = hipEventRecord(s->kev, s->kst);
    if (e != hipSuccess) return set_err(TCI_ERR_DEVICE, "serving grid launch: %s", hipGetErrorString(e));
    s->running = true;
    return TCI_OK;
}

static bool served_done(const tci_server *s, const tci_batch *b)
{
    return __atomic_load_n(s->done + b->ticket % (uint64_t)s->cap, __ATOMIC_ACQUIRE) == b->ticket + 1;
}

static tci_status served_query(tci_batch *b)
{
    tci_server *s = b->server;
    if (!s) {
        // the server stopped: it drained every ticket unless it failed first
        if (b->lost) return set_err(TCI_ERR_DEVICE, "the server stopped on an error before this request ran");
        return TCI_OK;
    }
    if (served_done(s, b)) return TCI_OK;
    const tci_status e = serve_ensure(s);
    return e != TCI_OK ? e : TCI_ERR_PENDING;
}

static void server_release(tci_server *s)
{
    if (s->kev) (void)hipEventDestroy(s->kev);
    if (s->d_g) (void)hipFree(s->d_g);
    if (s->ring) (void)hipHostFree(s->ring);
    if (s->done) (void)hipHostFree(s->done);
    if (s->ctl) (void)hipHostFree(s->ctl);
    if (s->kst) (void)hipStreamDestroy(s->kst);
    for (ServeChunk *c : s->chunks) {
        if (c->inuse == 0) serve_chunk_free(s->ctx, c);
        else c->orphan = true;   // freed with the last request that holds a block of it
    }
    delete s;
}

tci_status tci_server_start(tci_ctx *ctx, const tci_problem *shape, int32_t slots, int32_t idle_ms, int32_t team,
                            tci_server **out)
{
    if (!ctx || !shape || !out || slots < 1 || slots > (1 << 20) || idle_ms < 1 || idle_ms > 60000 ||
        (team != 0 && team != 256 && team != 512))
        return set_err(TCI_ERR_ARG, "bad arguments");
    *out = nullptr;
    SolveCaps c;
    std::string msg;
    int st = check_problem(*shape, c, &msg);
    if (st != TCI_OK) return set_err((tci_status)st, "shape: %s", msg.c_str());
    HIPCHK(hipSetDevice(ctx->device));
    tci_server *s = new tci_server();
    s->ctx = ctx;
    s->cap = slots;
    s->team = team ? team : ctx->team == 512 ? 512 : 256;
    s->dfree.resize(64);
    s->hfree.resize(64);
    auto fail = [&](hipError_t e, const char *what) {
        server_release(s);
        return set_err(TCI_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    };
    hipError_t e;
    int ncu = 0, khz = 0;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device)) != hipSuccess)
        return fail(e, "device attributes");
    if ((e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device)) != hipSuccess)
        return fail(e, "device attributes");
    // a full grid holds every CU's register file, and the runtime's own copy
    // kernels (downloads of results) would wait for it to leave: keep a few
    // workgroup slots free for them
    const int reserve = ncu / 32 > 4 ? ncu / 32 : 4;
    s->blocks = (s->team == 512 ? 1 : 2) * ncu - reserve;
    if (s->blocks < 1) s->blocks = 1;
    s->idle_ticks = (uint64_t)idle_ms * (uint64_t)(khz > 0 ? khz : 100000);
    {
        // the grid's stream on a hardware queue of its own (a stream with a
        // CU mask is never given a shared one): a queue it shared would hold
        // every later command of the other stream behind the resident grid
        const char *qv = getenv("TCI_SERVE_QUEUE");
        if (qv && !strcmp(qv, "shared")) {
            e = hipStreamCreateWithFlags(&s->kst, hipStreamNonBlocking);
        } else {
            std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
            for (int q = 0; q < ncu; ++q) mask[(size_t)q / 32] |= 1u << (q % 32);
            e = hipExtStreamCreateWithCUMask(&s->kst, (uint32_t)mask.size(), mask.data());
        }
        if (e != hipSuccess) return fail(e, "stream");
    }
    if ((e = hipEventCreateWithFlags(&s->kev, hipEventDisableTiming)) != hipSuccess) return fail(e, "event");
    const unsigned hf = hipHostMallocCoherent | hipHostMallocMapped;
    if ((e = hipHostMalloc((void **)&s->ctl, sizeof(ServeCtl), hf)) != hipSuccess) return fail(e, "hipHostMalloc");
    if ((e = hipHostMalloc((void **)&s->done, sizeof(uint64_t) * (size_t)slots, hf)) != hipSuccess)
        return fail(e, "hipHostMalloc");
    if ((e = hipHostMalloc((void **)&s->ring, sizeof(ServeSlot) * (size_t)slots, hf)) != hipSuccess)
        return fail(e, "hipHostMalloc");
    memset(s->ctl, 0, sizeof(ServeCtl));
    memset(s->done, 0, sizeof(uint64_t) * (size_t)slots);
    memset(s->ring, 0, sizeof(ServeSlot) * (size_t)slots);
    if ((e = hipHostGetDevicePointer((void **)&s->d_ctl, s->ctl, 0)) != hipSuccess) return fail(e, "device alias");
    if ((e = hipHostGetDevicePointer((void **)&s->d_done, s->done, 0)) != hipSuccess) return fail(e, "device alias");
    if ((e = hipHostGetDevicePointer((void **)&s->d_ring, s->ring, 0)) != hipSuccess) return fail(e, "device alias");
    if ((e = hipMalloc((void **)&s->d_g, sizeof(ServeDev))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMemsetAsync(s->d_g, 0, sizeof(ServeDev), s->kst)) != hipSuccess) return fail(e, "hipMemset");
    if ((e = hipStreamSynchronize(s->kst)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    // one LDS plan for the grid, from the largest problem the caller expects
    // (a larger one still runs: what does not fit stays in its arena)
    int64_t lut = 0;
    if (shape->f.kind == TCI_F_LORENTZ_INT) {
        lut = 2;
        for (int t = 0; t < shape->L; ++t) lut += (int64_t)shape->localdims[t] * shape->localdims[t];
    }
    const int64_t sbytes = ((((int64_t)NSETKIND * c.L * c.setcap * c.L) + 15) & ~(int64_t)15) + 4 * (int64_t)NSETKIND * c.L;
    const int stat = serve_static_lds(s->team);
    if (stat < 0) return fail(hipErrorInvalidValue, "serving kernel attributes");
    s->plan = make_plan(ctx, (int64_t)lds_for_team(ctx, s->team) - stat, c.mrcap, c.mrcap * c.L, sbytes, lut);
    s->slot.assign((size_t)slots, nullptr);
    *out = s;
    return TCI_OK;
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// one request staged and its ring entry written, not yet published
static tci_status server_stage(tci_server *s, const tci_problem *p, tci_batch **out)
{
    *out = nullptr;
    const size_t k = (size_t)(s->next % (uint64_t)s->cap);
    if (s->slot[k]) return set_err(TCI_ERR_OOM, "server queue full: free finished requests first");
    SolveCaps c;
    std::string msg;
    const int chk = check_problem(*p, c, &msg);
    if (chk != TCI_OK) return set_err((tci_status)chk, "%s", msg.c_str());
    HIPCHK(hipSetDevice(s->ctx->device));
    size_t ib = 0, sb = 0, rb = 0;
    layout_solve(c, nullptr, nullptr, nullptr, nullptr, &ib, &sb, &rb);
    // device block: [descriptor][inputs][state][report], each 256-aligned;
    // staging block: [descriptor][inputs] as in the device block, then the
    // report's host copy
    const size_t oin = align_up(sizeof(SolveDesc), 256);
    const size_t ost = oin + align_up(ib, 256), orp = ost + align_up(sb, 256);
    const size_t dbytes = orp + align_up(rb, 256);
    const size_t staged = align_up(oin + ib, 16);
    const size_t hrep = align_up(staged, 256);
    const size_t hbytes = hrep + align_up(rb, 256);
    tci_batch *b = new tci_batch();
    b->ctx = s->ctx;
    b->count = 1;
    b->served = true;
    b->dcls = serve_class(dbytes);
    b->hcls = serve_class(hbytes);
    tci_status st = serve_block_get(s, b->dcls, true, &b->dblk);
    if (st == TCI_OK) st = serve_block_get(s, b->hcls, false, &b->hblk);
    if (st != TCI_OK) {
        b->server = s;
        batch_release(b);
        return st;
    }
    uint8_t *dblk = b->dblk.dev, *hblk = b->hblk.ptr;
    memset(hblk, 0, staged);
    SolveDesc &D = *(SolveDesc *)hblk;
    size_t x, y, z;
    layout_solve(c, &D, dblk + oin, dblk + ost, dblk + orp, &x, &y, &z);
    fill_desc_scalars(*p, D);
    pack_inputs(*p, c, D, hblk + oin, dblk + oin);
    b->desc.assign(1, D);
    b->rep_off.assign(1, 0);
    b->d_desc = (SolveDesc *)dblk;
    b->d_in = dblk + oin;
    b->d_state = dblk + ost;
    b->d_rep = dblk + orp;
    b->in_bytes = ib;
    b->state_bytes = sb;
    b->rep_bytes = rb;
    b->desc_bytes = sizeof(SolveDesc);
    b->h_rep_pin = hblk + hrep;
    ServeSlot &E = s->ring[k];
    E.src = b->hblk.dev;
    E.dst = dblk;
    E.rep = dblk + orp;
    E.rep_out = b->hblk.dev + hrep;
    E.bytes = staged;
    E.rep_bytes = align_up(rb, 16);
    b->server = s;
    b->ticket = s->next;
    s->slot[k] = b;
    s->pending.push_back(b);
    s->next += 1;
    *out = b;
    return TCI_OK;
}

tci_status tci_server_submit(tci_server *s, const tci_problem *p, tci_batch **out)
{
    if (!s || !p || !out) return set_err(TCI_ERR_ARG, "bad arguments");
    const tci_status st = server_stage(s, p, out);
    if (st != TCI_OK) return st;
    __atomic_store_n(&s->ctl->tail, s->next, __ATOMIC_RELEASE);   // publish
    return serve_ensure(s);
}

tci_status tci_server_submit_many(tci_server *s, int32_t count, const tci_problem *problems, tci_batch **out,
                                  int32_t *accepted)
{
    if (!s || !accepted || count < 0 || (count > 0 && (!problems || !out))) return set_err(TCI_ERR_ARG, "bad arguments");
    *accepted = 0;
    tci_status st = TCI_OK;
    for (int32_t q = 0; q < count; ++q) {
        st = server_stage(s, problems + q, out + q);
        if (st != TCI_OK) break;
        *accepted = q + 1;
    }
    if (*accepted > 0) {
        __atomic_store_n(&s->ctl->tail, s->next, __ATOMIC_RELEASE);   // publish them all
        const tci_status e = serve_ensure(s);
        if (e != TCI_OK) return e;
    }
    // a full ring or budget stops the burst without an error: retire and go on
    return st == TCI_ERR_OOM ? TCI_OK : st;
}

tci_status tci_server_poll(tci_server *s, tci_batch **out, int32_t cap, int32_t *count)
{
    if (!s || !count || cap < 0 || (cap > 0 && !out)) return set_err(TCI_ERR_ARG, "bad arguments");
    *count = 0;
    for (size_t i = 0; i < s->pending.size() && *count < cap;) {
        tci_batch *b = s->pending[i];
        if (served_done(s, b)) {
            b->polled = true;
            out[(*count)++] = b;
            s->pending[i] = s->pending.back();
            s->pending.pop_back();
        } else {
            ++i;
        }
    }
    // requests still waiting and the grid gone idle: launch it again
    if (!s->pending.empty()) return serve_ensure(s);
    return TCI_OK;
}

tci_status tci_server_stop(tci_server *s)
{
    if (!s) return TCI_OK;
    tci_status r = TCI_OK;
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    if (s->ctl) __atomic_store_n(&s->ctl->stop, 1u, __ATOMIC_RELEASE);
    // a grid that was leaving idle when stop came may leave tickets behind:
    // the next launch sees stop and drains them all, so two rounds suffice
    for (int round = 0; round < 3 && r == TCI_OK; ++round) {
        if (s->running) {
            const hipError_t e = hipEventSynchronize(s->kev);
            if (e != hipSuccess) {
                r = set_err(TCI_ERR_DEVICE, "serving grid: %s", hipGetErrorString(e));
                break;
            }
            s->running = false;
        }
        bool pending = false;
        if (s->ctl)
            for (tci_batch *b : s->slot)
                if (b && !served_done(s, b)) pending = true;
        if (!pending) break;
        r = serve_ensure(s);
    }
    for (tci_batch *b : s->slot)
        if (b) {
            // every ticket ran: its report comes to the host now, and its
            // blocks stay with it (their chunks outlive the server).  After
            // a failure, a ticket without its done mark never ran: its
            // queries, waits and results report the failure from then on
            if (r == TCI_OK) {
                const tci_status w = fetch_reports(b);
                if (w != TCI_OK) r = w;
            }
            if (!b->fetched && !served_done(s, b)) b->lost = true;
            b->server = nullptr;
        }
    server_release(s);
    return r;
}

}  // extern "C"
