// tci_kernels.hip — gfx950 kernels of the TCI2 sweep.
//
//   tci_solve_kernel     one workgroup = one whole crossinterpolate2 solve
//                        (tensorci2.jl:943-953), state resident in HBM
//   tci_rrlu_kernel      one workgroup = one rrlu!(A) (matrixlu.jl:194-204),
//                        a batch of uneven matrices in one launch
//   tci_lu_unpack_kernel L = tril(A[:,1:p]), U = triu(A[1:p,:]) with the unit
//                        diagonal (matrixlu.jl:162-174)
//   tci_kron_center_kernel / tci_eval_kernel
//                        Pi-tensor evaluation (batcheval.jl:32-61)
//
// Workgroups are 256 threads = 4 wave64s.  Team reductions: DPP/shuffle
// within a wave, then one LDS slot per wave read back by every thread, so a
// reduction costs one barrier.  Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "tci_devteam.h"
#include "tci_engine.h"
#include "tci_kernels.h"

namespace tci {


constexpr int kTeam = 256;

// ---------------------------------------------------------------- solves
// Dynamic LDS of a solve workgroup (sizes fixed per launch by LdsPlan):
//   [stage: 2·stage_cap doubles][rows: tab_cap bytes][cols: tab_cap bytes]
//   [sets: sets_cap bytes (the pivot sets, then their counts)]
//   [Lorentzian value table: lut_cap doubles][Pi area: pi_cap doubles]
// Anything that does not fit stays in the solve's HBM arena.
template <int NT, int FK, int WPE>
__device__ __forceinline__ void solve_on_team(const SolveDesc &D, const LdsPlan &plan, EngShared &S, double *dyn)
{
    DevTeam<NT, WPE> tm;
    double *stage = dyn;
    uint8_t *tabs = (uint8_t *)(dyn + 2 * plan.stage_cap);
    uint8_t *lsets = tabs + 2 * plan.tab_cap;
    double *lut = (double *)(lsets + plan.sets_cap);
    double *pi = lut + plan.lut_cap;
    if (D.mrcap <= plan.stage_cap) {
        tm.init(stage, stage + D.mrcap);
        tm.stage_lds = true;
    } else {
        tm.init(D.stage, D.stage + D.mrcap);
    }
    if ((int64_t)D.mrcap * D.L <= plan.tab_cap) {
        tm.lrows = tabs;
        tm.lcols = tabs + plan.tab_cap;
    }
    const int64_t setbytes = (int64_t)NSETKIND * D.L * D.setcap * D.L;
    if (((setbytes + 15) & ~(int64_t)15) + 4 * NSETKIND * D.L <= plan.sets_cap) {
        tm.lsets = lsets;
        tm.lsetn = (int32_t *)(lsets + ((setbytes + 15) & ~(int64_t)15));
    }
    tm.llut = plan.lut_cap ? lut : nullptr;
    tm.llut_cap = plan.lut_cap;
    tm.lpi = pi;
    tm.lpi_cap = plan.pi_cap;
    Engine<DevTeam<NT, WPE>, FK> eng(tm, D, S);
    eng.run();
}

template <int NT, int FK, int WPE>
__global__ __launch_bounds__(NT, WPE) void tci_solve_kernel(const SolveDesc *__restrict__ descs,
                                                       int count, LdsPlan plan)
{
    const int s = blockIdx.x;
    if (s >= count) return;
    __shared__ EngShared S;
    extern __shared__ double dyn[];
    solve_on_team<NT, FK, WPE>(descs[s], plan, S, dyn);
}

// ------------------------------------------------------------------ serving
// Slot-level serving: a persistent grid whose workgroups each take one
// published solve at a time from a ring of requests.  The host publishes
// ticket t by writing ring[t % cap] and the staged image it names (coherent
// host memory, before) and then ctl->tail = t + 1.  One workgroup at a time, at most
// every kServePollTicks, reads the host's words and raises the device mirror
// g->dtail, so the idle workgroups poll device memory, not the bus.  A
// workgroup claims the next ticket below the mirror by a compare-and-swap on
// g->head, so a ticket is claimed only once it is published and every claimed
// ticket is run; after the solve its done slot (coherent host memory) becomes
// t + 1.  The grid leaves as a whole: with stop seen, once every published
// ticket is claimed; otherwise once no workgroup has run, claimed or finished
// a solve for idle_ticks (the first to see that marks the epoch closing and
// the others follow; one that claimed meanwhile still runs its ticket).  A
// ticket published while the grid leaves is run by the next launch.
constexpr uint64_t kServePollTicks = 50;   // 0.5 us of the 100 MHz wall clock

// n bytes (a multiple of 16, both ends 16-byte aligned) by the whole team
template <int NT>
__device__ __forceinline__ void team_copy16(uint8_t *dst, const uint8_t *src, uint64_t n)
{
    for (uint64_t o = (uint64_t)threadIdx.x * 16; o < n; o += (uint64_t)NT * 16)
        *(uint4 *)(dst + o) = *(const uint4 *)(src + o);
}

// the engine run of one served request, compiled apart from the claim loop
// so that the loop's live values do not crowd the engine's registers
template <int NT, int WPE>
__device__ __noinline__ void serve_solve(const SolveDesc *D, LdsPlan plan)
{
    __shared__ EngShared S;
    extern __shared__ double dyn[];
    solve_on_team<NT, 0, WPE>(*D, plan, S, dyn);
}

template <int NT, int WPE>
__global__ __launch_bounds__(NT, WPE) void tci_serve_kernel(const ServeSlot *__restrict__ ring, ServeDev *g,
                                                            const ServeCtl *ctl, uint64_t *done, int cap,
                                                            LdsPlan plan, uint64_t idle_ticks, uint32_t epoch)
{
    __shared__ uint64_t s_ticket;
    __shared__ int s_go;
    __shared__ ServeSlot s_slot;
    extern __shared__ double dyn[];
    const uint64_t born = (uint64_t)wall_clock64();
    for (;;) {
        if (threadIdx.x == 0) {
            int go = 0, nap = 1;
            uint64_t t = 0;
            for (;;) {
                uint64_t tail = __hip_atomic_load(&g->dtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                uint64_t h = __hip_atomic_load(&g->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t now = (uint64_t)wall_clock64();
                if (h >= tail) {
                    uint64_t lp = __hip_atomic_load(&g->lastpoll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (now - lp >= kServePollTicks &&
                        __hip_atomic_compare_exchange_strong(&g->lastpoll, &lp, now, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        // stop first: a tail read after it holds every ticket
                        // published before the stop
                        const bool st =
                            __hip_atomic_load(&ctl->stop, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
                        const uint64_t ht = __hip_atomic_load(&ctl->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
                        if (st) __hip_atomic_store(&g->dstop, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_fetch_max(&g->dtail, ht, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                        if (ht > tail) tail = ht;
                    }
                }
                const bool stopping =
                    __hip_atomic_load(&g->dstop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
                if (!stopping && __hip_atomic_load(&g->closing, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch)
                    break;
                if (h < tail) {
                    if (__hip_atomic_compare_exchange_strong(&g->head, &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_fetch_add(&g->busy, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(&g->last, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        t = h;
                        go = 1;
                        break;
                    }
                    continue;
                }
                if (stopping) break;
                if (__hip_atomic_load(&g->busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                    // the last finish's clock is published before its busy decrement
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    const uint64_t last = __hip_atomic_load(&g->last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint64_t ref = last > born ? last : born;
                    if (now > ref && now - ref > idle_ticks) {
                        __hip_atomic_store(&g->closing, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
                // back off while idle: about 0.9 us, doubling to about 55 us,
                // so a mostly idle grid does not keep the memory channel of
                // these words busy under the solves still running
                for (int k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(32);
                if (nap < 64) nap *= 2;
            }
            if (go) {
                // the ring entry and its staged image were written before
                // the publication: read them after it
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                const uint64_t *e = (const uint64_t *)(ring + t % (uint64_t)cap);
                uint64_t *w = (uint64_t *)&s_slot;
                for (int k = 0; k < (int)(sizeof(ServeSlot) / 8); ++k)
                    w[k] = __hip_atomic_load(e + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            s_ticket = t;
            s_go = go;
        }
        __syncthreads();
        if (!s_go) return;
        const uint64_t ticket = s_ticket;
        // the descriptor and inputs into the request's device block
        team_copy16<NT>(s_slot.dst, s_slot.src, s_slot.bytes);
        __threadfence();
        __syncthreads();
        serve_solve<NT, WPE>((const SolveDesc *)s_slot.dst, plan);
        __syncthreads();
        team_copy16<NT>(s_slot.rep_out, s_slot.rep, s_slot.rep_bytes);
        __threadfence_system();   // every thread's part of the report before the done mark
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_store(done + ticket % (uint64_t)cap, ticket + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&g->last, (uint64_t)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_sub(&g->busy, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

static const void *serve_kernel_ptr(int team)
{
    return team == 512 ? (const void *)tci_serve_kernel<512, 2> : (const void *)tci_serve_kernel<256, 2>;
}

int serve_static_lds(int team)
{
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, serve_kernel_ptr(team)) != hipSuccess) return -1;
    return (int)a.sharedSizeBytes;
}

hipError_t launch_serve(const ServeSlot *ring, ServeDev *g, const ServeCtl *ctl, uint64_t *done, int cap,
                        const LdsPlan &plan, uint64_t idle_ticks, uint32_t epoch, int team, int blocks,
                        hipStream_t st)
{
    const size_t lds = lds_plan_bytes(plan);
    int lds_blk = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    const int stat = serve_static_lds(team);
    if (stat < 0 || (size_t)stat + lds > (size_t)lds_blk) return hipErrorInvalidValue;
    const void *k = serve_kernel_ptr(team);
    if ((e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess) return e;
    if (team == 512)
        hipLaunchKernelGGL((tci_serve_kernel<512, 2>), dim3(blocks), dim3(512), lds, st, ring, g, ctl, done, cap,
                           plan, idle_ticks, epoch);
    else
        hipLaunchKernelGGL((tci_serve_kernel<256, 2>), dim3(blocks), dim3(256), lds, st, ring, g, ctl, done, cap,
                           plan, idle_ticks, epoch);
    return hipGetLastError();
}

// ------------------------------------------------------------------ rrLU
// order: the job each workgroup takes (largest first, so the long
// factorizations start early and the launch ends evenly), or null
__global__ __launch_bounds__(kTeam) void tci_rrlu_kernel(tci_lu_job *__restrict__ jobs,
                                                         int64_t count, int lds_stage_cap,
                                                         double *__restrict__ gstage,
                                                         int64_t gstage_stride,
                                                         const int32_t *__restrict__ order)
{
    if ((int64_t)blockIdx.x >= count) return;
    const int64_t s = order ? (int64_t)order[blockIdx.x] : (int64_t)blockIdx.x;
    extern __shared__ double stage[];
    tci_lu_job &J = jobs[s];
    const int m = (int)J.m, n = (int)J.n;
    const int mx = m > n ? m : n;
    DevTeam<kTeam> tm;
    if (mx <= lds_stage_cap) tm.init(stage, stage + mx);
    else tm.init(gstage + gstage_stride * s, gstage + gstage_stride * s + mx);
    double err, st[2];
    // the streaming factorization: the packed factors are the output
    const int p = rrlu_team(tm, J.A, m, n, J.maxrank, J.reltol, J.abstol, J.leftorth != 0,
                            J.rowperm, J.colperm, nullptr, &err, st);
    // 1-based permutations (rrLU.rowpermutation / colpermutation)
    for (int i = tm.tid(); i < m; i += kTeam) J.rowperm[i] += 1;
    for (int j = tm.tid(); j < n; j += kTeam) J.colperm[j] += 1;
    if (tm.tid() == 0) {
        J.status = p < 0 ? -p : 0;
        J.npivot = p < 0 ? 0 : p;
        J.error = err;
    }
}

// The register-tile factorization of one job (tci_regtile.h) with tile shape
// T, then rrlu!'s packed output written to J.A: the L columns (recorded by
// original row in Lsc, m × mr) and U rows (by original column in Usc,
// mr × n) placed at their final positions, the pivots on the diagonal and
// the trailing block from the registers.
template <class T, class Team>
__device__ void tile_job(Team &tm, tci_lu_job &J, int m, int n, double *Lsc, double *Usc, double *pv)
{
    int64_t mr = J.maxrank;
    if (m < mr) mr = m;
    if (n < mr) mr = n;
    T R;
    R.setup(tm, m, n);
    R.load(tm, J.A, m);
    __syncthreads();   // every load of A done before A is rewritten
    double err, st[2];
    const int p = R.factor(tm, J.maxrank, J.reltol, J.abstol, J.leftorth != 0, pv, &err, st, Lsc,
                           J.rowperm, J.colperm, Usc, mr);
    R.store_trailing(tm, J.A, m, p > 0 ? p : 0);
    __syncthreads();   // permutations, pivots and recorded factors complete
    const int pp = p > 0 ? p : 0;
    for (int64_t q = tm.tid(); q < (int64_t)m * pp; q += tm.nt()) {
        const int k = (int)(q / m), pos = (int)(q - (int64_t)k * m);
        if (pos > k) J.A[pos + (int64_t)m * k] = Lsc[J.rowperm[pos] + (int64_t)m * k];
        else if (pos == k) J.A[k + (int64_t)m * k] = pv[k];
    }
    for (int64_t q = tm.tid(); q < (int64_t)pp * n; q += tm.nt()) {
        const int pj = (int)(q / pp), k = (int)(q - (int64_t)pj * pp);
        if (pj > k) J.A[k + (int64_t)m * pj] = Usc[k + mr * J.colperm[pj]];
    }
    __syncthreads();
    for (int i = tm.tid(); i < m; i += tm.nt()) J.rowperm[i] += 1;
    for (int j = tm.tid(); j < n; j += tm.nt()) J.colperm[j] += 1;
    if (tm.tid() == 0) {
        J.status = p < 0 ? -p : 0;
        J.npivot = pp;
        J.error = err;
    }
}

// rrlu!(A) of the jobs that fit a 128×128 register tile, one workgroup per
// job: the matrix stays in registers during the factorization (the solve
// engine's tile, here with its packed factors written out).  idx: the job
// indices; scratch + soff[q]: the job's Lsc (m × mr) then Usc (mr × n).
__global__ __launch_bounds__(kTeam, 2) void tci_rrlu_tile_kernel(tci_lu_job *__restrict__ jobs,
                                                                const int32_t *__restrict__ idx,
                                                                int64_t count, double *__restrict__ scratch,
                                                                const int64_t *__restrict__ soff)
{
    if ((int64_t)blockIdx.x >= count) return;
    using Team = DevTeam<kTeam, 2>;
    __shared__ double pv[128];
    tci_lu_job &J = jobs[idx[blockIdx.x]];
    const int m = (int)J.m, n = (int)J.n;
    Team tm;
    tm.init(nullptr, nullptr);
    int64_t mr = J.maxrank;
    if (m < mr) mr = m;
    if (n < mr) mr = n;
    double *Lsc = scratch + soff[blockIdx.x];
    double *Usc = Lsc + (int64_t)m * mr;
    if (TileLU<Team, 1, 8>::fits(m, n)) tile_job<TileLU<Team, 1, 8>>(tm, J, m, n, Lsc, Usc, pv);
    else if (TileLU<Team, 1, 32>::fits(m, n)) tile_job<TileLU<Team, 1, 32>>(tm, J, m, n, Lsc, Usc, pv);
    else if (TileLU<Team, 2, 8>::fits(m, n)) tile_job<TileLU<Team, 2, 8>>(tm, J, m, n, Lsc, Usc, pv);
    else tile_job<TileLU<Team, 2, 32>>(tm, J, m, n, Lsc, Usc, pv);
}

// entry q of the unpacked factors: q < m·p an entry of L, else of U.  The
// bodies below are shared by the standalone kernels and the fused TT sweep,
// so both compute every entry with the same operations
__device__ __forceinline__ void lu_unpack_at(int64_t q, const double *A, int64_t m, int64_t p, int leftorth,
                                             double *L, int64_t ldl, double *U, int64_t ldu)
{
    const int64_t nl = m * p;
    if (q < nl) {
        int64_t j = q / m, i = q - j * m;
        double v = i >= j ? A[i + m * j] : 0.0;
        if (leftorth && i == j) v = 1.0;
        L[i + ldl * j] = v;
    } else {
        int64_t r = q - nl;
        int64_t j = r / p, i = r - j * p;
        double v = i <= j ? A[i + m * j] : 0.0;
        if (!leftorth && i == j) v = 1.0;
        U[i + ldu * j] = v;
    }
}

__global__ void tci_lu_unpack_kernel(const double *__restrict__ A, int64_t m, int64_t n, int64_t p,
                                     int leftorth, double *__restrict__ L, int64_t ldl,
                                     double *__restrict__ U, int64_t ldu)
{
    const int64_t nl = m * p, nu = p * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nl + nu;
         q += (int64_t)gridDim.x * blockDim.x)
        lu_unpack_at(q, A, m, p, leftorth, L, ldl, U, ldu);
}

// ------------------------------------------------------------ products
// C = A·B (M×K times K×N, column-major), one thread per entry of C, each
// sum in ascending k from 0.0 with a rounding per product and per sum (the
// order of the oracle's loops; the reference calls BLAS here).
__device__ __forceinline__ void gemm_at(int64_t q, int64_t M, int64_t K, const double *A, int64_t lda,
                                        const double *B, int64_t ldb, double *Cm, int64_t ldc)
{
    const int64_t j = q / M, i = q - j * M;
    double acc = 0.0;
    for (int64_t k = 0; k < K; ++k) acc = acc + A[i + lda * k] * B[k + ldb * j];
    Cm[i + ldc * j] = acc;
}

__global__ void tci_gemm_kernel(int64_t M, int64_t N, int64_t K, const double *__restrict__ A, int64_t lda,
                                const double *__restrict__ B, int64_t ldb, double *__restrict__ Cm, int64_t ldc)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < M * N;
         q += (int64_t)gridDim.x * blockDim.x)
        gemm_at(q, M, K, A, lda, B, ldb, Cm, ldc);
}

hipError_t launch_gemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                       double *Cm, int64_t ldc, hipStream_t st)
{
    const int64_t tot = M * N;
    if (tot == 0) return hipSuccess;
    int64_t blocks = (tot + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(tci_gemm_kernel, dim3((unsigned)blocks), dim3(256), 0, st, M, N, K, A, lda, B, ldb, Cm, ldc);
    return hipGetLastError();
}

// ---------------------------------------------------------- LUCI factors
// left(luci) and right(luci) of MatrixLUCI (matrixluci.jl:40-84) from the
// unpacked factors L (m×p, lda m) and U (p×n, lda p) of one rrLU with
// 1-based permutations rowperm/colperm (int32).  One thread per row of left
// (position i, written to row rowperm[i]-1) and one per column of right
// (position j, written to column colperm[j]-1); each substitution or dot
// product runs in a fixed order (ascending t for products from 0.0, back to
// front for the triangular solves).
__device__ __forceinline__ void luci_at(int64_t q, int64_t m, int64_t p, int leftorth, const int32_t *rowperm,
                                        const int32_t *colperm, const double *L, const double *U, double *left,
                                        int64_t ldl, double *right, int64_t ldr)
{
    if (q < m) {
        const int64_t i = q, r = rowperm[i] - 1;
        double *out = left + r;   // row r of left, stride ldl
        if (leftorth) {
            // colstimespivotinv (48-57): [I_p; L21·L11⁻¹]; x·L11 = L[i,1:p]
            if (i < p) {
                for (int64_t j = 0; j < p; ++j) out[ldl * j] = i == j ? 1.0 : 0.0;
            } else {
                for (int64_t j = p - 1; j >= 0; --j) {
                    double s = L[i + m * j];
                    for (int64_t t = j + 1; t < p; ++t) s = s - out[ldl * t] * L[t + m * j];
                    out[ldl * j] = s / L[j + m * j];
                }
            }
        } else {
            // colmatrix (40-42): L · U[:,1:p]
            for (int64_t j = 0; j < p; ++j) {
                double s = 0.0;
                const int64_t tmax = i < j ? i : j;
                for (int64_t t = 0; t <= tmax && t < p; ++t) s = s + L[i + m * t] * U[t + p * j];
                out[ldl * j] = s;
            }
        }
    } else {
        const int64_t j = q - m, c = colperm[j] - 1;
        double *out = right + ldr * c;   // column c of right
        if (leftorth) {
            // rowmatrix (44-46): L[1:p,:] · U
            for (int64_t i = 0; i < p; ++i) {
                double s = 0.0;
                const int64_t tmax = i < j ? i : j;
                for (int64_t t = 0; t <= tmax; ++t) s = s + L[i + m * t] * U[t + p * j];
                out[i] = s;
            }
        } else {
            // pivotinvtimesrows (59-68): [I_p  U11⁻¹·U12]
            if (j < p) {
                for (int64_t i = 0; i < p; ++i) out[i] = i == j ? 1.0 : 0.0;
            } else {
                for (int64_t i = p - 1; i >= 0; --i) {
                    double s = U[i + p * j];
                    for (int64_t t = i + 1; t < p; ++t) s = s - U[i + p * t] * out[t];
                    out[i] = s / U[i + p * i];
                }
            }
        }
    }
}

__global__ void tci_luci_kernel(int64_t m, int64_t n, int64_t p, int leftorth, const int32_t *__restrict__ rowperm,
                                const int32_t *__restrict__ colperm, const double *__restrict__ L,
                                const double *__restrict__ U, double *__restrict__ left, int64_t ldl,
                                double *__restrict__ right, int64_t ldr)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m + n;
         q += (int64_t)gridDim.x * blockDim.x)
        luci_at(q, m, p, leftorth, rowperm, colperm, L, U, left, ldl, right, ldr);
}

hipError_t launch_luci(int64_t m, int64_t n, int64_t p, int leftorth, const int32_t *rowperm, const int32_t *colperm,
                       const double *L, const double *U, double *left, int64_t ldl, double *right, int64_t ldr,
                       hipStream_t st)
{
    const int64_t tot = m + n;
    if (tot == 0 || p == 0) return hipSuccess;
    int64_t blocks = (tot + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(tci_luci_kernel, dim3((unsigned)blocks), dim3(256), 0, st, m, n, p, leftorth, rowperm, colperm,
                       L, U, left, ldl, right, ldr);
    return hipGetLastError();
}

// left(lu) and right(lu) with permute=true (matrixlu.jl:374-392): row i of
// L goes to row rowperm[i]-1 of left, column j of U to column colperm[j]-1
// of right (L m×p with lda m, U p×n with lda p; permutations 1-based)
__device__ __forceinline__ void lu_permute_at(int64_t q, int64_t m, int64_t p, const int32_t *rowperm,
                                              const int32_t *colperm, const double *L, const double *U,
                                              double *left, int64_t ldl, double *right, int64_t ldr)
{
    const int64_t nl = m * p;
    if (q < nl) {
        const int64_t j = q / m, i = q - j * m;
        left[(rowperm[i] - 1) + ldl * j] = L[i + m * j];
    } else {
        const int64_t r = q - nl, j = r / p, i = r - j * p;
        right[i + ldr * (colperm[j] - 1)] = U[i + p * j];
    }
}

__global__ void tci_lu_permute_kernel(int64_t m, int64_t n, int64_t p, const int32_t *__restrict__ rowperm,
                                      const int32_t *__restrict__ colperm, const double *__restrict__ L,
                                      const double *__restrict__ U, double *__restrict__ left, int64_t ldl,
                                      double *__restrict__ right, int64_t ldr)
{
    const int64_t nl = m * p, nu = p * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nl + nu;
         q += (int64_t)gridDim.x * blockDim.x)
        lu_permute_at(q, m, p, rowperm, colperm, L, U, left, ldl, right, ldr);
}

// ------------------------------------------------------ fused TT sweep
// The host-driven sweep (per bond: rrlu!, unpack, left/right of rrLU or
// MatrixLUCI, the product into the neighbour; tensortrain.jl:155-182,
// conversion.jl:105-141) in one workgroup: the streaming rrLU of the team
// (whose packed result equals every other rrLU kernel's) and the entry bodies
// of the standalone kernels above, so every value has the same bits.  Each
// bond's new dimension p is returned to every thread by the factorization,
// so the shapes of the next bond need no memory round trip.
constexpr int kTtTeam = 256;

__device__ __forceinline__ double tt_nan_max(double a, double b)
{
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

__global__ __launch_bounds__(kTtTeam) void tci_tt_sweep_kernel(TtSweepArgs a)
{
    extern __shared__ double stage[];
    DevTeam<kTtTeam> tm;
    const int L = a.L;
    const bool fw = a.forward != 0;
    int64_t carried = -1;   // the previous bond's new dimension
    for (int it = 0; it < L - 1; ++it) {
        const int b = fw ? it : L - 2 - it;
        const int s = fw ? b : b + 1;
        const int64_t m = fw ? (carried >= 0 ? carried : a.links[b]) * a.d[b] : a.links[b + 1];
        const int64_t n = fw ? a.links[b + 1] : a.d[b + 1] * (carried >= 0 ? carried : a.links[b + 2]);
        double *A = a.cores + a.off[s];
        const int64_t mx = m > n ? m : n;
        if (mx <= a.stage_cap) tm.init(stage, stage + mx);
        else tm.init(a.gstage, a.gstage + mx);
        // a bond matrix that fits is factored in LDS (the same operations on
        // a copy with a shorter latency) and copied back
        double *Aw = A;
        if (m * n <= a.mat_cap) {
            Aw = stage + 2 * (int64_t)a.stage_cap;
            for (int64_t q = tm.tid(); q < m * n; q += kTtTeam) Aw[q] = A[q];
            tm.sync();
        }
        double err, st2[2];
        const int p = rrlu_team(tm, Aw, (int)m, (int)n, a.maxrank, a.reltol, a.abstol, fw, a.perm, a.perm + m, nullptr,
                                &err, st2);
        if (Aw != A) {
            tm.sync();
            for (int64_t q = tm.tid(); q < m * n; q += kTtTeam) A[q] = Aw[q];
        }
        if (p < 0) {
            if (tm.tid() == 0) {
                a.status[0] = -p;
                a.status[1] = b;
            }
            return;
        }
        int32_t *rp = a.perm, *cp = a.perm + m;
        for (int64_t i = tm.tid(); i < m; i += kTtTeam) rp[i] += 1;
        for (int64_t j = tm.tid(); j < n; j += kTtTeam) cp[j] += 1;
        tm.sync();
        // running pivot errors (the diagonal of the packed factors, one
        // thread per entry) and the pivots
        for (int k = tm.tid(); k < p; k += kTtTeam) a.pe[k] = tt_nan_max(a.pe[k], fabs(A[k + m * k]));
        if (tm.tid() == 0) a.pe[p] = tt_nan_max(a.pe[p], err);
        int32_t *pv = a.piv + (int64_t)b * 2 * a.R;
        for (int k = tm.tid(); k < p; k += kTtTeam) {
            pv[k] = rp[k];
            pv[p + k] = cp[k];
        }
        for (int64_t q = tm.tid(); q < m * p + (int64_t)p * n; q += kTtTeam)
            lu_unpack_at(q, A, m, p, fw ? 1 : 0, a.wl, m, a.wu, p);
        tm.sync();
        double *lft = fw ? A : a.wc, *rgt = fw ? a.wc : A;
        if (a.method == 0) {
            for (int64_t q = tm.tid(); q < m * p + (int64_t)p * n; q += kTtTeam)
                lu_permute_at(q, m, p, rp, cp, a.wl, a.wu, lft, m, rgt, p);
        } else if (p > 0) {
            for (int64_t q = tm.tid(); q < m + n; q += kTtTeam)
                luci_at(q, m, p, fw ? 1 : 0, rp, cp, a.wl, a.wu, lft, m, rgt, p);
        }
        tm.sync();
        if (fw) {
            const int64_t nn = a.d[b + 1] * a.links[b + 2];
            double *nx = a.cores + a.off[b + 1];
            for (int64_t q = tm.tid(); q < (int64_t)p * nn; q += kTtTeam) gemm_at(q, p, n, a.wc, p, nx, n, a.wp, p);
            tm.sync();
            for (int64_t q = tm.tid(); q < (int64_t)p * nn; q += kTtTeam) nx[q] = a.wp[q];
        } else {
            const int64_t mm = a.links[b] * a.d[b];
            double *pv2 = a.cores + a.off[b];
            for (int64_t q = tm.tid(); q < mm * p; q += kTtTeam) gemm_at(q, mm, m, pv2, mm, a.wc, m, a.wp, mm);
            tm.sync();
            for (int64_t q = tm.tid(); q < mm * p; q += kTtTeam) pv2[q] = a.wp[q];
        }
        if (tm.tid() == 0) a.links[b + 1] = p;
        carried = p;
        tm.sync();
    }
}

hipError_t launch_tt_sweep(const TtSweepArgs &a, hipStream_t st)
{
    const size_t lds = sizeof(double) * (2 * (size_t)a.stage_cap + (size_t)a.mat_cap);
    hipError_t e = hipFuncSetAttribute((const void *)tci_tt_sweep_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(tci_tt_sweep_kernel, dim3(1), dim3(kTtTeam), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_lu_permute(int64_t m, int64_t n, int64_t p, const int32_t *rowperm, const int32_t *colperm,
                             const double *L, const double *U, double *left, int64_t ldl, double *right, int64_t ldr,
                             hipStream_t st)
{
    const int64_t tot = m * p + p * n;
    if (tot == 0) return hipSuccess;
    int64_t blocks = (tot + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(tci_lu_permute_kernel, dim3((unsigned)blocks), dim3(256), 0, st, m, n, p, rowperm, colperm,
                       L, U, left, ldl, right, ldr);
    return hipGetLastError();
}

// ------------------------------------------------------------ evaluation
// rowsx[r = i + nI·c] = left[i] ⊕ central digits of c (first fastest)
__global__ void tci_kron_center_kernel(const uint8_t *__restrict__ left, int64_t nI, int nl,
                                       const int32_t *__restrict__ cdims, int M, int64_t nc,
                                       uint8_t *__restrict__ rowsx, int stride)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nI * nc;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = r / nI, i = r - c * nI;
        uint8_t *dst = rowsx + r * stride;
        for (int t = 0; t < nl; ++t) dst[t] = left[i * nl + t];
        for (int t = 0; t < M; ++t) {
            int64_t dt = cdims[t];
            int64_t q = c / dt;
            dst[nl + t] = (uint8_t)(c - q * dt + 1);
            c = q;
        }
    }
}

// out[i + nr·j] = f(rows[i] ⊕ cols[j]) with a NaN-propagating max|out| per block
__global__ void tci_eval_kernel(int kind, const double *__restrict__ params,
                                const int32_t *__restrict__ iparams, int L,
                                const uint8_t *__restrict__ rows, int64_t nrow, int rstride, int nl,
                                const uint8_t *__restrict__ cols, int64_t ncol, int cstride,
                                double *__restrict__ out, double *__restrict__ blockmax)
{
    __shared__ double red[4];
    double lmax = 0.0;
    const int64_t tot = nrow * ncol;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = q / nrow, i = q - j * nrow;
        tci_point x;
        x.a = rows + i * rstride;
        x.b = cols + j * cstride;
        x.na = nl;
        x.L = L;
        double v = tci_functor_eval(kind, params, iparams, &x);
        out[q] = v;
        lmax = jl_max(lmax, fabs(v));
    }
    for (int off = 32; off >= 1; off >>= 1) lmax = jl_max(lmax, __shfl_xor(lmax, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = jl_max(r, red[w]);
        blockmax[blockIdx.x] = r;
    }
}

// --------------------------------------------------------- TT evaluation
// evaluate(tt, x) (abstracttensortrain.jl:124-132) at npts points: for each
// point the row vector v (χ_0 = 1 entries of 1) times each site's slice
// T_b[:, x_b, :], every product summed in ascending bond order from 0.0.
// A workgroup takes np points; their running vectors live in LDS (two
// buffers of np·chimax doubles), and thread t computes entries (p, j) of
// the next vector.  Site b's core starts at cores + coff[b]; its shape is
// dims3[3b..3b+2].  With fkind != 0, f is evaluated at the same points and
// the workgroup writes the NaN-propagating maxima of |TT - f| and |f| to
// blockmax[2·block], [2·block + 1].
__global__ __launch_bounds__(256) void tci_tt_eval_kernel(
    int L, const int64_t *__restrict__ dims3, const int64_t *__restrict__ coff,
    const double *__restrict__ cores, const uint8_t *__restrict__ pts, int64_t npts, int np,
    int chimax, int fkind, const double *__restrict__ params, const int32_t *__restrict__ iparams,
    double *__restrict__ out_tt, double *__restrict__ out_f, double *__restrict__ blockmax)
{
    extern __shared__ double vbuf[];
    __shared__ double red[8];
    const int64_t p0 = (int64_t)blockIdx.x * np;
    const int cn = (int)(npts - p0 < np ? npts - p0 : np);
    double *V = vbuf, *W = vbuf + (int64_t)np * chimax;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int q = tid; q < cn; q += nt) V[(int64_t)q * chimax] = 1.0;
    __syncthreads();
    for (int b = 0; b < L; ++b) {
        const int64_t d0 = dims3[3 * b], d = dims3[3 * b + 1], d2 = dims3[3 * b + 2];
        const double *T = cores + coff[b];
        for (int64_t o = tid; o < (int64_t)cn * d2; o += nt) {
            const int64_t q = o / d2, j = o - q * d2;
            const int64_t s = pts[(p0 + q) * L + b] - 1;
            const double *Tc = T + d0 * (s + d * j);
            const double *v = V + q * chimax;
            double acc = 0.0;
            for (int64_t i = 0; i < d0; ++i) acc = acc + v[i] * Tc[i];
            W[q * chimax + j] = acc;
        }
        __syncthreads();
        double *t = V; V = W; W = t;
    }
    double le = 0.0, lf = 0.0;
    for (int q = tid; q < cn; q += nt) {
        const double tt = V[(int64_t)q * chimax];
        out_tt[p0 + q] = tt;
        if (fkind) {
            tci_point x;
            x.a = pts + (p0 + q) * L;
            x.b = x.a;
            x.na = L;
            x.L = L;
            const double fv = tci_functor_eval(fkind, params, iparams, &x);
            if (out_f) out_f[p0 + q] = fv;
            le = jl_max(le, fabs(tt - fv));
            lf = jl_max(lf, fabs(fv));
        }
    }
    if (!fkind) return;
    for (int off = 32; off >= 1; off >>= 1) {
        le = jl_max(le, __shfl_xor(le, off));
        lf = jl_max(lf, __shfl_xor(lf, off));
    }
    if ((tid & 63) == 0) {
        red[2 * (tid >> 6)] = le;
        red[2 * (tid >> 6) + 1] = lf;
    }
    __syncthreads();
    if (tid == 0) {
        double e = red[0], f = red[1];
        for (int w = 1; w < nt / 64; ++w) {
            e = jl_max(e, red[2 * w]);
            f = jl_max(f, red[2 * w + 1]);
        }
        blockmax[2 * blockIdx.x] = e;
        blockmax[2 * blockIdx.x + 1] = f;
    }
}

// _floatingzone (globalsearch.jl:43-95) from each start point: one
// workgroup per point (grid-stride).  Per sweep, the right environments
// R_b = T_b[:, x_b, :]·R_{b+1} (R_L = [1], each entry summed in ascending
// bond order) are built from the current point; then for each site p, with
// l the row vector of the sites before p at the current point,
//   w_σ[k] = Σ_j l[j] T_p[j,σ,k]  (ascending j),  TT_σ = Σ_k w_σ[k] R_{p+1}[k]
// (ascending k), err_σ = |f(x with x_p = σ) − TT_σ|; x_p moves to Julia's
// argmax of err (first maximum, NaN above every number) and the running
// error takes the NaN-propagating max.  A sweep that leaves the error
// unchanged, or lifts it above earlystoptol, ends the search; so does a
// NaN error (with which the reference would sweep without end).
// Work area per workgroup (wstride doubles): renv (Σ_{b=1..L} χ_b, at
// roff[b-1]) | l (chimax) | w (dmax·chimax) | err (dmax); digit rows
// (dmax·L bytes per workgroup) hold the candidate points.
__global__ __launch_bounds__(256) void tci_floatingzone_kernel(
    int L, const int64_t *__restrict__ dims3, const int64_t *__restrict__ coff, const double *__restrict__ cores,
    const int64_t *__restrict__ roff, int chimax, int dmax, int fkind, const double *__restrict__ params,
    const int32_t *__restrict__ iparams, uint8_t *__restrict__ pts, double *__restrict__ errs, int64_t npts,
    int64_t nsweeps, double earlystoptol, double *__restrict__ work, int64_t wstride, uint8_t *__restrict__ digits)
{
    __shared__ uint8_t piv[256];
    __shared__ double s_maxerr;
    __shared__ int s_best, s_stop;
    const int tid = threadIdx.x, nt = blockDim.x;
    double *renv = work + (int64_t)blockIdx.x * wstride;
    double *lv = renv + roff[L];
    double *w = lv + chimax;
    double *err = w + (int64_t)dmax * chimax;
    uint8_t *rows = digits + (int64_t)blockIdx.x * dmax * L;
    for (int64_t q = blockIdx.x; q < npts; q += gridDim.x) {
        for (int t = tid; t < L; t += nt) piv[t] = pts[q * L + t];
        __syncthreads();
        for (int64_t sweep = 0;; ++sweep) {
            // right environments of the current point
            if (tid == 0) renv[roff[L - 1]] = 1.0;   // R_L
            __syncthreads();
            for (int b = L - 1; b >= 1; --b) {
                const int64_t d0 = dims3[3 * b], d = dims3[3 * b + 1], d2 = dims3[3 * b + 2];
                const double *Tc = cores + coff[b] + d0 * (int64_t)(piv[b] - 1);
                const double *Rn = renv + roff[b];
                double *Rb = renv + roff[b - 1];
                for (int64_t j = tid; j < d0; j += nt) {
                    double acc = 0.0;
                    for (int64_t k = 0; k < d2; ++k) acc = acc + Tc[j + d0 * d * k] * Rn[k];
                    Rb[j] = acc;
                }
                __syncthreads();
            }
            if (sweep == 0) {
                // the start point's error: TT(x) = T_1[1, x_1, :]·R_2
                if (tid == 0) {
                    const int64_t d = dims3[1], d2 = dims3[2];
                    const double *Tc = cores + coff[0] + (int64_t)(piv[0] - 1);
                    const double *Rn = renv + roff[0];
                    double acc = 0.0;
                    for (int64_t k = 0; k < d2; ++k) acc = acc + (0.0 + 1.0 * Tc[d * k]) * Rn[k];
                    tci_point x;
                    x.a = piv;
                    x.b = piv;
                    x.na = L;
                    x.L = L;
                    s_maxerr = fabs(tci_functor_eval(fkind, params, iparams, &x) - acc);
                }
                __syncthreads();
            }
            const double prev = s_maxerr;
            if (tid == 0) lv[0] = 1.0;
            __syncthreads();
            for (int p = 0; p < L; ++p) {
                const int64_t d0 = dims3[3 * p], d = dims3[3 * p + 1], d2 = dims3[3 * p + 2];
                const double *T = cores + coff[p];
                const double *Rn = renv + roff[p];   // R_{p+1}
                for (int64_t o = tid; o < d * d2; o += nt) {
                    const int64_t s = o / d2, k = o - s * d2;
                    const double *Tc = T + d0 * (s + d * k);
                    double acc = 0.0;
                    for (int64_t j = 0; j < d0; ++j) acc = acc + lv[j] * Tc[j];
                    w[s * chimax + k] = acc;
                }
                for (int64_t o = tid; o < d * L; o += nt) {
                    const int64_t s = o / L, t = o - s * L;
                    rows[s * L + t] = t == p ? (uint8_t)(s + 1) : piv[t];
                }
                __syncthreads();
                for (int64_t s = tid; s < d; s += nt) {
                    double acc = 0.0;
                    for (int64_t k = 0; k < d2; ++k) acc = acc + w[s * chimax + k] * Rn[k];
                    tci_point x;
                    x.a = rows + s * L;
                    x.b = x.a;
                    x.na = L;
                    x.L = L;
                    err[s] = fabs(tci_functor_eval(fkind, params, iparams, &x) - acc);
                }
                __syncthreads();
                if (tid == 0) {
                    int best = 0;
                    double bv = err[0], mx = err[0];
                    for (int s = 1; s < (int)d; ++s) {
                        const double e = err[s];
                        // isless(bv, e): NaN is above every number, the first NaN stays
                        if (bv == bv && (e != e || bv < e)) {
                            best = s;
                            bv = e;
                        }
                        mx = jl_max(mx, e);
                    }
                    piv[p] = (uint8_t)(best + 1);
                    s_best = best;
                    s_maxerr = jl_max(mx, s_maxerr);
                }
                __syncthreads();
                for (int64_t k = tid; k < d2; k += nt) lv[k] = w[s_best * chimax + k];
                __syncthreads();
            }
            if (tid == 0) {
                const double me = s_maxerr;
                s_stop = me == prev || me > earlystoptol || me != me || sweep + 1 >= nsweeps;
            }
            __syncthreads();
            if (s_stop) break;
        }
        for (int t = tid; t < L; t += nt) pts[q * L + t] = piv[t];
        if (tid == 0) errs[q] = s_maxerr;
        __syncthreads();
    }
}

hipError_t launch_floatingzone(int L, const int64_t *dims3, const int64_t *coff, const double *cores,
                               const int64_t *roff, int chimax, int dmax, int fkind, const double *params,
                               const int32_t *iparams, uint8_t *pts, double *errs, int64_t npts, int64_t nsweeps,
                               double earlystoptol, double *work, int64_t wstride, uint8_t *digits, int blocks,
                               hipStream_t st)
{
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(tci_floatingzone_kernel, dim3((unsigned)blocks), dim3(256), 0, st, L, dims3, coff, cores, roff,
                       chimax, dmax, fkind, params, iparams, pts, errs, npts, nsweeps, earlystoptol, work, wstride,
                       digits);
    return hipGetLastError();
}

// ------------------------------------------------------------- launchers
int tt_eval_points_per_block(int chimax)
{
    // two running-vector buffers of np·chimax doubles in at most 64 KiB
    int np = (int)((64 * 1024) / (16 * (int64_t)(chimax > 0 ? chimax : 1)));
    if (np > 256) np = 256;
    return np < 1 ? 1 : np;
}

hipError_t launch_tt_eval(int L, const int64_t *dims3, const int64_t *coff, const double *cores,
                          const uint8_t *pts, int64_t npts, int chimax, int fkind, const double *params,
                          const int32_t *iparams, double *out_tt, double *out_f, double *blockmax,
                          hipStream_t st)
{
    if (npts <= 0) return hipSuccess;
    const int np = tt_eval_points_per_block(chimax);
    const int64_t blocks = (npts + np - 1) / np;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    const size_t lds = (size_t)2 * np * chimax * sizeof(double);
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tci_tt_eval_kernel, dim3((unsigned)blocks), dim3(256), lds, st, L, dims3, coff, cores,
                       pts, npts, np, chimax, fkind, params, iparams, out_tt, out_f, blockmax);
    return hipGetLastError();
}

size_t lds_plan_bytes(const LdsPlan &p)
{
    return (size_t)2 * p.stage_cap * sizeof(double) + 2 * (size_t)p.tab_cap + (size_t)p.sets_cap +
           (size_t)p.lut_cap * sizeof(double) + (size_t)p.pi_cap * sizeof(double);
}

// the kernel variant a (team, functor kind) pair runs
static const void *solve_kernel_ptr(int team, int fkind)
{
    switch (team) {
    case 64:
        return fkind == TCI_F_QUANTICS_GAUSS_COS ? (const void *)tci_solve_kernel<64, TCI_F_QUANTICS_GAUSS_COS, 2>
                                                 : (const void *)tci_solve_kernel<64, 0, 2>;
    case 1024:
        return fkind == TCI_F_LORENTZ_INT ? (const void *)tci_solve_kernel<1024, TCI_F_LORENTZ_INT, 4>
                                          : (const void *)tci_solve_kernel<1024, 0, 4>;
    case 512:
        return fkind == TCI_F_LORENTZ_INT ? (const void *)tci_solve_kernel<512, TCI_F_LORENTZ_INT, 2>
                                          : (const void *)tci_solve_kernel<512, 0, 2>;
    default:
        switch (fkind) {
        case TCI_F_LORENTZ_INT: return (const void *)tci_solve_kernel<256, TCI_F_LORENTZ_INT, 2>;
        case TCI_F_LORENTZ_WEIGHTED: return (const void *)tci_solve_kernel<256, TCI_F_LORENTZ_WEIGHTED, 2>;
        case TCI_F_QUANTICS_GAUSS_COS: return (const void *)tci_solve_kernel<256, TCI_F_QUANTICS_GAUSS_COS, 2>;
        case TCI_F_GAUSS_SUM: return (const void *)tci_solve_kernel<256, TCI_F_GAUSS_SUM, 2>;
        default: return (const void *)tci_solve_kernel<256, 0, 2>;
        }
    }
}

int solve_static_lds(int team, int fkind)
{
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, solve_kernel_ptr(team, fkind)) != hipSuccess) return -1;
    return (int)a.sharedSizeBytes;
}

template <int NT, int FK, int WPE>
static void launch_one(const SolveDesc *descs, int count, const LdsPlan &plan, size_t lds, hipStream_t st)
{
    if (hipFuncSetAttribute((const void *)tci_solve_kernel<NT, FK, WPE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return;   // the launch guard in launch_solve reports the error
    hipLaunchKernelGGL((tci_solve_kernel<NT, FK, WPE>), dim3(count), dim3(NT), lds, st, descs, count, plan);
}

template <int NT, int WPE>
static void launch_nt(const SolveDesc *descs, int count, int fkind, const LdsPlan &plan, size_t lds,
                      hipStream_t st)
{
    switch (fkind) {
    case TCI_F_LORENTZ_INT: launch_one<NT, TCI_F_LORENTZ_INT, WPE>(descs, count, plan, lds, st); break;
    case TCI_F_LORENTZ_WEIGHTED: launch_one<NT, TCI_F_LORENTZ_WEIGHTED, WPE>(descs, count, plan, lds, st); break;
    case TCI_F_QUANTICS_GAUSS_COS: launch_one<NT, TCI_F_QUANTICS_GAUSS_COS, WPE>(descs, count, plan, lds, st); break;
    case TCI_F_GAUSS_SUM: launch_one<NT, TCI_F_GAUSS_SUM, WPE>(descs, count, plan, lds, st); break;
    default: launch_one<NT, 0, WPE>(descs, count, plan, lds, st); break;
    }
}

// fkind: the functor kind shared by every solve of the batch, or 0 (mixed)
hipError_t launch_solve(const SolveDesc *descs, int count, int team, int fkind, const LdsPlan &plan,
                        hipStream_t st)
{
    const size_t lds = lds_plan_bytes(plan);
    // a launch whose LDS (static + dynamic) exceeds the workgroup limit is
    // refused here, never sent to the device
    int lds_blk = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    const int stat = solve_static_lds(team, fkind);
    if (stat < 0 || (size_t)stat + lds > (size_t)lds_blk) return hipErrorInvalidValue;
    // team shapes: 256 threads at 256 VGPRs, two solves per CU (the default,
    // every functor specialized); 512 threads at 256 VGPRs, one solve per CU;
    // 1024 threads at 128 VGPRs
    switch (team) {
    case 64:
        if (fkind == TCI_F_QUANTICS_GAUSS_COS) launch_one<64, TCI_F_QUANTICS_GAUSS_COS, 2>(descs, count, plan, lds, st);
        else launch_one<64, 0, 2>(descs, count, plan, lds, st);
        break;
    case 1024:
        if (fkind == TCI_F_LORENTZ_INT) launch_one<1024, TCI_F_LORENTZ_INT, 4>(descs, count, plan, lds, st);
        else launch_one<1024, 0, 4>(descs, count, plan, lds, st);
        break;
    case 512:
        if (fkind == TCI_F_LORENTZ_INT) launch_one<512, TCI_F_LORENTZ_INT, 2>(descs, count, plan, lds, st);
        else launch_one<512, 0, 2>(descs, count, plan, lds, st);
        break;
    default: launch_nt<256, 2>(descs, count, fkind, plan, lds, st); break;
    }
    return hipGetLastError();
}

hipError_t launch_rrlu(tci_lu_job *jobs, int64_t count, int64_t max_side, double *gstage,
                       int64_t gstage_stride, const int32_t *order, hipStream_t st)
{
    const int cap = 48 * 1024 / 16;
    const size_t lds = max_side <= cap ? (size_t)2 * max_side * sizeof(double) : 0;
    hipLaunchKernelGGL(tci_rrlu_kernel, dim3((unsigned)count), dim3(kTeam), lds, st, jobs, count,
                       max_side <= cap ? (int)max_side : 0, gstage, gstage_stride, order);
    return hipGetLastError();
}

hipError_t launch_rrlu_tile(tci_lu_job *jobs, const int32_t *idx, int64_t count, double *scratch,
                            const int64_t *soff, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(tci_rrlu_tile_kernel, dim3((unsigned)count), dim3(kTeam), 0, st, jobs, idx, count,
                       scratch, soff);
    return hipGetLastError();
}

hipError_t launch_lu_unpack(const double *A, int64_t m, int64_t n, int64_t p, int leftorth,
                            double *L, int64_t ldl, double *U, int64_t ldu, hipStream_t st)
{
    int64_t tot = m * p + p * n;
    if (tot == 0) return hipSuccess;
    int64_t blocks = (tot + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(tci_lu_unpack_kernel, dim3((unsigned)blocks), dim3(256), 0, st, A, m, n, p,
                       leftorth, L, ldl, U, ldu);
    return hipGetLastError();
}

hipError_t launch_kron_center(const uint8_t *left, int64_t nI, int nl, const int32_t *cdims, int M,
                              int64_t nc, uint8_t *rowsx, int stride, hipStream_t st)
{
    int64_t tot = nI * nc;
    if (tot == 0) return hipSuccess;
    int64_t blocks = (tot + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(tci_kron_center_kernel, dim3((unsigned)blocks), dim3(256), 0, st, left, nI,
                       nl, cdims, M, nc, rowsx, stride);
    return hipGetLastError();
}

int eval_blocks(int64_t tot)
{
    int64_t b = (tot + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t launch_eval(int kind, const double *params, const int32_t *iparams, int L,
                       const uint8_t *rows, int64_t nrow, int rstride, int nl, const uint8_t *cols,
                       int64_t ncol, int cstride, double *out, double *blockmax, hipStream_t st)
{
    int blocks = eval_blocks(nrow * ncol);
    hipLaunchKernelGGL(tci_eval_kernel, dim3(blocks), dim3(256), 0, st, kind, params, iparams, L,
                       rows, nrow, rstride, nl, cols, ncol, cstride, out, blockmax);
    return hipGetLastError();
}

}  // namespace tci
