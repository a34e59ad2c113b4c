// tci_rrlu_grid.hip — rrlu!(A) of ONE large matrix spread over the whole
// chip (matrixlu.jl:141-181): a cooperative grid of one workgroup per CU,
// each owning the physical columns j ≡ g (mod G).
//
// The row and column swaps of addpivot! (matrixlu.jl:98-112) are kept as
// logical permutations (rowAt/rowPos, colAt/colPos, identical in every
// workgroup's LDS); the values are updated where they lie, and one gather at
// the end writes rrlu!'s packed matrix.  Per pivot:
//   1. every workgroup reduces the candidates of the previous pass (one per
//      workgroup) with the total order of submatrixargmax (largest a², then
//      the first in column-major order of LOGICAL positions; NaN never wins;
//      nothing left → logical (k, k)), so all workgroups agree on the pivot
//      and on the stop test (matrixlu.jl:150-160);
//   2. the pivot column x (divided by the pivot when leftorthogonal) is read
//      by every workgroup; the pivot row y (divided when not leftorthogonal)
//      by the owner of each column;
//   3. each workgroup updates its trailing columns, A[i,j] -= x_i·y_j with the
//      product and the difference rounded separately, and in the same pass
//      finds its best candidate for the next pivot;
//   4. one grid barrier.
// The division written back into the pivot column is deferred to the next
// step (the other workgroups read the undivided column during this one).
// Every result is the same bits as the one-workgroup factorization.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include "tci_kernels.h"

namespace cg = cooperative_groups;

namespace tci {


struct GridCand {
    double v;              // a², -inf for none
    double a;              // the candidate's value
    int32_t cpos, rpos;    // logical position
    int32_t pr, pc;        // physical row / column
};

// The candidates are the only data every workgroup's control flow depends
// on, so they travel through agent-scope atomic stores and loads (coherent
// across the XCDs' L2s) on top of the grid barrier's fences.
__device__ inline void cand_store(GridCand *dst, const GridCand &c)
{
    uint64_t *w = (uint64_t *)dst;
    __hip_atomic_store(w + 0, __double_as_longlong(c.v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 1, __double_as_longlong(c.a), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 2, (uint64_t)(uint32_t)c.cpos | ((uint64_t)(uint32_t)c.rpos << 32), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 3, (uint64_t)(uint32_t)c.pr | ((uint64_t)(uint32_t)c.pc << 32), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline GridCand cand_load(const GridCand *src)
{
    uint64_t *w = (uint64_t *)src;
    GridCand c;
    c.v = __longlong_as_double(__hip_atomic_load(w + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    c.a = __longlong_as_double(__hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const uint64_t p = __hip_atomic_load(w + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t q = __hip_atomic_load(w + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    c.cpos = (int32_t)(uint32_t)p;
    c.rpos = (int32_t)(uint32_t)(p >> 32);
    c.pr = (int32_t)(uint32_t)q;
    c.pc = (int32_t)(uint32_t)(q >> 32);
    return c;
}

__device__ inline bool cand_better(double v, int cpos, int rpos, const GridCand &b)
{
    return v > b.v || (v == b.v && (cpos < b.cpos || (cpos == b.cpos && rpos < b.rpos)));
}

__device__ inline bool cand_better(const GridCand &a, const GridCand &b)
{
    return cand_better(a.v, a.cpos, a.rpos, b);
}

// team-wide best of one candidate per thread (result in every thread)
template <int kGridTeam>
__device__ GridCand cand_reduce(GridCand c, GridCand *red)
{
    const int tid = threadIdx.x;
    for (int off = 32; off >= 1; off >>= 1) {
        GridCand o;
        o.v = __shfl_xor(c.v, off);
        o.a = __shfl_xor(c.a, off);
        o.cpos = __shfl_xor(c.cpos, off);
        o.rpos = __shfl_xor(c.rpos, off);
        o.pr = __shfl_xor(c.pr, off);
        o.pc = __shfl_xor(c.pc, off);
        if (cand_better(o, c)) c = o;
    }
    if ((tid & 63) == 0) red[tid >> 6] = c;
    __syncthreads();
    GridCand b = red[0];
    for (int w = 1; w < kGridTeam / 64; ++w)
        if (cand_better(red[w], b)) b = red[w];
    __syncthreads();
    return b;
}

template <int kGridTeam>
__global__ __launch_bounds__(kGridTeam, 1) void tci_rrlu_grid_kernel(tci_lu_job *__restrict__ job,
                                                                    double *__restrict__ out,
                                                                    GridCand *__restrict__ cand,
                                                                    int32_t *__restrict__ flags)
{
    constexpr int kGridRows = kGridMaxSide / kGridTeam;   // rows per thread
    cg::grid_group grid = cg::this_grid();
    extern __shared__ int32_t perm_lds[];
    __shared__ GridCand red[kGridTeam / 64];
    __shared__ GridCand s_best;
    __shared__ double s_a;
    __shared__ int s_nan;
    const int tid = threadIdx.x, g = blockIdx.x, G = gridDim.x;
    double *__restrict__ A = job->A;
    const int m = (int)job->m, n = (int)job->n;
    const bool lo = job->leftorth != 0;
    const double reltol = job->reltol, abstol = job->abstol;
    int64_t mr64 = job->maxrank;
    if (m < mr64) mr64 = m;
    if (n < mr64) mr64 = n;
    const int mr = (int)mr64;
    int32_t *rowPos = perm_lds, *rowAt = rowPos + m, *colPos = rowAt + m, *colAt = colPos + n;
    for (int i = tid; i < m; i += kGridTeam) rowPos[i] = rowAt[i] = i;
    for (int j = tid; j < n; j += kGridTeam) colPos[j] = colAt[j] = j;
    __syncthreads();

    // pass 0: the first candidate (the whole matrix)
    GridCand best;
    best.v = -INFINITY;
    best.a = NAN;
    best.cpos = best.rpos = INT32_MAX;
    best.pr = best.pc = -1;
    for (int j = g; j < n; j += G) {
        const double *col = A + (int64_t)m * j;
        for (int i = tid; i < m; i += kGridTeam) {
            const double a = col[i];
            const double v = a * a;
            if (cand_better(v, j, i, best)) {
                best.v = v;
                best.a = a;
                best.cpos = j;
                best.rpos = i;
                best.pr = i;
                best.pc = j;
            }
        }
    }
    best = cand_reduce<kGridTeam>(best, red);
    if (tid == 0) cand_store(cand + g, best);
    grid.sync();

    double maxerr = 0.0, err = NAN;
    int p = 0, pend_col = -1;   // the pivot column whose division is still to be written
    double pend_a = 0.0;
    while (p < mr) {
        const int k = p;
        GridCand *cur = cand + (int64_t)(k & 1) * G;
        // 1. the pivot
        GridCand c;
        c.v = -INFINITY;
        c.a = NAN;
        c.cpos = c.rpos = INT32_MAX;
        c.pr = c.pc = -1;
        for (int q = tid; q < G; q += kGridTeam) {
            const GridCand o = cand_load(cur + q);
            if (cand_better(o, c)) c = o;
        }
        c = cand_reduce<kGridTeam>(c, red);
        if (tid == 0) {
            s_best = c;
            s_a = c.a;
            if (c.pr < 0) {
                // nothing but NaN left: (k, k) (matrixlu.jl:9, 21-31), whose
                // value is then NaN too
                s_best.pr = rowAt[k];
                s_best.pc = colAt[k];
                s_a = NAN;
            }
        }
        __syncthreads();
        const int pr = s_best.pr, pc = s_best.pc;
        const double a = s_a;
        err = fabs(a);
        if ((err < reltol * maxerr || err < abstol) && p > 0) break;   // uniform in the grid
        maxerr = (maxerr != maxerr || err != err) ? maxerr + err : (maxerr > err ? maxerr : err);
        // the division of the previous pivot column, now that nobody reads it
        if (lo && pend_col >= 0 && (pend_col % G) == g) {
            double *col = A + (int64_t)m * pend_col;
            for (int i = tid; i < m; i += kGridTeam)
                if (rowPos[i] > k - 1) col[i] = col[i] / pend_a;
        }
        __syncthreads();
        // logical swaps k ↔ pivot (swaprow!/swapcol!)
        if (tid == 0) {
            const int pp = rowPos[pr], rk = rowAt[k];
            rowAt[pp] = rk;
            rowPos[rk] = pp;
            rowAt[k] = pr;
            rowPos[pr] = k;
            const int qq = colPos[pc], ck = colAt[k];
            colAt[qq] = ck;
            colPos[ck] = qq;
            colAt[k] = pc;
            colPos[pc] = k;
        }
        __syncthreads();
        // 2. x over my rows (trailing rows: logical position > k)
        double x[kGridRows];
        uint32_t tmask = 0;
        const double *pcol = A + (int64_t)m * pc;
#pragma unroll
        for (int r = 0; r < kGridRows; ++r) {
            const int i = tid + kGridTeam * r;
            x[r] = 0.0;
            if (i < m && rowPos[i] > k) {
                tmask |= 1u << r;
                const double v = pcol[i];
                x[r] = lo ? v / a : v;
            }
        }
        // 3. update my trailing columns, with the next candidate
        best.v = -INFINITY;
        best.a = NAN;
        best.cpos = best.rpos = INT32_MAX;
        best.pr = best.pc = -1;
        for (int j = g; j < n; j += G) {
            const int cp = colPos[j];
            if (cp <= k) continue;
            double *col = A + (int64_t)m * j;
            double y = col[pr];
            if (!lo) y = y / a;   // written back below, once every thread has read it
#pragma unroll
            for (int r = 0; r < kGridRows; ++r) {
                if (tmask & (1u << r)) {
                    const int i = tid + kGridTeam * r;
                    const double t = x[r] * y;
                    const double v = col[i] - t;
                    col[i] = v;
                    const double v2 = v * v;
                    if (v2 > best.v || v2 == best.v) {
                        const int rp = rowPos[i];
                        if (cand_better(v2, cp, rp, best)) {
                            best.v = v2;
                            best.a = v;
                            best.cpos = cp;
                            best.rpos = rp;
                            best.pr = i;
                            best.pc = j;
                        }
                    }
                }
            }
        }
        best = cand_reduce<kGridTeam>(best, red);   // (its barriers also order the reads of row pr)
        if (!lo)
            for (int j = g + G * tid; j < n; j += G * kGridTeam)
                if (colPos[j] > k) A[pr + (int64_t)m * j] = A[pr + (int64_t)m * j] / a;
        if (tid == 0) cand_store(cand + (int64_t)((k + 1) & 1) * G + g, best);
        pend_col = pc;
        pend_a = a;
        ++p;
        grid.sync();
    }
    // the last division, then the gather of the packed matrix
    if (lo && pend_col >= 0 && (pend_col % G) == g) {
        double *col = A + (int64_t)m * pend_col;
        for (int i = tid; i < m; i += kGridTeam)
            if (rowPos[i] > p - 1) col[i] = col[i] / pend_a;
    }
    if (tid == 0) s_nan = 0;
    grid.sync();
    int nanL = 0, nanU = 0;
    for (int J = g; J < n; J += G) {
        const double *src = A + (int64_t)m * colAt[J];
        double *dst = out + (int64_t)m * J;
        for (int I = tid; I < m; I += kGridTeam) {
            const double v = src[rowAt[I]];
            dst[I] = v;
            if (v != v) {
                if (J < p && I >= J) nanL = 1;
                if (I < p && I <= J) nanU = 1;
            }
        }
    }
    if (nanL) atomicOr(&s_nan, 1);
    if (nanU) atomicOr(&s_nan, 2);
    __syncthreads();
    if (tid == 0 && s_nan) atomicOr(flags, s_nan);
    grid.sync();
    if (g == 0) {
        for (int i = tid; i < m; i += kGridTeam) job->rowperm[i] = rowAt[i] + 1;
        for (int j = tid; j < n; j += kGridTeam) job->colperm[j] = colAt[j] + 1;
        if (tid == 0) {
            const int f = atomicAdd(flags, 0);
            // lu.L is checked first (matrixlu.jl:164-169)
            job->status = (f & 1) ? TCI_ERR_NAN_L : (f & 2) ? TCI_ERR_NAN_U : TCI_OK;
            job->npivot = p;
            job->error = p >= (m < n ? m : n) ? 0.0 : err;
        }
    }
}

size_t rrlu_grid_lds(int64_t m, int64_t n) { return (size_t)4 * 2 * (size_t)(m + n); }

// threads per workgroup: 1024 (16 waves per CU keep enough loads in
// flight for the streamed update); TCI_RRLU_GRID_TEAM = 256 or 512 overrides
static int grid_team()
{
    static int t = [] {
        const char *e = getenv("TCI_RRLU_GRID_TEAM");
        const int v = e ? atoi(e) : 1024;
        return (v == 256 || v == 512 || v == 1024) ? v : 1024;
    }();
    return t;
}

static const void *grid_kernel(int team)
{
    switch (team) {
    case 256: return (const void *)tci_rrlu_grid_kernel<256>;
    case 512: return (const void *)tci_rrlu_grid_kernel<512>;
    default: return (const void *)tci_rrlu_grid_kernel<1024>;
    }
}

int rrlu_grid_default_team() { return grid_team(); }

static int valid_team(int team) { return (team == 256 || team == 512 || team == 1024) ? team : grid_team(); }

int rrlu_grid_blocks(int device, size_t lds, int team)
{
    team = valid_team(team);
    int ncu = 0, per = 0;
    const void *k = grid_kernel(team);
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, team, lds) != hipSuccess) return 0;
    return per < 1 ? 0 : ncu;   // one workgroup per CU
}

size_t rrlu_grid_cand_bytes(int blocks) { return sizeof(GridCand) * 2 * (size_t)blocks; }

hipError_t launch_rrlu_grid(tci_lu_job *job, int64_t m, int64_t n, double *out, void *cand, int32_t *flags,
                            int blocks, int team, hipStream_t st)
{
    if (m < 1 || n < 1 || m > kGridMaxSide || n > kGridMaxSide || blocks < 1) return hipErrorInvalidValue;
    team = valid_team(team);
    const size_t lds = rrlu_grid_lds(m, n);
    const void *k = grid_kernel(team);
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    void *args[] = {(void *)&job, (void *)&out, (void *)&cand, (void *)&flags};
    return hipLaunchCooperativeKernel(k, dim3((unsigned)blocks), dim3(team), args, (unsigned)lds, st);
}

}  // namespace tci
