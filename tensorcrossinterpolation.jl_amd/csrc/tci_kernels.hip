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

#include "tci_engine.h"
#include "tci_kernels.h"

namespace tci {

template <int NT>
struct DevTeam {
    static constexpr int NW = NT / 64;
    int tid;
    int nt;
    double *xs, *ys;
    ArgBest *red_a;   // 2 × NW
    double *red_d;    // 2 × NW
    int par_a, par_d;

    __device__ void sync() { __syncthreads(); }

    __device__ ArgBest argmax(ArgBest b)
    {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            ArgBest o;
            o.v = __shfl_xor(b.v, off);
            o.lin = (int64_t)__shfl_xor((long long)b.lin, off);
            o.a = __shfl_xor(b.a, off);
            if (arg_better(o, b)) b = o;
        }
        ArgBest *slot = red_a + par_a * NW;
        if ((tid & 63) == 0) slot[tid >> 6] = b;
        __syncthreads();
        ArgBest r = slot[0];
#pragma unroll
        for (int w = 1; w < NW; ++w)
            if (arg_better(slot[w], r)) r = slot[w];
        par_a ^= 1;
        return r;
    }

    __device__ double max_nan(double v)
    {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v = jl_max(v, __shfl_xor(v, off));
        double *slot = red_d + par_d * NW;
        if ((tid & 63) == 0) slot[tid >> 6] = v;
        __syncthreads();
        double r = slot[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) r = jl_max(r, slot[w]);
        par_d ^= 1;
        return r;
    }

    __device__ int any(int v) { return __syncthreads_or(v); }
};

constexpr int kTeam = 256;

// ---------------------------------------------------------------- solves
__global__ __launch_bounds__(kTeam) void tci_solve_kernel(const SolveDesc *__restrict__ descs,
                                                          int count, int lds_stage_cap)
{
    const int s = blockIdx.x;
    if (s >= count) return;
    __shared__ EngShared S;
    __shared__ ArgBest red_a[2 * (kTeam / 64)];
    __shared__ double red_d[2 * (kTeam / 64)];
    extern __shared__ double stage[];
    const SolveDesc &D = descs[s];
    DevTeam<kTeam> tm;
    tm.tid = threadIdx.x;
    tm.nt = kTeam;
    tm.par_a = 0;
    tm.par_d = 0;
    tm.red_a = red_a;
    tm.red_d = red_d;
    if (D.mrcap <= lds_stage_cap) {
        tm.xs = stage;
        tm.ys = stage + D.mrcap;
    } else {
        tm.xs = D.stage;
        tm.ys = D.stage + D.mrcap;
    }
    Engine<DevTeam<kTeam>> eng(tm, D, S);
    eng.run();
}

// ------------------------------------------------------------------ rrLU
__global__ __launch_bounds__(kTeam) void tci_rrlu_kernel(tci_lu_job *__restrict__ jobs,
                                                         int64_t count, int lds_stage_cap,
                                                         double *__restrict__ gstage,
                                                         int64_t gstage_stride)
{
    const int64_t s = blockIdx.x;
    if (s >= count) return;
    __shared__ ArgBest red_a[2 * (kTeam / 64)];
    __shared__ double red_d[2 * (kTeam / 64)];
    extern __shared__ double stage[];
    tci_lu_job &J = jobs[s];
    const int m = (int)J.m, n = (int)J.n;
    const int mx = m > n ? m : n;
    DevTeam<kTeam> tm;
    tm.tid = threadIdx.x;
    tm.nt = kTeam;
    tm.par_a = 0;
    tm.par_d = 0;
    tm.red_a = red_a;
    tm.red_d = red_d;
    if (mx <= lds_stage_cap) {
        tm.xs = stage;
        tm.ys = stage + mx;
    } else {
        tm.xs = gstage + gstage_stride * s;
        tm.ys = tm.xs + mx;
    }
    double err, st[2];
    int p = rrlu_team(tm, J.A, m, n, J.maxrank, J.reltol, J.abstol, J.leftorth != 0, J.rowperm,
                      J.colperm, nullptr, &err, st);
    // 1-based permutations (rrLU.rowpermutation / colpermutation)
    for (int i = tm.tid; i < m; i += kTeam) J.rowperm[i] += 1;
    for (int j = tm.tid; j < n; j += kTeam) J.colperm[j] += 1;
    if (tm.tid == 0) {
        J.status = p < 0 ? -p : 0;
        J.npivot = p < 0 ? 0 : p;
        J.error = err;
    }
}

__global__ void tci_lu_unpack_kernel(const double *__restrict__ A, int64_t m, int64_t n, int64_t p,
                                     int leftorth, double *__restrict__ L, int64_t ldl,
                                     double *__restrict__ U, int64_t ldu)
{
    const int64_t nl = m * p, nu = p * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nl + nu;
         q += (int64_t)gridDim.x * blockDim.x) {
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

// ------------------------------------------------------------- launchers
hipError_t launch_solve(const SolveDesc *descs, int count, int max_mrcap, hipStream_t st)
{
    // staging vectors in LDS when 2·mrcap doubles fit in 48 KiB
    const int cap = 48 * 1024 / 16;
    const size_t lds = max_mrcap <= cap ? (size_t)2 * max_mrcap * sizeof(double) : 0;
    hipLaunchKernelGGL(tci_solve_kernel, dim3(count), dim3(kTeam), lds, st, descs, count,
                       max_mrcap <= cap ? max_mrcap : 0);
    return hipGetLastError();
}

hipError_t launch_rrlu(tci_lu_job *jobs, int64_t count, int64_t max_side, double *gstage,
                       int64_t gstage_stride, hipStream_t st)
{
    const int cap = 48 * 1024 / 16;
    const size_t lds = max_side <= cap ? (size_t)2 * max_side * sizeof(double) : 0;
    hipLaunchKernelGGL(tci_rrlu_kernel, dim3((unsigned)count), dim3(kTeam), lds, st, jobs, count,
                       max_side <= cap ? (int)max_side : 0, gstage, gstage_stride);
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
