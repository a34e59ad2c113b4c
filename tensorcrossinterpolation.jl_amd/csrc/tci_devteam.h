// tci_devteam.h — the gfx950 workgroup as a "Team" of the solve engine
// (tci_engine.h): team barriers and reductions, wave-level DPP reductions and
// ballots, and the LDS slots of the register-tile rrLU (tci_regtile.h).
#pragma once
#include <hip/hip_runtime.h>

#include "tci_common.h"

namespace tci {

// register-tile shape per team size (tci_regtile.h): wave w, lane l holds rows
// l + 64·i (i < RA) of columns w + NW·j (j < CB); capacity (64·RA)×(NW·CB)
template <int NT> struct TileShape;
template <> struct TileShape<64> { static constexpr int RA = 1, CB = 64; };     // 64×64 (one wave)
template <> struct TileShape<256> { static constexpr int RA = 2, CB = 32; };    // 128×128
template <> struct TileShape<512> { static constexpr int RA = 2, CB = 16; };    // 128×128
template <> struct TileShape<1024> { static constexpr int RA = 2, CB = 8; };    // 128×128

// one 64-bit value moved across lanes by a DPP pattern (both halves)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
// DPP controls: quad_perm [1,0,3,2] and [2,3,0,1], half-row mirror, row mirror
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

// WPE (waves per SIMD the kernel is compiled for) is part of the type so
// that every kernel variant gets its own copy of the engine's functions,
// register-allocated for that variant's budget
template <int NT_, int WPE_ = 2>
struct DevTeam {
    static constexpr int NT = NT_, WS = 64, NW = NT / 64, WPE = WPE_;
    static constexpr int RA = TileShape<NT>::RA, CB = TileShape<NT>::CB;
    static constexpr int TMR = WS * RA;   // rows of one wave's staged column
    // LDS of the register-tile rrLU (one set per kernel instantiation)
    __device__ static double *lds_a() { __shared__ double b[NW]; return b; }
    __device__ static uint32_t *lds_b() { __shared__ uint32_t b[NW]; return b; }
    __device__ static double *lds_x() { __shared__ double b[NW * TMR]; return b; }
    __device__ static double *lds_y() { __shared__ double b[NW * CB]; return b; }
    __device__ double &tA(int w) const { return lds_a()[w]; }
    __device__ uint32_t &tB(int w) const { return lds_b()[w]; }
    __device__ double &tX(int w, int r) const { return lds_x()[w * TMR + r]; }
    __device__ double &tY(int w, int j) const { return lds_y()[w * CB + j]; }

    // wave-uniform max of non-NaN doubles
    __device__ double wave_max(double v) const
    {
        v = fmax(v, dpp_f64<kDppXor1>(v));
        v = fmax(v, dpp_f64<kDppXor2>(v));
        v = fmax(v, dpp_f64<kDppHalfMirror>(v));
        v = fmax(v, dpp_f64<kDppMirror>(v));
        return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)),
                    fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
    }
    __device__ uint32_t wave_min(uint32_t v) const
    {
        v = min(v, dpp_u32<kDppXor1>(v));
        v = min(v, dpp_u32<kDppXor2>(v));
        v = min(v, dpp_u32<kDppHalfMirror>(v));
        v = min(v, dpp_u32<kDppMirror>(v));
        const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
        const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
        return min(min(a, b), min(c, d));
    }
    __device__ bool wave_any(bool p) const { return __ballot(p) != 0; }
    // lanes where p holds, as a wave-uniform bit mask
    __device__ uint64_t wave_ballot(bool p) const { return __ballot(p); }
    // lane l's v in every lane (l wave-uniform)
    __device__ int wave_bcast(int v, int l) const { return __builtin_amdgcn_readlane(v, l); }
    __device__ double wave_bcast_f64(double v, int l) const { return readlane_f64(v, l); }
    // lowest lane where p holds, -1 if none (wave-uniform)
    __device__ int wave_first(bool p) const
    {
        const unsigned long long b = __ballot(p);
        return b ? (int)__builtin_ctzll(b) : -1;
    }
    // as wave_sync, for LDS only (does not wait for global memory operations)
    __device__ void wave_sync_lds() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
    // makes this wave's earlier LDS and global writes visible to its later
    // reads in any lane (no barrier with the other waves)
    __device__ void wave_sync() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // thread coordinates come from the hardware registers, so a team passed
    // by reference costs no memory traffic to read them
    __device__ int tid() const { return (int)threadIdx.x; }
    __device__ static constexpr int nt() { return NT; }
    __device__ int lane() const { return (int)(threadIdx.x & 63); }
    __device__ static constexpr int wsize() { return 64; }
    __device__ int wave() const { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }
    __device__ static constexpr int nwaves() { return NW; }
    double *xs, *ys;
    bool stage_lds;       // xs and ys are in LDS
    double *lpi;          // LDS area for Pi matrices
    int64_t lpi_cap;      // its size in doubles (0: none)
    uint8_t *lrows, *lcols;
    uint8_t *lsets;       // LDS pivot sets (null: in HBM)
    int32_t *lsetn;       // LDS set counts
    double *llut;         // LDS table of Lorentzian values (null: none)
    int64_t llut_cap;
    // team-reduction slots, double-buffered by par_a / par_d
    __device__ static ArgBest *lds_ra() { __shared__ ArgBest b[2 * NW]; return b; }
    __device__ static double *lds_rd() { __shared__ double b[2 * NW]; return b; }
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
        ArgBest *slot = lds_ra() + par_a * NW;
        if (lane() == 0) slot[wave()] = b;
        __syncthreads();
        ArgBest r = slot[0];
#pragma unroll
        for (int w = 1; w < NW; ++w)
            if (arg_better(slot[w], r)) r = slot[w];
        par_a ^= 1;
        return r;
    }

    // NaN-propagating max over the team (the wave part by DPP, as wave_max)
    __device__ double max_nan(double v)
    {
        v = jl_max(v, dpp_f64<kDppXor1>(v));
        v = jl_max(v, dpp_f64<kDppXor2>(v));
        v = jl_max(v, dpp_f64<kDppHalfMirror>(v));
        v = jl_max(v, dpp_f64<kDppMirror>(v));
        v = jl_max(jl_max(readlane_f64(v, 0), readlane_f64(v, 16)),
                   jl_max(readlane_f64(v, 32), readlane_f64(v, 48)));
        double *slot = lds_rd() + par_d * NW;
        if (lane() == 0) slot[wave()] = v;
        __syncthreads();
        double r = slot[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) r = jl_max(r, slot[w]);
        par_d ^= 1;
        return r;
    }

    __device__ int uniform(int x) const { return __builtin_amdgcn_readfirstlane(x); }
    __device__ uint32_t uniform_u32(uint32_t x) const { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
    __device__ double uniform_f64(double x) const { return readlane_f64(x, 0); }

    __device__ int any(int v) { return __syncthreads_or(v); }
    // a promise that p points into LDS (the compiler then emits LDS
    // instructions instead of generic ones for accesses through it); it
    // must hold
    __device__ static void assume_lds(const void *p)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        __builtin_assume(__builtin_amdgcn_is_shared(p));
#endif
    }
    __device__ uint64_t clock() const { return wall_clock64(); }
    __device__ void init(double *x, double *y)
    {
        xs = x;
        ys = y;
        stage_lds = false;
        par_a = 0;
        par_d = 0;
        lpi = nullptr;
        lpi_cap = 0;
        lrows = nullptr;
        lcols = nullptr;
        lsets = nullptr;
        lsetn = nullptr;
        llut = nullptr;
        llut_cap = 0;
    }
};

}  // namespace tci
