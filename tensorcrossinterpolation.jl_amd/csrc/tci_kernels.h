// tci_kernels.h — launchers of the gfx950 kernels (defined in tci_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tci.h"
#include "tci_desc.h"

namespace tci {

// on-chip (LDS) placement of one solve workgroup's hot buffers
struct LdsPlan {
    int32_t stage_cap;   // rrLU staging vectors: 2 × stage_cap doubles (0: in HBM)
    int32_t tab_cap;     // bytes per index table (rows, cols) (0: in HBM)
    int32_t sets_cap;    // bytes for the pivot sets and their counts (0: in HBM)
    int32_t lut_cap;     // doubles for the Lorentzian value table (0: none)
    int64_t pi_cap;      // doubles available for a Pi matrix (0: in HBM)
};
size_t lds_plan_bytes(const LdsPlan &p);
// static LDS bytes of the solve kernel variant for (team, functor kind), -1 on error
int solve_static_lds(int team, int fkind);
hipError_t launch_solve(const SolveDesc *descs_device, int count, int team, int fkind,
                        const LdsPlan &plan, hipStream_t st);
// slot-level serving: the queue's control words in coherent host memory
struct ServeCtl {
    uint64_t tail;   // host: tickets published (ring entries valid below it)
    uint32_t stop;   // host: 1 = leave once idle
    uint32_t pad;
};
// the serving grid's own words in device memory (zero when the server starts)
struct ServeDev {
    uint64_t head;       // next ticket to claim
    uint64_t dtail;      // mirror of ServeCtl::tail (only ever raised)
    uint64_t lastpoll;   // wall clock of the last read of the host's words
    uint64_t last;       // wall clock of the last claim or finished solve
    uint32_t busy;       // workgroups running a solve
    uint32_t closing;    // the launch epoch that is leaving for lack of work
    uint32_t dstop;      // the launch epoch that has seen ServeCtl::stop
    uint32_t pad;
};
// one published request (a ring entry in coherent host memory).  The
// claiming workgroup copies the staged image (the descriptor, then the
// inputs) into the device block, runs the solve, and copies the report out
// before it sets the done mark
struct ServeSlot {
    const uint8_t *src;    // the staged image (device alias of pinned host memory)
    uint8_t *dst;          // the request's device block: the descriptor at dst
    const uint8_t *rep;    // the report in the device block
    uint8_t *rep_out;      // its host copy (device alias of pinned host memory)
    uint64_t bytes;        // staged bytes, a multiple of 16
    uint64_t rep_bytes;    // report bytes, a multiple of 16
};
int serve_static_lds(int team);
// a persistent grid of `blocks` workgroups (generic functor kernel) that runs
// published tickets; it leaves, all of it, once stop is set and every
// published ticket is claimed, or once no solve has run, been claimed or
// finished for idle_ticks of the wall clock.  epoch: this launch's number
// (each launch a new one, from 1)
hipError_t launch_serve(const ServeSlot *ring, ServeDev *g, const ServeCtl *ctl, uint64_t *done, int cap,
                        const LdsPlan &plan, uint64_t idle_ticks, uint32_t epoch, int team, int blocks,
                        hipStream_t st);
// order (device, count entries, or null): the job index each workgroup takes
hipError_t launch_rrlu(tci_lu_job *jobs_device, int64_t count, int64_t max_side, double *gstage,
                       int64_t gstage_stride, const int32_t *order, hipStream_t st);
// the register-tile rrLU of the jobs idx[0..count) (each m, n <= 128);
// scratch + soff[q]: m·mr + mr·n doubles for job idx[q] (mr = min(m, n, maxrank))
constexpr int kTileSide = 128;
hipError_t launch_rrlu_tile(tci_lu_job *jobs_device, const int32_t *idx, int64_t count, double *scratch,
                            const int64_t *soff, hipStream_t st);
// rrlu!(A) of one matrix (each side <= kGridMaxSide) over a cooperative grid
// of `blocks` workgroups (one per CU, rrlu_grid_blocks); the packed result
// goes to out (m×n), the job's permutations, npivot, error and status are
// written; cand: rrlu_grid_cand_bytes(blocks) bytes; flags: one int32, zero
// on entry
constexpr int kGridMaxSide = 8192;
size_t rrlu_grid_lds(int64_t m, int64_t n);
// team: 256, 512 or 1024 threads (anything else: the default,
// TCI_RRLU_GRID_TEAM or 1024); the cooperative grid size for that LDS (one
// workgroup per CU), 0 if none fits
int rrlu_grid_default_team();
int rrlu_grid_blocks(int device, size_t lds, int team);
size_t rrlu_grid_cand_bytes(int blocks);
hipError_t launch_rrlu_grid(tci_lu_job *job, int64_t m, int64_t n, double *out, void *cand, int32_t *flags,
                            int blocks, int team, hipStream_t st);
// One whole TT sweep of factorizations (the host-driven sweep's steps, the
// same operations per entry) in a single workgroup, for trains whose bonds
// are small enough that per-bond host round trips would dominate
struct TtSweepArgs {
    int32_t L, method, forward, stage_cap;   // stage_cap: doubles per LDS staging vector
    int32_t mat_cap;                         // doubles of LDS for the bond matrix (0: factored in place)
    int64_t maxrank, R;                      // R: the largest bond on entry (pivot slots per bond: 2R)
    double reltol, abstol;
    int64_t *links;                          // device, L+1; each bond's new dimension written back
    const int32_t *d;                        // device, L
    const int64_t *off;                      // device, L: each site's slot in cores (doubles)
    double *cores, *wl, *wu, *wc, *wp, *gstage;
    int32_t *perm;                           // device, m+n of the largest bond
    int32_t *piv;                            // device, (L-1)·2R: bond b's row, then column pivots (1-based)
    double *pe;                              // device, R+1: running pivot errors (zero on entry)
    int32_t *status;                         // device, 2: status and bond of a NaN verdict (zero on entry)
};
hipError_t launch_tt_sweep(const TtSweepArgs &a, hipStream_t st);
// C = A·B with each sum in ascending k (M×K · K×N, column-major)
hipError_t launch_gemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                       double *Cm, int64_t ldc, hipStream_t st);
// left(luci) (m×p, ld ldl) and right(luci) (p×n, ld ldr) of MatrixLUCI from
// the unpacked L (m×p) and U (p×n) and the 1-based permutations
hipError_t launch_luci(int64_t m, int64_t n, int64_t p, int leftorth, const int32_t *rowperm, const int32_t *colperm,
                       const double *L, const double *U, double *left, int64_t ldl, double *right, int64_t ldr,
                       hipStream_t st);
// left(lu)/right(lu) with the permutations applied, from the unpacked L, U
hipError_t launch_lu_permute(int64_t m, int64_t n, int64_t p, const int32_t *rowperm, const int32_t *colperm,
                             const double *L, const double *U, double *left, int64_t ldl, double *right, int64_t ldr,
                             hipStream_t st);
hipError_t launch_lu_unpack(const double *A, int64_t m, int64_t n, int64_t p, int leftorth,
                            double *L, int64_t ldl, double *U, int64_t ldu, hipStream_t st);
hipError_t launch_kron_center(const uint8_t *left, int64_t nI, int nl, const int32_t *cdims, int M,
                              int64_t nc, uint8_t *rowsx, int stride, hipStream_t st);
int eval_blocks(int64_t tot);
// TT evaluation at npts points (and f, |TT - f| and |f| maxima when fkind != 0);
// blockmax holds 2 doubles per block of tt_eval_points_per_block(chimax) points
int tt_eval_points_per_block(int chimax);
hipError_t launch_tt_eval(int L, const int64_t *dims3, const int64_t *coff, const double *cores,
                          const uint8_t *pts, int64_t npts, int chimax, int fkind, const double *params,
                          const int32_t *iparams, double *out_tt, double *out_f, double *blockmax,
                          hipStream_t st);
// _floatingzone from each of npts points (in/out, npts×L digits), one
// workgroup per point; work: wstride doubles per workgroup, digits: dmax·L
// bytes per workgroup; roff: L+1 offsets of the right environments
hipError_t launch_floatingzone(int L, const int64_t *dims3, const int64_t *coff, const double *cores,
                               const int64_t *roff, int chimax, int dmax, int fkind, const double *params,
                               const int32_t *iparams, uint8_t *pts, double *errs, int64_t npts, int64_t nsweeps,
                               double earlystoptol, double *work, int64_t wstride, uint8_t *digits, int blocks,
                               hipStream_t st);
hipError_t launch_eval(int kind, const double *params, const int32_t *iparams, int L,
                       const uint8_t *rows, int64_t nrow, int rstride, int nl, const uint8_t *cols,
                       int64_t ncol, int cstride, double *out, double *blockmax, hipStream_t st);

}  // namespace tci
