// Launchers of the CDNA4 kernels (kernels.hip). Every launcher enqueues on the given stream and returns the HIP
// status of the launch.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "nppe_internal.h"

namespace nppe
{

// ---- batched launches -----------------------------------------------------------------------------------------
// While a LaunchRecorder is set on the calling thread, the launches of the kernels that have a batched form (the
// whole-row sweep, the whole-row and marching restrictions, the coarsest solve, the row sum of <r, z>, the whole-row
// FP64 Krylov passes and the separate finalisation) append a record -- the batched kernel, the system's grid and block, its arguments --
// instead of launching. A batched solve concatenates one record of each of several systems into one launch
// (launch_batched), whose block g - start[q] of system q does what that block does in the system's own launch.
// Kinds of records: a sweep or a restriction gives the same bits for any tiling, so a batch may re-tile it
// (retile_record); the other kernels' reductions follow their grid, which stays as recorded.
enum
{
    REC_FIXED = 0,
    REC_ROW_SWEEP = 1,
    REC_RESTRICT_ROW = 2,
    REC_RESTRICT_MARCH = 3,
    REC_SWEEP_RESTRICT = 4,
    REC_POST2 = 5,   // the post-smoothing pair in one march (its chunk length only)
    REC_PRE2 = 6     // the pre-smoothing and restriction in one march (its chunk length only, even)
};
struct LaunchRecord
{
    const void *fn = nullptr;   // the batched kernel: kernel(const Args *args, const int *start, int nsys)
    int grid = 0, threads = 0;
    int phase = 0, level = 0, seq = 0;   // where in the inner iteration (record_slot), then the order there
    int kind = REC_FIXED;
    int tm = 0, k = 0, flags = 0;        // whole-row kernels: tile rows, 64-lane steps per row, variant bits
    std::vector<unsigned char> args;     // the system's Args
};
struct LaunchRecorder
{
    int phase = 0, level = 0, seq = 0;
    std::vector<LaunchRecord> recs;
};
// Places of the inner iteration, in the order it runs them: the V-cycle's way down (level ascending), the coarsest
// solve, the way up (level descending), the Krylov update.
enum { SLOT_DOWN = 0, SLOT_COARSE = 1, SLOT_UP = 2, SLOT_KRYLOV = 3 };
// The record re-tiled: tm tile rows (whole-row sweeps; 0 keeps them) and p planes per chunk (made even for the
// restrictions, which sum plane pairs). False, and the record unchanged, for a fixed record or a tile shape the
// kernel has no form for.
bool retile_record(LaunchRecord &r, int tm, int p);
// Tile rows a batch may give the whole-row sweep record r.
void retile_rows(const LaunchRecord &r, std::vector<int> &tms);
void set_launch_recorder(LaunchRecorder *r);   // null: launch again
// Kernel-attached timing: the next whole-row sweep launched on this thread records start on its dispatch's start and
// stop on its end (hipExtLaunchKernel's events: the kernel's own duration, with no event packets around it) and
// clears the pair; launch_events_pending says whether a pair set here is still waiting for such a launch.
void set_launch_events(hipEvent_t start, hipEvent_t stop);
bool launch_events_pending();
void record_slot(int phase, int level);        // the place of the next records (nothing without a recorder)
// args: nsys records' Args back to back (the kernel's Args type); start: nsys + 1 block offsets, start[0] = 0,
// start[nsys] = grid, each system at least one block.
hipError_t launch_batched(hipStream_t s, const void *fn, int grid, int threads, const void *args, const int *start,
                          int nsys);

// ---- FP64 fine level ------------------------------------------------------------------------------------------
// box[order[q]] = src[q] (0 when src is null) for the n unknowns; with sumSlot >= 0 also sums src^2 into
// st->sums[sumSlot].
hipError_t launch_scatter64(hipStream_t s, int64_t n, const int32_t *order, const double *src, double *box,
                            SolveState *st, int sumSlot, double *partials);
// A solve without a guess: b64[order[q]] = src[q], r32[order[q]] = (float) src[q], where given r64[order[q]] =
// src[q] and x64[order[q]] = 0 (a pass begun by launch_krylov_init with residualIsB reads b and 0.0 for them), and
// the sum of src^2 into st->sums[0].
hipError_t launch_scatter_rhs64(hipStream_t s, int64_t n, const int32_t *order, const double *src, double *b64,
                                double *r64, float *r32, double *x64, SolveState *st, double *partials);
// dst[q] = box[order[q]] - (meanSlot >= 0 ? st->sums[meanSlot] / n : 0).
// gate: only once st->accepted (a true residual met the target).
hipError_t launch_gather64(hipStream_t s, int64_t n, const int32_t *order, const double *box, double *dst,
                           const SolveState *st, int meanSlot, int gate = 0);
// st->sums[slot] = sum of box[order[q]].
// gate: only once st->accepted, into sums[2].
hipError_t launch_sum64(hipStream_t s, int64_t n, const int32_t *order, const double *box, SolveState *st, int slot,
                        double *partials, int gate = 0);
// Outer refinement step: xn = x + e (e may be null) is written to xout (a different buffer), then r = b - A xn in
// FP64 with the reference's exact coefficients, stored as FP32 in r32 and/or FP64 in r64; st->sums[1] = ||r||^2.
// gate: only once the Krylov loop is done (st->done); then also st->accepted = skip or ||r||^2 < acc2, and with
// sumX (3D) st->sums[2] = the sum of x over the unknowns (the mean projection's sum, in the same pass).
hipError_t launch_update_residual64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty,
                                    const double *x, const float *e, double *xout, const double *b, float *r32,
                                    double *r64, SolveState *st, double *partials, int gate = 0, int sumX = 0);
// y = A x in FP64 (the reference's SpMV, row sums in ascending column order).
hipError_t launch_apply64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *x,
                          double *y);
// FP64 Jacobi-PCG kernels (Eigen's ConjugateGradient loop, ConjugateGradient.h:57-88).
hipError_t launch_cg64_init(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *r,
                            double *p, SolveState *st, double *partials);
hipError_t launch_cg64_spmv(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *p,
                            double *q, SolveState *st, double *partials);
hipError_t launch_cg64_update(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, double *x,
                              double *r, const double *p, const double *q, SolveState *st, double *partials);
hipError_t launch_cg64_direction(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty,
                                 const double *r, double *p, SolveState *st);

// ---- FP32 multigrid ------------------------------------------------------------------------------------------
// One full red-black Gauss-Seidel sweep x_out = S(x_in) (c1: colour updated first; zero: x_in taken as 0), in one
// pass: tiles of 64 x tileRows (8 or 16) cells marching along the slow axis in chunks of planesPerChunk planes.
// With rzPartials non-null the sweep is the CG's last level-0 sweep and also reduces <b, x_out> into st (as
// launch_cg_rz(b, x_out) does). The whole-row sweep (rbgs_family RBGS_ROW) writes one partial per interior row and
// plane, which launch_rz_rows_finalize then sums; the other sweeps complete the reduction themselves.
// With prolongFrom non-null the sweep first takes x_in += P x_c (x_c = prolongXc on level prolongFrom): folded
// into the whole-row sweep's loads (x_in itself is not written), or as a separate prolongation pass before it.
hipError_t launch_rbgs(hipStream_t s, const Level32 &L, const float *xin, float *xout, const float *b, int c1,
                       int zero, int tileRows, int planesPerChunk, SolveState *st, double *rzPartials = nullptr,
                       const Level32 *prolongFrom = nullptr, const float *prolongXc = nullptr);
hipError_t launch_rz_rows_finalize(hipStream_t s, const Box &B, SolveState *st, double *partials);
// Both post-smoothing sweeps of a level in one march: x_out = S(S(x_in + P x_c)) (x_c on level Lc; colour c1 first
// in both), the second sweep three planes behind the first. x_out (another buffer than x_in) is bit for bit that of
// the two launch_rbgs calls.
// rbgs2_level says whether a level has it (3D, uniform, rows of at most 128 cells, the whole-row family;
// NPPE_FUSE_POST=0 turns it off: the tests compare the two).
bool rbgs2_level(const Level32 &L, const float *xin, const float *xout, const float *b);
int64_t rbgs2_blocks(const Box &B, int planesPerChunk);
hipError_t launch_rbgs2(hipStream_t s, const Level32 &L, const float *xin, float *xout, const float *b, int c1,
                        int planesPerChunk, SolveState *st, const Level32 &Lc, const float *xc);
// Which sweep kernel launch_rbgs runs for these operands, and the tile heights it accepts for it (every tile shape
// gives bit-identical results: a colour's updates are independent of each other).
enum { RBGS_ROW = 0, RBGS_MARCH = 1, RBGS_2D = 2 };
int rbgs_family(const Level32 &L, const float *xin, const float *xout, const float *b);
void rbgs_tile_candidates(int family, std::vector<int> &tileRows);
// Blocks of one sweep launch of that family and tile shape.
// (folding: the sweep that folds the prolongation in, whose tiles the whole-row family picks on its own)
int64_t rbgs_blocks(const Box &B, int family, int tileRows, int planesPerChunk, bool folding = false,
                    bool uniform = true);
// The tile rows the whole-row sweep runs for the tile height asked.
int rbgs_row_tile(const Box &B, int tileRows, bool folding, bool uniform);
hipError_t launch_residual_restrict(hipStream_t s, const Level32 &Lf, const float *x, const float *b,
                                    const Level32 &Lc, float *bc, const SolveState *st);
// A plain whole-row sweep x_out = S(x_in) and the restriction of its residual b - A x_out into bc in one march
// (the last pre-smoothing sweep of a uniform 3D level with rows of at most 256 cells): x_out and bc bit for bit
// those of launch_rbgs followed by launch_residual_restrict. Tiles: the largest of rbgs_restrict_rows at most
// tileRows, planesPerChunk made even. rbgs_restrict_fusable says whether the level has it (NPPE_FUSE_RESTRICT=0
// turns it off: the tests compare the two).
bool rbgs_restrict_fusable(const Level32 &Lf, const float *xin, const float *xout, const float *b, const float *bc);
void rbgs_restrict_rows(int k, std::vector<int> &tms);
hipError_t launch_rbgs_restrict(hipStream_t s, const Level32 &Lf, const float *xin, float *xout, const float *b,
                                int c1, const Level32 &Lc, float *bc, int tileRows, int planesPerChunk,
                                const SolveState *st);
int64_t rbgs_restrict_blocks(const Box &B, int tileRows, int planesPerChunk);
// Both pre-smoothing sweeps from x = 0 and the restriction of the residual in one march: x_out = S(S_0(b)) (colour c1
// first in both) and bc = R (b - A x_out), bit for bit those of launch_rbgs (zero guess) followed by
// launch_rbgs_restrict; x_out is the only x written or read. rbgs_pre_restrict_level says whether a level has it
// (rows of at most 128 cells where rbgs_restrict_fusable holds; NPPE_FUSE_PRE=0 turns it off: the tests compare the
// two); planesPerChunk is made even.
bool rbgs_pre_restrict_level(const Level32 &Lf, const float *xin, const float *xout, const float *b, const float *bc);
int64_t rbgs_pre_restrict_blocks(const Box &B, int planesPerChunk);
hipError_t launch_rbgs_pre_restrict(hipStream_t s, const Level32 &Lf, float *xout, const float *b, int c1,
                                    const Level32 &Lc, float *bc, int planesPerChunk, const SolveState *st);
hipError_t launch_prolong(hipStream_t s, const Level32 &Lf, float *x, const Level32 &Lc, const float *xc,
                          const SolveState *st);
// Coarsest level: x[order[i]] = sum_j M[i][j] b[order[j]] in FP64 (M: the dense pseudo-inverse, n x n).
hipError_t launch_coarse_solve(hipStream_t s, int n, const int32_t *order, const double *M, const float *b, float *x,
                               const SolveState *st);

// ---- FP32 CG around the V-cycle ------------------------------------------------------------------------------
// <r, z> -> beta (0 on the first direction), rz.
hipError_t launch_cg_rz(hipStream_t s, const Level32 &L, const float *r, const float *z, SolveState *st,
                        double *partials);
// p = z + beta p (p = z on the first direction).
hipError_t launch_cg_direction(hipStream_t s, const Level32 &L, const float *z, float *p, SolveState *st);
// q = A p, <p, q> -> alpha.
hipError_t launch_cg_spmv(hipStream_t s, const Level32 &L, const float *p, float *q, SolveState *st,
                          double *partials);
// Whether launch_cg_dirspmv serves this level (3D boxes); otherwise launch_cg_direction + launch_cg_spmv.
bool cg_fused_direction(const Level32 &L);
// p = z + beta p and q = A p with <p, q> -> alpha in one pass; p is read from one of (pA, pB) and written to the
// other (pA written on even inner iterations).
hipError_t launch_cg_dirspmv(hipStream_t s, const Level32 &L, const float *z, float *pA, float *pB, float *q,
                             SolveState *st, double *partials);
// FP64 Krylov around the FP32 preconditioner. The directions live in a ring of kXRing FP32 box vectors (slot j at
// ring + j * stride; stride a multiple of 16): iteration k writes slot k mod kXRing. Part 1: p = z + beta p formed
// in FP64, stored in FP32, <p, A p> -> alpha (kept in st->alphas too), A p in the reference's FP64 arithmetic.
hipError_t launch_cg_dir64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const float *z,
                           float *ring, int64_t stride, SolveState *st, double *partials);
// Part 2: q = A p from the same FP32 p, r -= alpha q (FP64), r32 = (float) r, <r, r> -> convergence flag; on the
// iteration that fills the ring (k mod kXRing = kXRing - 1) also x += alpha_j p_j for the ring's iterations, oldest
// first -- the per-iteration updates, in their order.
// b: the right-hand side, read for r on the first update of a pass that began at x = 0 (SolveState::zeroStart, the
// residual then still b); x is read as 0.0 by the first full ring of such a pass.
hipError_t launch_cg_update64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const float *ring,
                              int64_t stride, double *x, double *r, float *r32, const double *b, SolveState *st,
                              double *partials);
// The rest of the deferred update once the recursion is done: x += alpha_j p_j for the slots j < iters mod kXRing
// (the iterations since the last full ring), at the unknowns. Nothing before st->done. In a pass that began at
// x = 0 (SolveState::zeroStart) x is not read while no full ring has been applied (the same sums from 0.0), and a
// pass without iterations writes that 0.
hipError_t launch_flush_x(hipStream_t s, const Box &B, const Tables64 &T, const float *ring, int64_t stride, double *x,
                          const SolveState *st);
// e += alpha p (e = alpha p on the first iteration), r -= alpha q, <r, r> -> convergence flag; p is pA on even
// inner iterations and pB on odd ones (pA == pB for an in-place direction).
hipError_t launch_cg_update(hipStream_t s, const Level32 &L, float *e, float *r, const float *pA, const float *pB,
                            const float *q, SolveState *st, double *partials);
// Inner-solve set-up from the outer residual norm st->sums[1]: thr2 = red^2 * sums[1], counters cleared.
hipError_t launch_inner_setup(hipStream_t s, SolveState *st, double reduction, int maxIters);
// FP64 CG set-up: thr2, iteration cap, counters and the accepted flag cleared.
hipError_t launch_state_init(hipStream_t s, SolveState *st, double thr2, int maxIters);
// The FP64 Krylov set-up from |b|^2 in st->sums[0] and the guess's true residual |r|^2 in st->sums[1] (residualIsB:
// the residual is b itself, no guess, and sums[1] is not read).
hipError_t launch_krylov_init(hipStream_t s, SolveState *st, double tol, double ref, double aimFactor, int maxIters,
                              int residualIsB);

}  // namespace nppe
