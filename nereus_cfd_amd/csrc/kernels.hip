// CDNA4 (gfx950) kernels of the pressure Poisson backend.
//
// Every stencil kernel walks the interior of one level's box in 64 (fast) x 4 (mid) tiles of one slow plane: a
// wavefront covers 64 consecutive cells of a row, so each load and store is one coalesced 256 B (FP32) or 512 B
// (FP64) line pair. Block ids are remapped so that each of the eight XCDs gets a contiguous run of planes and the
// i-1 / i+1 neighbour planes of a tile are in the same XCD's L2. Reductions are deterministic: each block writes a
// partial sum, and the last block to finish (ticket from one atomic counter) adds the partials in index order and
// publishes the scalar, so a repeated solve is bit-identical.
//
// The operator is the reference's matrix, built on the fly from the cell types of the cell and its six neighbours
// and from 1D spacing tables (PressureSolver3D.cpp:47-87, :145-170; PressureSolver2D.cpp:117-143):
//   FLUID neighbour             -> diag -= c; off-diagonal c when the neighbour is an unknown (not on the ring)
//   EMPTY / PRESSURE neighbour  -> diag -= c * rf, rf = FaceGap / (Width/2) on the domain ring, 1 inside
//   SOLID / INFLOW / OUTFLOW    -> nothing (Neumann)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <type_traits>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "kernels.h"
#include "op64.h"

namespace nppe
{
namespace
{

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}


// A reduction's destination: the block partials, the state its finalisation updates, the FinOp to apply, and
// whether the kernel finalises itself (fused) or a one-block k_finalize launch follows it.
struct Red
{
    double *partials;
    SolveState *st;
    int op;
    int fused;
};

__device__ void fin_apply(SolveState *st, double *partials, int nblk, int nv, int op, bool coherent);

// Block sum of NV per-thread values, written by thread 0 to partials[a * nblk + blk]. Fused, the last block of
// the grid to finish (a ticket counter in the state) then sums the partials of all blocks in index order and
// applies the FinOp (fin_apply); every thread of every block of the grid must call it. There the partials are
// agent-scope atomic stores (written through to the device coherence point), each completed before its block
// takes a ticket, and the last block reads them with agent-scope atomic loads; a device-scope fence instead would
// write back the whole L2 of the XCD from every block. The completion wait lengthens every block, so large grids
// leave the finalisation to a k_finalize launch. Values are never combined atomically: every reduction is
// deterministic.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], const Red &red, int nblk, int blk)
{
    __shared__ double sh[NV][16];
    __shared__ int lastBlock;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int a = 0; a < NV; a++)
    {
        const double x = wave_sum(v[a]);
        if (lane == 0) sh[a][w] = x;
    }
    __syncthreads();
    if (!red.fused)
    {
        if (tid == 0)
#pragma unroll
            for (int a = 0; a < NV; a++)
            {
                double x = 0.0;
                for (int q = 0; q < nw; q++) x += sh[a][q];
                red.partials[(size_t)a * nblk + blk] = x;
            }
        return;
    }
    if (tid == 0)
    {
#pragma unroll
        for (int a = 0; a < NV; a++)
        {
            double x = 0.0;
            for (int q = 0; q < nw; q++) x += sh[a][q];
            __hip_atomic_store(&red.partials[(size_t)a * nblk + blk], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __builtin_amdgcn_s_waitcnt(0);   // the partial stores have completed
        lastBlock = __hip_atomic_fetch_add(&red.st->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                    (unsigned)(nblk - 1);
    }
    __syncthreads();
    if (!lastBlock) return;
    fin_apply(red.st, red.partials, nblk, NV, red.op, true);
}

// Logical block id after the XCD remap: consecutive hardware workgroup ids go round-robin to the eight XCDs, so
// XCD x receives ids x, x+8, ...; those are given the x-th contiguous share of the logical blocks, and each
// logical block a contiguous run of tiles -- one XCD sweeps a contiguous run of planes.
__device__ __forceinline__ int logical_block()
{
    const int b = blockIdx.x, nb = gridDim.x;
    const int xcd = b & 7, r = b >> 3, per = nb >> 3, extra = nb & 7;
    return xcd < extra ? xcd * (per + 1) + r : extra * (per + 1) + (xcd - extra) * per + r;
}

struct TileRange
{
    int begin, end, lb;
};

__device__ __forceinline__ TileRange my_tiles(int ntiles)
{
    const int lb = logical_block();
    const int G = gridDim.x;
    return TileRange{(int)((int64_t)lb * ntiles / G), (int)((int64_t)(lb + 1) * ntiles / G), lb};
}

// Interior cell of this thread in tile t; false outside the interior.
__device__ __forceinline__ bool tile_cell(const Box &B, int t, int &s, int &m, int &f)
{
    const int tf = t % B.tilesF;
    const int t2 = t / B.tilesF;
    const int tm = t2 % B.tilesM;
    const int tp = t2 / B.tilesM;
    f = 1 + tf * kBlockFast + (threadIdx.x & (kBlockFast - 1));
    m = 1 + tm * kBlockMid + (threadIdx.x / kBlockFast);
    s = B.threeD ? 1 + tp : 0;
    return f <= B.nf - 2 && m <= B.nm - 2;
}

// ---- FP32 operator -------------------------------------------------------------------------------------------
// Coefficients of the cell at (s, m, f) in direction order slow+, slow-, mid+, mid-, fast+, fast-: c[q] for a
// coupled neighbour (coupled bit q set), and the diagonal (negative, or 0 for an isolated cell). Built from the
// stencil code: the common cell (no Dirichlet neighbour) needs only its code8 byte; a cell beside a Dirichlet
// cell reads its code16 word.
struct Op32
{
    float c[6];
    float diag;
    unsigned coupled;
};

__device__ __forceinline__ Op32 op32(const Level32 &L, int s, int m, int f, int64_t idx, unsigned code)
{
    Op32 o;
    o.coupled = code & 63u;
    float ring[6];
    if (L.uniform)
    {
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            o.c[q] = L.cu;
            ring[q] = L.cr;
        }
    }
    else
    {
        const Tables32 &T = L.T;
        const float wS = T.wS[s], wM = T.wM[m], wF = T.wF[f];
        const float aS = wM * wF, aM = wS * wF, aF = wS * wM;
        o.c[0] = aS * T.gS[s];
        o.c[1] = aS * T.gS[s - (L.B.threeD ? 1 : 0)];
        o.c[2] = aM * T.gM[m];
        o.c[3] = aM * T.gM[m - 1];
        o.c[4] = aF * T.gF[f];
        o.c[5] = aF * T.gF[f - 1];
        ring[0] = ring[1] = aS * T.hS[s];
        ring[2] = ring[3] = aM * T.hM[m];
        ring[4] = ring[5] = aF * T.hF[f];
    }
    float diag = 0.f;
    if (!(code & 128u))
    {
#pragma unroll
        for (int q = 0; q < 6; q++) diag -= ((o.coupled >> q) & 1u) ? o.c[q] : 0.f;
    }
    else
    {
        const unsigned c16 = L.c16[idx];
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned st = (c16 >> (2 * q)) & 3u;
            diag -= st == 3u ? ring[q] : (st != 0u ? o.c[q] : 0.f);
        }
    }
    o.diag = diag;
    return o;
}

__device__ __forceinline__ float offsum32(const Box &B, const Op32 &o, const float *x, int64_t idx)
{
    const int64_t off[6] = {B.sS, -B.sS, B.sM, -B.sM, 1, -1};
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 6; q++)
    {
        if (q < 2 && !B.threeD) continue;
        const float v = x[idx + off[q]];
        acc += ((o.coupled >> q) & 1u) ? o.c[q] * v : 0.f;
    }
    return acc;
}

// 1/n for the n coupled neighbours of an interior cell of a uniform level (no Dirichlet neighbour): there the
// diagonal is -cu n, so a Gauss-Seidel update is (sum of coupled x - b / cu) / n and A x = cu (sum - n x).
// Selected in registers rather than read from a table: a table in memory is a dependent load in every update.
__device__ __forceinline__ float inv_count(unsigned code)
{
    const unsigned n = __popc(code & 63u);
    float r = n == 1u ? 1.f : 0.f;
    r = n == 2u ? 0.5f : r;
    r = n == 3u ? 1.f / 3.f : r;
    r = n == 4u ? 0.25f : r;
    r = n == 5u ? 0.2f : r;
    r = n == 6u ? 1.f / 6.f : r;
    return r;
}

__device__ __forceinline__ bool fast_cell(const Level32 &L, unsigned code) { return L.uniform && !(code & 128u); }

// Sum of the coupled neighbours' x (uniform fast path).
__device__ __forceinline__ float nbsum32(const Box &B, unsigned code, const float *x, int64_t idx)
{
    float acc = 0.f;
    if (B.threeD)
    {
        acc += (code & 1u) ? x[idx + B.sS] : 0.f;
        acc += (code & 2u) ? x[idx - B.sS] : 0.f;
    }
    acc += (code & 4u) ? x[idx + B.sM] : 0.f;
    acc += (code & 8u) ? x[idx - B.sM] : 0.f;
    acc += (code & 16u) ? x[idx + 1] : 0.f;
    acc += (code & 32u) ? x[idx - 1] : 0.f;
    return acc;
}

// (A x)_idx in FP32.
__device__ __forceinline__ float apply32(const Level32 &L, int s, int m, int f, int64_t idx, unsigned code,
                                         const float *x)
{
    if (fast_cell(L, code)) return L.cu * (nbsum32(L.B, code, x, idx) - (float)__popc(code & 63u) * x[idx]);
    const Op32 o = op32(L, s, m, f, idx, code);
    return offsum32(L.B, o, x, idx) + o.diag * x[idx];
}

// ---- scalar finalisation -----------------------------------------------------------------------------------

// What the last block of a reducing kernel does with the block partials (fin_apply).
enum FinOp
{
    FIN_SUM0 = 0,     // sums[0] = t0
    FIN_SUM1,         // sums[1] = t0
    FIN_SUM2,         // sums[2] = t0
    FIN_TRUE_RES,     // only after the Krylov loop is done: sums[1] = t0; accepted = skip or t0 < acc2
    FIN_SUM2_ACC,     // only once a true residual was accepted: sums[2] = t0
    FIN_TRUE_RES_SUM, // FIN_TRUE_RES, and sums[2] = t1 (the sum of x over the unknowns, for the mean projection)
    FIN_RZ,           // beta = t0 / rz (0 on the first direction); rz = t0
    FIN_PQ,           // alpha = rz / t0; first = 0
    FIN_RR,           // rr = t0; iters++; done when rr < thr2 (1) or at the cap (2)
    FIN_CG64_INIT,    // rz = t0; iters = 0
    FIN_CG64_PQ,      // alpha = rz / t0
    FIN_CG64_UPDATE   // rr = t0; converged -> done = 1 (not counted, as Eigen); else beta = t1 / rz, rz = t1, iters++
};

__device__ __forceinline__ bool fin_in_loop(int op) { return op >= FIN_RZ; }

// Run by every thread of one block (the last block of a fused reducing kernel, or k_finalize): sums the nblk
// partials of each value in a fixed order (thread t takes entries t, t + 4 bd, ... into four interleaved sums;
// a separate launch reads them in 16-byte pairs),
// applies op, and resets the ticket for the next reduction. coherent: read with agent-scope atomic loads (the
// partials were written by the same kernel's other blocks).
__device__ void fin_apply(SolveState *st, double *partials, int nblk, int nv, int op, bool coherent)
{
    __shared__ double sh[2][16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6, bd = blockDim.x;
    auto ld = [&](const double *p) {
        return coherent ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
    };
    for (int a = 0; a < nv; a++)
    {
        const double *pa = partials + (size_t)a * nblk;
        double x;
        if (!coherent && ((uintptr_t)pa & 15u) == 0)
        {
            // a separate launch over many partials: 16-byte loads (pairs t, t + bd, ... into four pair sums), twice
            // the bytes in flight of the scalar loop; a lone last partial is added by thread 0
            const double2 *p2 = reinterpret_cast<const double2 *>(pa);
            const int n2 = nblk >> 1;
            double2 y0 = make_double2(0.0, 0.0), y1 = y0, y2 = y0, y3 = y0;
            int q = tid;
            for (; q + 3 * bd < n2; q += 4 * bd)
            {
                const double2 a0 = p2[q], a1 = p2[q + bd], a2 = p2[q + 2 * bd], a3 = p2[q + 3 * bd];
                y0.x += a0.x; y0.y += a0.y;
                y1.x += a1.x; y1.y += a1.y;
                y2.x += a2.x; y2.y += a2.y;
                y3.x += a3.x; y3.y += a3.y;
            }
            for (; q < n2; q += bd)
            {
                const double2 a0 = p2[q];
                y0.x += a0.x; y0.y += a0.y;
            }
            x = ((y0.x + y0.y) + (y1.x + y1.y)) + ((y2.x + y2.y) + (y3.x + y3.y));
            if (tid == 0 && (nblk & 1)) x += pa[nblk - 1];
        }
        else
        {
            double x0 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
            int q = tid;
            for (; q + 3 * bd < nblk; q += 4 * bd)
            {
                x0 += ld(pa + q);
                x1 += ld(pa + q + bd);
                x2 += ld(pa + q + 2 * bd);
                x3 += ld(pa + q + 3 * bd);
            }
            for (; q < nblk; q += bd) x0 += ld(pa + q);
            x = (x0 + x1) + (x2 + x3);
        }
        x = wave_sum(x);
        if (lane == 0) sh[a][w] = x;
    }
    __syncthreads();
    if (tid != 0) return;
    __hip_atomic_store(&st->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (fin_in_loop(op) && st->done) return;
    if ((op == FIN_TRUE_RES || op == FIN_TRUE_RES_SUM) && !st->done) return;
    if (op == FIN_SUM2_ACC && !st->accepted) return;
    double t[2] = {0.0, 0.0};
    for (int a = 0; a < nv; a++)
        for (int q = 0; q < nw; q++) t[a] += sh[a][q];
    if (fin_in_loop(op) && (!isfinite(t[0]) || !isfinite(t[1])))
    {
        // a NaN or infinity in the recursion (an inconsistent singular system drives p^T A p to round-off and
        // alpha past the range): stop the loop; no comparison with a NaN would ever end it
        st->done = kDoneNonFinite;
        return;
    }
    switch (op)
    {
    case FIN_SUM0: st->sums[0] = t[0]; break;
    case FIN_SUM1: st->sums[1] = t[0]; break;
    case FIN_SUM2: st->sums[2] = t[0]; break;
    case FIN_TRUE_RES:
    case FIN_TRUE_RES_SUM:
        st->sums[1] = t[0];
        st->accepted = (st->skip || t[0] < st->acc2) ? 1 : 0;
        if (op == FIN_TRUE_RES_SUM) st->sums[2] = t[1];
        break;
    case FIN_SUM2_ACC: st->sums[2] = t[0]; break;
    case FIN_RZ:
        st->beta = (st->first || st->rz == 0.0) ? 0.0 : t[0] / st->rz;
        st->rz = t[0];
        break;
    case FIN_PQ:
        st->alpha = t[0] != 0.0 ? st->rz / t[0] : 0.0;
        st->alphas[st->iters % kXRing] = st->alpha;
        st->first = 0;
        break;
    case FIN_RR:
        st->rr = t[0];
        st->iters += 1;
        if (t[0] < st->thr2) st->done = 1;
        else if (st->iters >= st->maxIters) st->done = 2;
        break;
    case FIN_CG64_INIT:
        st->rz = t[0];
        st->iters = 0;
        break;
    case FIN_CG64_PQ: st->alpha = t[0] != 0.0 ? st->rz / t[0] : 0.0; break;
    case FIN_CG64_UPDATE:
        st->rr = t[0];
        if (t[0] < st->thr2)
            st->done = 1;   // Eigen breaks before counting this pass (ConjugateGradient.h:78-79)
        else
        {
            st->beta = st->rz != 0.0 ? t[1] / st->rz : 0.0;
            st->rz = t[1];
            st->iters += 1;
            if (st->iters >= st->maxIters) st->done = 2;
        }
        break;
    default: break;
    }
}

__global__ __launch_bounds__(1024) void k_finalize(SolveState *st, double *partials, int nblk, int nv, int op)
{
    fin_apply(st, partials, nblk, nv, op, false);
}

// The <r, z> of the whole-row sweep from its partials, one per interior row and plane, in two fixed stages: block q
// sums the rows [q seg, (q + 1) seg) (thread t the entries t, t + 256, ... of them), and the last block to finish
// sums the block sums in index order and applies FIN_RZ (block_sum's fused finalisation). The order depends on the
// row count only, never on the sweep's tiles.
__device__ __forceinline__ void rz_rows_sum_body(SolveState *st, const double *__restrict__ rows, int nrows, int seg,
                                                 double *blockSums, int blk, int nblk)
{
    const int e = min(nrows, (blk + 1) * seg);
    double v[1] = {0.0};
    for (int q = blk * seg + (int)threadIdx.x; q < e; q += blockDim.x) v[0] += rows[q];
    block_sum<1>(v, Red{blockSums, st, (int)FIN_RZ, 1}, nblk, blk);
}

__global__ __launch_bounds__(kBlockThreads) void k_rz_rows_sum(SolveState *st, const double *__restrict__ rows,
                                                               int nrows, int seg, double *blockSums)
{
    rz_rows_sum_body(st, rows, nrows, seg, blockSums, blockIdx.x, gridDim.x);
}

// One block for a separate finalisation: 256 threads, or 1024 for many partials.
inline int fin_threads(int nblk) { return nblk > 8192 ? 1024 : kBlockThreads; }

// ---- FP64 kernels --------------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlockThreads) void k_scatter64(int64_t n, const int32_t *__restrict__ order,
                                                             const double *__restrict__ src, double *__restrict__ box,
                                                             Red red, int withSum)
{
    double v2 = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlockThreads)
    {
        const double v = src ? src[q] : 0.0;
        box[order[q]] = v;
        v2 += v * v;
    }
    if (!withSum) return;
    double a[1] = {v2};
    block_sum<1>(a, red, gridDim.x, blockIdx.x);
}

// The right-hand side of a solve without a guess: b64 = b, r32 = (float) b, and (where given) r64 = b and x = 0 at
// every unknown (the residual of the zero guess, b - A 0, is b bit for bit; a pass that starts from the zero guess
// reads b for it and 0.0 for x instead, SolveState::zeroStart), and <b, b> reduced as k_scatter64 reduces it.
__global__ __launch_bounds__(kBlockThreads) void k_scatter_rhs64(int64_t n, const int32_t *__restrict__ order,
                                                                 const double *__restrict__ src,
                                                                 double *__restrict__ b64, double *__restrict__ r64,
                                                                 float *__restrict__ r32, double *__restrict__ x64,
                                                                 Red red)
{
    double v2 = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlockThreads)
    {
        const double v = src[q];
        const int32_t o = order[q];
        b64[o] = v;
        if (r64) r64[o] = v;
        r32[o] = (float)v;
        if (x64) x64[o] = 0.0;
        v2 += v * v;
    }
    double a[1] = {v2};
    block_sum<1>(a, red, gridDim.x, blockIdx.x);
}

__global__ __launch_bounds__(kBlockThreads) void k_gather64(int64_t n, const int32_t *__restrict__ order,
                                                            const double *__restrict__ box, double *__restrict__ dst,
                                                            const SolveState *st, int meanSlot, int gate)
{
    if (gate && !st->accepted) return;
    const double mean = meanSlot >= 0 ? st->sums[meanSlot] / (double)n : 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlockThreads)
        dst[q] = box[order[q]] - mean;
}

__global__ __launch_bounds__(kBlockThreads) void k_sum64(int64_t n, const int32_t *__restrict__ order,
                                                         const double *__restrict__ box, Red red, int gate)
{
    if (gate && !red.st->accepted) return;   // every block alike: no partial, no ticket
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlockThreads)
        acc += box[order[q]];
    double a[1] = {acc};
    block_sum<1>(a, red, gridDim.x, blockIdx.x);
}

// One refinement step: xn = x + e (e may be null), written to xout, and r = b - A xn in FP64 with the reference's
// exact coefficients (neighbour values xn formed on the fly, so x and xout must differ), stored as FP32 in r32
// and/or FP64 in r64; partial sums of r^2.
__global__ __launch_bounds__(kBlockThreads) void k_update_residual64(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                                     const double *__restrict__ x,
                                                                     const float *__restrict__ e,
                                                                     double *__restrict__ xout,
                                                                     const double *__restrict__ b,
                                                                     float *__restrict__ r32, double *__restrict__ r64,
                                                                     Red red, int gate, int sumX)
{
#pragma clang fp contract(off)
    if (gate && !red.st->done) return;   // every block alike: no partial, no ticket
    const TileRange tr = my_tiles(B.blocks());
    double rr = 0.0, xs = 0.0;   // <r, r>; with sumX the sum of x over the unknowns
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        double r = 0.0;
        if (ty[idx] == kFluid)
        {
            double ax;
            if (e)
            {
                ax = apply64(B, T, ty, s, m, f, idx, [&](int64_t d) { return x[idx + d] + (double)e[idx + d]; });
                xout[idx] = x[idx] + (double)e[idx];
            }
            else
            {
                ax = apply64(B, T, ty, s, m, f, idx, [&](int64_t d) { return x[idx + d]; });
            }
            r = b[idx] - ax;
            xs += e ? x[idx] + (double)e[idx] : x[idx];
        }
        if (r32) r32[idx] = (float)r;
        if (r64) r64[idx] = r;
        rr += r * r;
    }
    if (sumX)
    {
        double a[2] = {rr, xs};
        block_sum<2>(a, red, gridDim.x, tr.lb);
    }
    else
    {
        double a[1] = {rr};
        block_sum<1>(a, red, gridDim.x, tr.lb);
    }
}

// The deferred update's rest at the end of a pass (launch_flush_x): x += alpha_j p_j for the ring slots j < iters mod
// kXRing, at the unknowns, oldest first. Four cells per thread (elements 4q - 1 .. 4q + 2: element -1 of every
// box array is 16-byte aligned, and the allocations extend past the box), a float4 of each direction and two
// double2 of x. A pass that began at x = 0 (SolveState::zeroStart) has not written x while no full ring has been
// applied: the sums then start from 0.0 instead of reading it -- the same additions, the same bits -- and a pass
// that took no iteration writes that 0 at the unknowns.
__global__ __launch_bounds__(kBlockThreads) void k_flush_x(int64_t quads, const uint8_t *__restrict__ c8,
                                                           const float *__restrict__ ring, int64_t stride,
                                                           double *__restrict__ x, const SolveState *st)
{
#pragma clang fp contract(off)
    if (!st->done) return;
    const int iters = st->iters, rem = iters % kXRing;
    const bool zero = st->zeroStart && iters < kXRing;
    if (rem == 0 && !zero) return;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += step)
    {
        const int64_t i = 4 * q - 1;
        const uchar4 c = *reinterpret_cast<const uchar4 *>(c8 + i);
        const bool u0 = c.x & 64u, u1 = c.y & 64u, u2 = c.z & 64u, u3 = c.w & 64u;
        if (!(u0 || u1 || u2 || u3)) continue;
        double2 xa = make_double2(0.0, 0.0), xb = make_double2(0.0, 0.0);
        if (!zero)
        {
            xa = *reinterpret_cast<const double2 *>(x + i);
            xb = *reinterpret_cast<const double2 *>(x + i + 2);
        }
        double v0 = xa.x, v1 = xa.y, v2 = xb.x, v3 = xb.y;
        for (int j = 0; j < rem; j++)
        {
            const double a = st->alphas[j];
            const float4 p = *reinterpret_cast<const float4 *>(ring + (size_t)j * (size_t)stride + i);
            v0 = v0 + a * (double)p.x;
            v1 = v1 + a * (double)p.y;
            v2 = v2 + a * (double)p.z;
            v3 = v3 + a * (double)p.w;
        }
        *reinterpret_cast<double2 *>(x + i) = make_double2(u0 ? v0 : xa.x, u1 ? v1 : xa.y);
        *reinterpret_cast<double2 *>(x + i + 2) = make_double2(u2 ? v2 : xb.x, u3 ? v3 : xb.y);
    }
}

// The FP64 Krylov recursion's two passes on any box (2D grids), one thread per cell over the tiles, with the
// reference's FP64 operator (apply64). Part 1: p = (float)(z + beta p_old) at the unknowns and <p, A p>; p is
// pointwise, so the neighbours' directions are formed from their z and p_old as well and no pass reads a p it
// writes. Part 2: q = A p, x += alpha p, r -= alpha q, r32 = (float) r and <r, r>.
__global__ __launch_bounds__(kBlockThreads) void k_cg_dir64_g(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                              const float *__restrict__ z, float *__restrict__ ring,
                                                              int64_t stride, const SolveState *st, Red red)
{
#pragma clang fp contract(off)
    if (st->done) return;
    const int first = st->first;
    const double beta = st->beta;
    const int slot = st->iters % kXRing;
    float *__restrict__ pn = ring + (size_t)slot * (size_t)stride;
    const float *__restrict__ po = ring + (size_t)((slot + kXRing - 1) % kXRing) * (size_t)stride;
    auto PV = [&](int64_t i) { return (double)(float)(first ? (double)z[i] : (double)z[i] + beta * (double)po[i]); };
    const TileRange tr = my_tiles(B.blocks());
    double pq = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (!(T.c8[idx] & 64u)) continue;
        const double pc = PV(idx);
        const double v = apply64(B, T, ty, s, m, f, idx, [&](int64_t d) { return PV(idx + d); });
        pn[idx] = (float)pc;
        pq += pc * v;
    }
    double a[1] = {pq};
    block_sum<1>(a, red, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg_update64_g(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                                 const float *__restrict__ ring, int64_t stride,
                                                                 double *__restrict__ x, double *__restrict__ r,
                                                                 float *__restrict__ r32, const double *__restrict__ b0,
                                                                 const SolveState *st, Red red)
{
#pragma clang fp contract(off)
    if (st->done) return;
    const double alpha = st->alpha;
    const int slot = st->iters % kXRing;
    const bool flush = slot == kXRing - 1;
    // a pass begun at x = 0: the residual is still b before its first update, x still 0 before its first ring
    const double *__restrict__ rin = st->zeroStart && st->iters == 0 ? b0 : r;
    const bool x0 = st->zeroStart && st->iters < kXRing;
    const float *__restrict__ p = ring + (size_t)slot * (size_t)stride;
    const TileRange tr = my_tiles(B.blocks());
    double rr = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (!(T.c8[idx] & 64u)) continue;
        const double pc = p[idx];
        const double q = apply64(B, T, ty, s, m, f, idx, [&](int64_t d) { return (double)p[idx + d]; });
        const double rv = rin[idx] - alpha * q;
        r[idx] = rv;
        r32[idx] = (float)rv;
        rr += rv * rv;
        if (flush)
        {
            double xv = x0 ? 0.0 : x[idx];
#pragma unroll
            for (int j = 0; j < kXRing; j++)
                xv = xv + st->alphas[j] * (j == kXRing - 1 ? pc : (double)ring[(size_t)j * (size_t)stride + idx]);
            x[idx] = xv;
        }
    }
    double a[1] = {rr};
    block_sum<1>(a, red, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_apply64(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                           const double *__restrict__ x, double *__restrict__ y)
{
    const TileRange tr = my_tiles(B.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        double v = 0.0;
        if (ty[idx] == kFluid) v = apply64(B, T, ty, s, m, f, idx, [&](int64_t d) { return x[idx + d]; });
        y[idx] = v;
    }
}

// Jacobi: Eigen's DiagonalPreconditioner takes 1/value of the diagonal entry, or 1 where the row has none.
__device__ __forceinline__ double inv_diag64(const Op64 &o) { return o.diag != 0.0 ? 1.0 / o.dval : 1.0; }

__device__ __forceinline__ double inv_diag_at(const Box &B, const Tables64 &T, const uint8_t *ty, int s, int m, int f,
                                              int64_t idx)
{
#pragma clang fp contract(off)
    if (T.unitCoef)
    {
        const unsigned code = T.c8[idx];
        if (!(code & 128u))
        {
            const unsigned n = __popc(code & 63u);
            return n ? 1.0 / (T.scale * (-(double)n)) : 1.0;
        }
    }
    return inv_diag64(op64(B, T, ty, s, m, f, idx));
}

__global__ __launch_bounds__(kBlockThreads) void k_cg64_init(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                             const double *__restrict__ r, double *__restrict__ p,
                                                             Red red)
{
    const TileRange tr = my_tiles(B.blocks());
    double rz = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        double pv = 0.0;
        if (ty[idx] == kFluid)
        {
            pv = inv_diag_at(B, T, ty, s, m, f, idx) * r[idx];
            rz += r[idx] * pv;
        }
        p[idx] = pv;
    }
    double a[1] = {rz};
    block_sum<1>(a, red, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg64_spmv(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                             const double *__restrict__ p, double *__restrict__ q,
                                                             const SolveState *st, Red red)
{
    if (st->done) return;
    const TileRange tr = my_tiles(B.blocks());
    double pq = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        double v = 0.0;
        if (ty[idx] == kFluid)
        {
            v = apply64(B, T, ty, s, m, f, idx, [&](int64_t d) { return p[idx + d]; });
            pq += p[idx] * v;
        }
        q[idx] = v;
    }
    double a[1] = {pq};
    block_sum<1>(a, red, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg64_update(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                               double *__restrict__ x, double *__restrict__ r,
                                                               const double *__restrict__ p,
                                                               const double *__restrict__ q, const SolveState *st,
                                                               Red red)
{
    if (st->done) return;
    const double alpha = st->alpha;
    const TileRange tr = my_tiles(B.blocks());
    double rr = 0.0, rz = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (ty[idx] != kFluid) continue;
        x[idx] += alpha * p[idx];
        const double rv = r[idx] - alpha * q[idx];
        r[idx] = rv;
        rr += rv * rv;
        rz += rv * (inv_diag_at(B, T, ty, s, m, f, idx) * rv);
    }
    double a[2] = {rr, rz};
    block_sum<2>(a, red, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg64_direction(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                                  const double *__restrict__ r, double *__restrict__ p,
                                                                  const SolveState *st)
{
    if (st->done) return;
    const double beta = st->beta;
    const TileRange tr = my_tiles(B.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (ty[idx] != kFluid) continue;
        p[idx] = inv_diag_at(B, T, ty, s, m, f, idx) * r[idx] + beta * p[idx];
    }
}

// ---- FP32 multigrid kernels ----------------------------------------------------------------------------------

// One full red-black Gauss-Seidel sweep in one pass over memory (x_in -> x_out, never in place).
//
// A block owns a TF x TM tile of the mid/fast plane and marches along the slow axis over a chunk of planes. The
// old x of four consecutive planes sits in LDS with a two-cell halo. At plane p the first colour is updated on the
// tile plus a one-cell halo (the halo values are recomputed from the same old data the neighbouring block uses,
// so every block sees identical numbers), then the second colour on the tile of plane p-1, whose six neighbours
// are all first-colour values already final, and plane p-1 leaves for x_out. Colour of cell (s, m, f) is
// (s + m + f) & 1; c1 is the colour updated first (0 = red first for the pre-smoother, 1 for the adjoint
// post-smoother). zero: x_in is taken as 0 (the first sweep of a V-cycle level).
constexpr int kTF = 64;
constexpr int kLW = kTF + 4;   // LDS row length: the tile plus a two-cell halo each side

template <int TM>
struct Rbgs
{
    static constexpr int kLoad = ((TM + 4) * kLW + kBlockThreads - 1) / kBlockThreads;   // x loads per thread
    static constexpr int kC1 = ((TM + 2) * 33 + kBlockThreads - 1) / kBlockThreads;      // first-colour cells
    static constexpr int kC2 = (TM * 32 + kBlockThreads - 1) / kBlockThreads;            // second-colour cells
    static constexpr int kStore = (TM * kTF + kBlockThreads - 1) / kBlockThreads;
};

// Per-thread staging of one plane's global data, loaded a step before it is used.
template <int TM>
struct RbgsRegs
{
    float x[Rbgs<TM>::kLoad];
    float b1[Rbgs<TM>::kC1];
    unsigned c1[Rbgs<TM>::kC1];
    float b2[Rbgs<TM>::kC2];
    unsigned c2[Rbgs<TM>::kC2];
};

struct RbgsCtx
{
    Level32 L;
    int m0, f0;
};

// Cell k of colour col among rows [r0, r0 + rows) and columns [cA, cA + 2*perRow) (local halo-2 coordinates) of
// plane pl: returns false when the slot is empty or outside the interior.
__device__ __forceinline__ bool rbgs_cell(const RbgsCtx &C, int pl, int col, int j, int r0, int rows, int cA,
                                          int perRow, int &lr, int &lc, int64_t &idx)
{
    if (j >= rows * perRow) return false;
    lr = r0 + j / perRow;
    const int k = j - (j / perRow) * perRow;
    const int m = C.m0 - 2 + lr;
    const int fA = C.f0 - 2 + cA;
    const int par = (col - pl - m - fA) & 1;
    lc = cA + 2 * k + par;
    const int f = C.f0 - 2 + lc;
    const Box &B = C.L.B;
    if (lc > cA + 2 * perRow - 1 || m < 1 || m > B.nm - 2 || f < 1 || f > B.nf - 2) return false;
    idx = B.at(pl, m, f);
    return true;
}

// The coarse correction a sweep folds into x where it reads it (x + P x_c, the injection prolongation): the coarse
// level's x and its strides (the parent of fine cell (s, m, f) is ((s + 1) / 2, (m + 1) / 2, (f + 1) / 2)).
struct Prolong
{
    const float *xc;
    int64_t sS;
    int sM, ns, nf;
};

// x of plane pl on the tile and its two-cell halo. With pro (2D levels), x + P x_c at the unknowns, as k_prolong
// forms it (the parent of a cell that is no unknown, or of one whose parent is none, adds 0: such x_c are 0).
template <int TM>
__device__ __forceinline__ void rbgs_fetch_x(const RbgsCtx &C, const float *xin, int zero, int pl,
                                             float (&v)[Rbgs<TM>::kLoad], const Prolong *pro = nullptr)
{
    const Box &B = C.L.B;
    const bool planeOk = !zero && pl >= 0 && pl < B.ns;
#pragma unroll
    for (int k = 0; k < Rbgs<TM>::kLoad; k++)
    {
        const int i = threadIdx.x + k * kBlockThreads;
        const int lr = i / kLW, lc = i - lr * kLW;
        const int m = C.m0 - 2 + lr, f = C.f0 - 2 + lc;
        const bool in = planeOk && lr < TM + 4 && m >= 0 && m < B.nm && f >= 0 && f < B.nf;
        float x = in ? xin[B.at(pl, m, f)] : 0.f;
        if (pro && in)
        {
            const unsigned c = C.L.c8[B.at(pl, m, f)];
            const float xcv = pro->xc[(int64_t)((m + 1) >> 1) * pro->sM + ((f + 1) >> 1)];
            x += (c & 64u) ? xcv : 0.f;
        }
        v[k] = x;
    }
}

template <int TM>
__device__ __forceinline__ void rbgs_put_x(float (*X)[kLW], const float (&v)[Rbgs<TM>::kLoad])
{
#pragma unroll
    for (int k = 0; k < Rbgs<TM>::kLoad; k++)
    {
        const int i = threadIdx.x + k * kBlockThreads;
        const int lr = i / kLW, lc = i - lr * kLW;
        if (lr < TM + 4) X[lr][lc] = v[k];
    }
}

// code and b of the first-colour cells of plane pl (tile + 1-cell halo) and second-colour cells of plane pl2.
template <int TM>
__device__ __forceinline__ void rbgs_fetch_cells(const RbgsCtx &C, const float *__restrict__ b, int pl, int c1,
                                                 int pl2, int c2, RbgsRegs<TM> &R)
{
    const Box &B = C.L.B;
    const bool ok1 = B.threeD ? (pl >= 1 && pl <= B.ns - 2) : pl == 0;
    const bool ok2 = B.threeD ? (pl2 >= 1 && pl2 <= B.ns - 2) : pl2 == 0;
#pragma unroll
    for (int k = 0; k < Rbgs<TM>::kC1; k++)
    {
        int lr, lc;
        int64_t idx;
        R.c1[k] = 0u;
        R.b1[k] = 0.f;
        if (ok1 && rbgs_cell(C, pl, c1, threadIdx.x + k * kBlockThreads, 1, TM + 2, 1, 33, lr, lc, idx))
        {
            R.c1[k] = C.L.c8[idx];
            R.b1[k] = b[idx];
        }
    }
#pragma unroll
    for (int k = 0; k < Rbgs<TM>::kC2; k++)
    {
        int lr, lc;
        int64_t idx;
        R.c2[k] = 0u;
        R.b2[k] = 0.f;
        if (ok2 && rbgs_cell(C, pl2, c2, threadIdx.x + k * kBlockThreads, 2, TM, 2, 32, lr, lc, idx))
        {
            R.c2[k] = C.L.c8[idx];
            R.b2[k] = b[idx];
        }
    }
}

// Gauss-Seidel update of one cell from LDS neighbours: Xc the cell's plane, Xp / Xm the planes above / below.
__device__ __forceinline__ float rbgs_update(const RbgsCtx &C, int pl, int lr, int lc, int64_t idx, unsigned code,
                                             float bv, float (*Xc)[kLW], float (*Xp)[kLW], float (*Xm)[kLW])
{
    if (fast_cell(C.L, code))
    {
        float acc = 0.f;
        if (C.L.B.threeD)
        {
            acc += (code & 1u) ? Xp[lr][lc] : 0.f;
            acc += (code & 2u) ? Xm[lr][lc] : 0.f;
        }
        acc += (code & 4u) ? Xc[lr + 1][lc] : 0.f;
        acc += (code & 8u) ? Xc[lr - 1][lc] : 0.f;
        acc += (code & 16u) ? Xc[lr][lc + 1] : 0.f;
        acc += (code & 32u) ? Xc[lr][lc - 1] : 0.f;
        return (acc - bv * C.L.icu) * inv_count(code);
    }
    const int m = C.m0 - 2 + lr, f = C.f0 - 2 + lc;
    const Op32 o = op32(C.L, pl, m, f, idx, code);
    float acc = 0.f;
    if (C.L.B.threeD)
    {
        acc += (o.coupled & 1u) ? o.c[0] * Xp[lr][lc] : 0.f;
        acc += (o.coupled & 2u) ? o.c[1] * Xm[lr][lc] : 0.f;
    }
    acc += (o.coupled & 4u) ? o.c[2] * Xc[lr + 1][lc] : 0.f;
    acc += (o.coupled & 8u) ? o.c[3] * Xc[lr - 1][lc] : 0.f;
    acc += (o.coupled & 16u) ? o.c[4] * Xc[lr][lc + 1] : 0.f;
    acc += (o.coupled & 32u) ? o.c[5] * Xc[lr][lc - 1] : 0.f;
    return o.diag != 0.f ? (bv - acc) / o.diag : 0.f;
}

template <int TM>
__device__ __forceinline__ void rbgs_colour1(const RbgsCtx &C, float (*X)[kLW], int c1, const RbgsRegs<TM> &R)
{
#pragma unroll
    for (int k = 0; k < Rbgs<TM>::kC1; k++)
    {
        if (!(R.c1[k] & 64u)) continue;
        int lr, lc;
        int64_t idx;
        rbgs_cell(C, 0, c1, threadIdx.x + k * kBlockThreads, 1, TM + 2, 1, 33, lr, lc, idx);
        X[lr][lc] = rbgs_update(C, 0, lr, lc, idx, R.c1[k], R.b1[k], X, X, X);
    }
}

template <int TM>
__device__ __forceinline__ void rbgs_colour2(const RbgsCtx &C, float (*X)[kLW], int c2, const RbgsRegs<TM> &R)
{
#pragma unroll
    for (int k = 0; k < Rbgs<TM>::kC2; k++)
    {
        if (!(R.c2[k] & 64u)) continue;
        int lr, lc;
        int64_t idx;
        rbgs_cell(C, 0, c2, threadIdx.x + k * kBlockThreads, 2, TM, 2, 32, lr, lc, idx);
        X[lr][lc] = rbgs_update(C, 0, lr, lc, idx, R.c2[k], R.b2[k], X, X, X);
    }
}

template <int TM>
__device__ __forceinline__ void rbgs_store(const RbgsCtx &C, float (*X)[kLW], float *xout)
{
    const Box &B = C.L.B;
#pragma unroll
    for (int k = 0; k < Rbgs<TM>::kStore; k++)
    {
        const int i = threadIdx.x + k * kBlockThreads;
        const int r = i / kTF, c = i - r * kTF;
        const int m = C.m0 + r, f = C.f0 + c;
        if (r < TM && m <= B.nm - 2 && f <= B.nf - 2) xout[B.at(0, m, f)] = X[r + 2][c + 2];
    }
}

// One full red-black Gauss-Seidel sweep of a 2D level in one pass over memory (x_in -> x_out, never in place).
//
// A block owns a 64 x TM tile of the plane. Its old x, with a two-cell halo, goes to LDS; the first colour is
// updated on the tile plus a one-cell halo (the halo values are recomputed from the same old data the neighbouring
// block uses, so every block sees identical numbers), then the second colour on the tile, whose four neighbours are
// all first-colour values already final, and the tile leaves for x_out. Colour of cell (m, f) is (m + f) & 1; c1 is
// updated first (0: red first, the pre-smoother; 1: its adjoint, the post-smoother). zero: x_in is taken as 0. PRO:
// x + P x_c wherever x is read (the coarse correction folded in). 3D levels sweep in the marches below.
template <int TM, bool PRO>
__global__ __launch_bounds__(kBlockThreads) void k_rbgs(Level32 L, const float *__restrict__ xin,
                                                        float *__restrict__ xout, const float *__restrict__ b, int c1,
                                                        int zero, int tilesF, const SolveState *st, Prolong pro)
{
    if (st->done) return;
    __shared__ float X[TM + 4][kLW];
    const int lb = logical_block();
    const int tf = lb % tilesF, tm = lb / tilesF;
    const RbgsCtx C{L, 1 + tm * TM, 1 + tf * kTF};
    const int c2 = c1 ^ 1;
    RbgsRegs<TM> R;
    float v[Rbgs<TM>::kLoad];
    rbgs_fetch_x<TM>(C, xin, zero, 0, v, PRO ? &pro : nullptr);
    rbgs_fetch_cells<TM>(C, b, 0, c1, 0, c2, R);
    rbgs_put_x<TM>(X, v);
    __syncthreads();
    rbgs_colour1<TM>(C, X, c1, R);
    __syncthreads();
    rbgs_colour2<TM>(C, X, c2, R);
    __syncthreads();
    rbgs_store<TM>(C, X, xout);
}

// ---- red-black Gauss-Seidel sweep, 2.5D march (3D boxes) -----------------------------------------------------
//
// Each lane owns a pair of neighbouring cells (f, f+1) of one row and marches with it along the slow axis, so on
// every plane it holds exactly one cell of each colour and no lane idles. Colour of (s, m, f) is (s + m + f) & 1 and
// c1 is the colour updated first. At plane step s the lane updates its first-colour cell of plane s from old values
// (x_in: slow neighbours from its registers, in-row neighbours from the neighbouring lanes by shuffle, the rows
// above and below from x_in through L1), publishes it in LDS, then updates its second-colour cell of plane s-1,
// whose six neighbours are first-colour values already final: slow ones from its own registers, in-row ones by
// shuffle, mid ones from LDS. Two extra waves march the rows just outside the tile and only publish first-colour
// values; lanes 0 and 63 also update the first-colour cells just left / right of the tile when the second-colour
// cell at the tile edge needs them. Every redundant value is computed from the same old data its owner uses, so
// blocks agree bit for bit and the sweep equals the two-pass red-black sweep. One barrier per plane.
constexpr int kPW = 128;   // tile width: 64 lanes x 2 cells

// Gauss-Seidel value of one cell from its b and neighbours (slow+, slow-, mid+, mid-, fast+, fast-). UNI: the
// level is uniform, so even a cell beside a Dirichlet cell needs no tables (its code16 gives the ring factors).
template <bool UNI>
__device__ __forceinline__ float gs_update(const Level32 &L, int s, int m, int f, int64_t idx, unsigned code, float bv,
                                           const float (&nb)[6])
{
    if (UNI)
    {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < 6; q++) acc += ((code >> q) & 1u) ? nb[q] : 0.f;
        if (!(code & 128u)) return (acc - bv * L.icu) * inv_count(code);
        const unsigned c16 = L.c16[idx];
        float d = 0.f;   // -diag / cu
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned st = (c16 >> (2 * q)) & 3u;
            d += st == 3u ? 2.f : (st != 0u ? 1.f : 0.f);
        }
        return d != 0.f ? (acc - bv * L.icu) / d : 0.f;
    }
    const Op32 o = op32(L, s, m, f, idx, code);
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 6; q++) acc += ((o.coupled >> q) & 1u) ? o.c[q] * nb[q] : 0.f;
    return o.diag != 0.f ? (bv - acc) / o.diag : 0.f;
}

template <int TM, bool UNI>
__global__ __launch_bounds__(64 * (TM + 2)) void k_rbgs_march(Level32 L, const float *__restrict__ xin,
                                                              float *__restrict__ xout, const float *__restrict__ b,
                                                              int c1, int zero, int P, int tilesF, int tilesM,
                                                              const SolveState *st)
{
    if (st->done) return;
    const Box &B = L.B;
    __shared__ float red[2][TM + 2][kPW + 2];   // first-colour values of rows m0-1 .. m0+TM, columns f0-1 .. f0+128
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;   // wave 0: row m0-1, wave TM+1: row m0+TM
    const int lb = logical_block();
    const int tf = lb % tilesF, t2 = lb / tilesF, tm = t2 % tilesM, chunk = t2 / tilesM;
    const int f0 = 1 + tf * kPW, m0 = 1 + tm * TM;
    const int m = m0 - 1 + wave;
    const bool haloRow = wave == 0 || wave == TM + 1;
    const bool rowOk = m >= 1 && m <= B.nm - 2;
    const int fa = f0 + 2 * lane;                     // the pair (fa, fa + 1)
    const bool okA = rowOk && fa <= B.nf - 2, okB = rowOk && fa + 1 <= B.nf - 2;
    const int s0 = 1 + chunk * P, s1 = min(s0 + P - 1, B.ns - 2);
    const bool leftLane = lane == 0 && !haloRow, rightLane = lane == 63 && !haloRow;
    const int fe = leftLane ? f0 - 1 : f0 + kPW;      // edge column of an edge lane
    const bool okE = (leftLane || rightLane) && rowOk && fe >= 1 && fe <= B.nf - 2;
    const int64_t rowBase = B.at(0, rowOk ? m : 1, 0);
    const bool rd = !zero;

    auto X = [&](int sp, int f) -> float {   // old value, 0 outside the box or for the zero start
        if (!rd || sp < 0 || sp >= B.ns) return 0.f;
        return xin[rowBase + (int64_t)sp * B.sS + f];
    };
    auto inPlane = [&](int sp) { return sp >= 1 && sp <= B.ns - 2; };
    // Old values of the pair and the edge column at planes sp-1, sp, sp+1 and (prefetched) sp+2.
    float xa[4], xb[4], xe[4];
    {
        const int sp = s0 - 1;
        xa[1] = okA ? X(sp - 1, fa) : 0.f;
        xb[1] = okB ? X(sp - 1, fa + 1) : 0.f;
        xe[1] = okE ? X(sp - 1, fe) : 0.f;
        xa[2] = okA ? X(sp, fa) : 0.f;
        xb[2] = okB ? X(sp, fa + 1) : 0.f;
        xe[2] = okE ? X(sp, fe) : 0.f;
        xa[3] = okA ? X(sp + 1, fa) : 0.f;
        xb[3] = okB ? X(sp + 1, fa + 1) : 0.f;
        xe[3] = okE ? X(sp + 1, fe) : 0.f;
    }
    // b and codes of the pair at the current plane (prefetched a step ahead) and at the previous one.
    float bA = 0.f, bB = 0.f, bPrevA = 0.f, bPrevB = 0.f;
    unsigned cA = 0u, cB = 0u, cPrevA = 0u, cPrevB = 0u;
    if (inPlane(s0 - 1))
    {
        const int64_t i = rowBase + (int64_t)(s0 - 1) * B.sS + fa;
        if (okA) { cA = L.c8[i]; bA = b[i]; }
        if (okB) { cB = L.c8[i + 1]; bB = b[i + 1]; }
    }
    // New first-colour values of the lane's column at planes sp-2, sp-1, sp.
    float R[3] = {0.f, 0.f, 0.f};
    for (int sp = s0 - 1; sp <= s1 + 1; sp++)
    {
#pragma unroll
        for (int k = 0; k < 3; k++)
        {
            xa[k] = xa[k + 1];
            xb[k] = xb[k + 1];
            xe[k] = xe[k + 1];
        }
        // prefetch for the next step: old x of plane sp+2, b and codes of plane sp+1
        xa[3] = okA ? X(sp + 2, fa) : 0.f;
        xb[3] = okB ? X(sp + 2, fa + 1) : 0.f;
        xe[3] = okE ? X(sp + 2, fe) : 0.f;
        float nbA = 0.f, nbB = 0.f;
        unsigned ncA = 0u, ncB = 0u;
        if (inPlane(sp + 1))
        {
            const int64_t i = rowBase + (int64_t)(sp + 1) * B.sS + fa;
            if (okA) { ncA = L.c8[i]; nbA = b[i]; }
            if (okB) { ncB = L.c8[i + 1]; nbB = b[i + 1]; }
        }
        R[0] = R[1];
        R[1] = R[2];
        const bool planeIn = inPlane(sp);
        // first colour at plane sp: element e1 of the pair; in-row neighbours are old values
        const int e1 = (c1 - sp - m - fa) & 1;
        float leftNb = __shfl_up(xb[1], 1, 64);
        float rightNb = __shfl_down(xa[1], 1, 64);
        if (lane == 0) leftNb = (rowOk && fa - 1 >= 0) ? (okE && leftLane ? xe[1] : X(sp, fa - 1)) : 0.f;
        if (lane == 63) rightNb = (rowOk && fa + 2 < B.nf) ? (okE && rightLane ? xe[1] : X(sp, fa + 2)) : 0.f;
        const int64_t idxA = rowBase + (int64_t)sp * B.sS + fa;
        float rnew = 0.f;
        {
            const unsigned code = e1 ? cB : cA;
            if (planeIn && (code & 64u))
            {
                const int64_t idx = idxA + e1;
                const float nb[6] = {e1 ? xb[2] : xa[2], e1 ? xb[0] : xa[0], rd ? xin[idx + B.sM] : 0.f,
                                     rd ? xin[idx - B.sM] : 0.f, e1 ? rightNb : xb[1], e1 ? xa[1] : leftNb};
                rnew = gs_update<UNI>(L, sp, m, fa + e1, idx, code, e1 ? bB : bA, nb);
            }
        }
        // the edge first-colour cell next to the tile-edge second-colour cell, same plane
        float redge = 0.f;
        if (okE && planeIn && ((c1 - sp - m - fe) & 1) == 0)
        {
            const int64_t idx = rowBase + (int64_t)sp * B.sS + fe;
            const unsigned code = L.c8[idx];
            if (code & 64u)
            {
                const float nb[6] = {xe[2], xe[0], rd ? xin[idx + B.sM] : 0.f, rd ? xin[idx - B.sM] : 0.f,
                                     leftLane ? xa[1] : X(sp, fe + 1), leftLane ? X(sp, fe - 1) : xb[1]};
                redge = gs_update<UNI>(L, sp, m, fe, idx, code, b[idx], nb);
            }
        }
        R[2] = rnew;
        // Publish plane sp's first-colour values (column 1 + 2 lane + e1; edges at 0 and kPW + 1). One barrier per
        // plane: it orders this step's writes of red[sp & 1] after the previous step's reads of the same buffer
        // (plane sp - 2), and the previous step's writes of red[q & 1] before this step's reads.
        __syncthreads();
        red[sp & 1][wave][1 + 2 * lane + e1] = rnew;
        if (leftLane) red[sp & 1][wave][0] = redge;
        if (rightLane) red[sp & 1][wave][kPW + 1] = redge;
        // second colour at plane q = sp - 1 on tile rows: the colour flips with the plane, so at q the first-colour
        // element is e1 ^ 1 and the second-colour element is e1
        const int q = sp - 1;
        if (!haloRow && q >= s0 && q <= s1)
        {
            const int eq = e1;
            const unsigned code = eq ? cPrevB : cPrevA;
            float out = eq ? xb[0] : xa[0];   // a non-unknown keeps its (zero) value
            const float rq = R[1];            // this pair's first-colour value at plane q (element eq ^ 1)
            if (code & 64u)
            {
                const int c = 1 + 2 * lane + eq;
                const float left = eq ? rq : red[q & 1][wave][c - 1];
                const float right = eq ? red[q & 1][wave][c + 1] : rq;
                const float nb[6] = {R[2], R[0], red[q & 1][wave + 1][c], red[q & 1][wave - 1][c], right, left};
                out = gs_update<UNI>(L, q, m, fa + eq, rowBase + (int64_t)q * B.sS + fa + eq, code,
                                     eq ? bPrevB : bPrevA, nb);
            }
            if (okA) xout[rowBase + (int64_t)q * B.sS + fa] = eq ? rq : out;
            if (okB) xout[rowBase + (int64_t)q * B.sS + fa + 1] = eq ? out : rq;
        }
        bPrevA = bA;
        bPrevB = bB;
        cPrevA = cA;
        cPrevB = cB;
        bA = nbA;
        bB = nbB;
        cA = ncA;
        cB = ncB;
    }
}

// gs_update<false> (op32's coefficients and diagonal, then the Gauss-Seidel quotient) from width-table entries the
// caller already holds: tS = (wS, gS[s], gS[s-1], hS) of the plane, tM of the row, tF of the column. The same
// products, sums and quotient in the same order, so the same value bit for bit.
__device__ __forceinline__ float gs_stretched(const uint16_t *c16, int64_t idx, unsigned code, float bv,
                                             const float (&nb)[6], float4 tS, float4 tM, float4 tF)
{
    const float wS = tS.x, wM = tM.x, wF = tF.x;
    const float aS = wM * wF, aM = wS * wF, aF = wS * wM;
    const float c[6] = {aS * tS.y, aS * tS.z, aM * tM.y, aM * tM.z, aF * tF.y, aF * tF.z};
    const unsigned coupled = code & 63u;
    float diag = 0.f;
    if (!(code & 128u))
    {
#pragma unroll
        for (int q = 0; q < 6; q++) diag -= ((coupled >> q) & 1u) ? c[q] : 0.f;
    }
    else
    {
        const float rS = aS * tS.w, rM = aM * tM.w, rF = aF * tF.w;
        const float ring[6] = {rS, rS, rM, rM, rF, rF};
        const unsigned cc = c16[idx];
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned st = (cc >> (2 * q)) & 3u;
            diag -= st == 3u ? ring[q] : (st != 0u ? c[q] : 0.f);
        }
    }
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 6; q++) acc += ((coupled >> q) & 1u) ? c[q] * nb[q] : 0.f;
    return diag != 0.f ? (bv - acc) / diag : 0.f;
}

// The coarse correction a whole-row sweep folds in (x_in + P x_c where x is used): the coarse level's iterate and
// box strides.
// ---- per-row code descriptors ---------------------------------------------------------------------------------
//
// A row whose interior cells 2 .. nf - 3 share one stencil code (most rows of a box: every row of an all-Neumann or
// all-outlet box) has its codes built from its descriptor (launch_row_codes) instead of read: a wave holds the
// descriptors of its row for the planes it marches over, lane j that of plane p0 + j (and p0 + 64 + j in the second
// word), and reads one out per plane with a wave-uniform lane index. A row without one (descriptor bit 31 clear, or
// no table) reads its codes as before.
struct RowCodes
{
    uint32_t a = 0u, b = 0u;
    int p0 = 0;
    // planes p0 .. pEnd of row mc of a box of ns planes (clamped into the box as the sweeps clamp their plane index)
    __device__ __forceinline__ void load(const uint32_t *rowc, int ns, int mc, int p0_, int pEnd)
    {
        p0 = p0_;
        if (!rowc) return;
        const uint32_t *rr = rowc + (size_t)mc * (size_t)ns;
        const int lane = threadIdx.x & 63;
        a = rr[min(max(p0 + lane, 0), ns - 1)];
        if (pEnd - p0 >= 64) b = rr[min(max(p0 + 64 + lane, 0), ns - 1)];
    }
    __device__ __forceinline__ void load(const Level32 &L, int mc, int p0_, int pEnd)
    {
        load(L.rowc, L.B.ns, mc, p0_, pEnd);
    }
    __device__ __forceinline__ uint32_t at(int pl) const
    {
        const int j = pl - p0;
        return j < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)a, j)
                      : (j < 128 ? (uint32_t)__builtin_amdgcn_readlane((int)b, j - 64) : 0u);
    }
};

// The pair (f, f + 1) codes (low byte: f) of a row with descriptor d: the first, middle or last code by column.
__device__ __forceinline__ unsigned row_pair_code(uint32_t d, bool first, bool loLast, bool hiLast)
{
    const unsigned cf = d & 255u, cmid = (d >> 8) & 255u, cl = (d >> 16) & 255u;
    const unsigned lo = first ? cf : (loLast ? cl : cmid);
    const unsigned hi = hiLast ? cl : cmid;
    return lo | (hi << 8);
}

// The codes of plane pl's pairs (first cells 1 + 2 (lane + 64 k), read at c8pl + off[k] for a row without a
// descriptor) into c: built where the row has a descriptor, else read.
template <int K>
__device__ __forceinline__ void row_codes_into(unsigned (&c)[K], const RowCodes &rc, int pl, const uint8_t *c8pl,
                                               const uint32_t (&off)[K], int lane, int nf)
{
    const uint32_t d = rc.at(pl);
    if (d >> 31)
    {
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const int fa = 1 + 2 * (lane + 64 * k);
            c[k] = row_pair_code(d, fa == 1, fa == nf - 2, fa + 1 == nf - 2);
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < K; k++) c[k] = (unsigned)*reinterpret_cast<const uint16_t *>(c8pl + off[k]);
    }
}

// ---- red-black Gauss-Seidel sweep, whole-row waves (3D uniform levels) -----------------------------------------
//
// As k_rbgs_march, but a wave spans the whole interior row: lane l owns the pairs (1 + 2p, 2 + 2p) for
// p = l, l + 64, ... (K pairs), so no tile edge exists along the fast axis and no edge values are recomputed. The
// block is TM rows plus one halo row above and below (halo waves publish first-colour values only). The old x and
// the first-colour values pass between waves through LDS double-buffered by plane parity, one barrier per plane
// step. Plane bases are scalar and per-lane offsets 32-bit; loads outside the interior are clamped into the box, so
// the loop has no bounds branches. Uniform levels only (constant coefficients).
//
// An update adds all six neighbours without testing the code's neighbour bits. A neighbour whose bit is clear is a
// ring cell or an interior cell that is no unknown, and every iterate holds +0 there (the buffers are cleared when a
// layout is set, every sweep writes 0 at non-unknowns, the prolongation adds at unknowns only), so the sum equals
// the sum over the coupled neighbours bit for bit. What the kernel makes up itself is zeroed where it stands for such
// a cell: the pairs past the end of a short row, the LDS columns beyond the row ends, and (prolongation folded in)
// the coarse correction at a fine cell that is no unknown. 1/n and the unknown mask come from a table in LDS
// indexed by code & 127. The wave index is made wave-uniform and a row's colour parity alternates with the plane,
// so the parity is a compile-time constant of each plane step: the kernel holds one plane loop per parity of the
// first plane, and a wave runs the one its row needs.
//
// Every global load of a plane is issued two planes (x, b, codes) or one plane (the halo waves' outer rows) ahead
// of its use: the old x of the mid-axis neighbour rows and of the neighbouring pairs comes from LDS (each wave
// publishes its own row), and the halo waves, whose outer neighbour row belongs to no wave of the block, prefetch
// that row. So no load is waited for in the plane that issues it.
// With dot non-null the sweep also writes <b, x_out> over the unknowns of each interior row and plane to
// dot[(plane - 1) * (nm - 2) + row - 1] (the CG's <r, z> when the sweep is the last one of the V-cycle on the
// finest level): per row, not per block, so the sum over them does not depend on the tile shape.
template <int TM, int K, bool ZERO, bool PRO, bool NU>
__device__ __forceinline__ void rbgs_row_body(const Level32 &L, const float *__restrict__ xin,
                                              float *__restrict__ xout, const float *__restrict__ b, int c1, int P,
                                              int tilesM, const SolveState *st, double *__restrict__ dot,
                                              const Prolong &pro, int lb)
{
    if (st->done) return;
    const Box &B = L.B;
    constexpr int kCols = 128 * K + 2;
    constexpr int kSlots = 64 * K + 2;
    __shared__ float red[2][TM + 2][kCols];      // first-colour values by column f, rows m0-1 .. m0+TM
    __shared__ float old[2][TM + 2][2][kSlots];  // old x of plane sp by parity, row, element (a = 0, b = 1), pair + 1
    __shared__ float2 tab[128];                  // by code & 127: (1/n, all ones) for an unknown, (0, 0) otherwise
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tm = lb % tilesM, chunk = lb / tilesM;
    const int m0 = 1 + tm * TM;
    const int m = m0 - 1 + wave;
    const bool haloRow = wave == 0 || wave == TM + 1;
    const bool rowIn = m >= 1 && m <= B.nm - 2;
    const bool ragged = B.nf - 2 < 128 * K;   // the row ends before the last lane's last pair
    const int mc = min(max(m, 0), B.nm - 1);
    // the row beyond a halo row (outside the block), clamped into the box
    const int mo = min(max(wave == 0 ? m - 1 : m + 1, 0), B.nm - 1);
    const int s0 = 1 + chunk * P, s1 = min(s0 + P - 1, B.ns - 2);
    const uint32_t sM = (uint32_t)B.sM;
    // per pair: first cell fa = 1 + 2p; load column fc = fa clamped to the last odd column whose pair stays in the
    // box (odd, so the pair load stays aligned)
    const int fcMax = ((B.nf - 2) & 1) ? B.nf - 2 : B.nf - 3;
    uint32_t off[K], offO[K];
    bool okA[K], okB[K];
    unsigned cm[K];   // the pair's codes where its elements are interior cells of the row (zero elsewhere)
    int fa[K];
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        fa[k] = 1 + 2 * (lane + 64 * k);
        okA[k] = rowIn && fa[k] <= B.nf - 2;
        okB[k] = rowIn && fa[k] + 1 <= B.nf - 2;
        cm[k] = (okA[k] ? 0xFFu : 0u) | (okB[k] ? 0xFF00u : 0u);
        const int fc = min(fa[k], fcMax);
        off[k] = (uint32_t)mc * sM + (uint32_t)fc;
        offO[k] = (uint32_t)mo * sM + (uint32_t)fc;
    }
    // PRO: the coarse parent of the pair (2p+1, 2p+2) is column p+1 of row (m+1)/2, plane (s+1)/2
    uint32_t offC[K], offCO[K];
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        const int Fc = min(lane + 64 * k + 1, pro.nf - 1);
        offC[k] = PRO ? (uint32_t)((mc + 1) >> 1) * (uint32_t)pro.sM + (uint32_t)Fc : 0u;
        offCO[k] = PRO ? (uint32_t)((mo + 1) >> 1) * (uint32_t)pro.sM + (uint32_t)Fc : 0u;
    }
    auto planeXC = [&](int pl) {
        const int sc = (min(max(pl, 0), B.ns - 1) + 1) >> 1;
        return pro.xc + (size_t)min(sc, pro.ns - 1) * (size_t)pro.sS;
    };
    // NU: the width-table entries of this lane's columns and of this wave's row, fixed over the march (indices
    // clamped into the interior: a cell outside it is no unknown and never uses them)
    float4 tF[K][2], tM = make_float4(0.f, 0.f, 0.f, 0.f);
    if (NU)
    {
        const Tables32 &T = L.T;
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int e = 0; e < 2; e++)
            {
                const int f = min(max(fa[k] + e, 1), B.nf - 2);
                tF[k][e] = make_float4(T.wF[f], T.gF[f], T.gF[f - 1], T.hF[f]);
            }
        const int mr = min(max(m, 1), B.nm - 2);
        tM = make_float4(T.wM[mr], T.gM[mr], T.gM[mr - 1], T.hM[mr]);
    }
    auto tablesS = [&](int pl) {
        const int sr = min(max(pl, 1), B.ns - 2);
        return make_float4(L.T.wS[sr], L.T.gS[sr], L.T.gS[sr - 1], L.T.hS[sr]);
    };
    for (int i = threadIdx.x; i < 2 * (TM + 2) * kCols; i += 64 * (TM + 2)) (&red[0][0][0])[i] = 0.f;
    for (int i = threadIdx.x; i < 2 * (TM + 2) * 2 * kSlots; i += 64 * (TM + 2)) (&old[0][0][0][0])[i] = 0.f;
    if (threadIdx.x < 128)
    {
        const unsigned c = threadIdx.x;
        tab[c] = make_float2((c & 64u) ? inv_count(c) : 0.f, __uint_as_float((c & 64u) ? ~0u : 0u));
    }
    auto plane = [&](const float *base, int pl) { return base + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto planeC = [&](int pl) { return L.c8 + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto plane_off = [&](int pl) { return (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    // one aligned load per pair (the arrays are allocated so that odd columns start pair-aligned)
    auto pair = [](const float *row, uint32_t o) { return *reinterpret_cast<const float2 *>(row + o); };
    auto pairC = [](const uint8_t *row, uint32_t o) { return (unsigned)*reinterpret_cast<const uint16_t *>(row + o); };
    // PRO: x + P x_c at the pair's unknowns (cc: the pair's masked codes)
    auto fold = [](float2 &x, float xc, unsigned cc) {
        x.x += (cc & 64u) ? xc : 0.f;
        x.y += (cc & (64u << 8)) ? xc : 0.f;
    };
    // (acc - b / cu) / n at an unknown, +0 elsewhere
    auto gs = [&](float acc, float bv, unsigned code) {
        const float2 t = tab[code & 127u];
        return __uint_as_float(__float_as_uint((acc - bv * L.icu) * t.x) & __float_as_uint(t.y));
    };
    // the update of an unknown with a Dirichlet neighbour: the diagonal from the two-bit neighbour states
    auto gsDirichlet = [&](float acc, float bv, int pl, uint32_t o) {
        const unsigned c16 = L.c16[(size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS + o];
        float d = 0.f;
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned stt = (c16 >> (2 * q)) & 3u;
            d += stt == 3u ? 2.f : (stt != 0u ? 1.f : 0.f);
        }
        return d != 0.f ? (acc - bv * L.icu) / d : 0.f;
    };
    // Per-plane values live in slots indexed by plane mod 4 (relative to the first plane s0-1), so the march
    // renames registers instead of moving them: a move of a register that a load in flight will write waits for
    // every load. x, b and codes are loaded two planes ahead, the halo waves' outer row one plane ahead.
    float2 X[4][K];    // old x of the pair (PRO: plus the coarse correction from the plane step before its use)
    float2 O[2][K];    // halo waves: old x of the outer row's pair
    float2 Bv[4][K];   // b of the pair
    unsigned C[4][K];  // codes of the pair (low byte: element a), masked with cm where used
    float R[4][K];     // first-colour value of this lane's pair
    float XC[4][K];    // PRO: the coarse correction of the pair
    float OC[2][K];    // PRO: and of the outer row's pair
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            X[j][k] = Bv[j][k] = make_float2(0.f, 0.f);
            O[j & 1][k] = make_float2(0.f, 0.f);
            R[j][k] = 0.f;
            C[j][k] = 0u;
            XC[j][k] = 0.f;
            OC[j & 1][k] = 0.f;
        }
    const int sp0 = s0 - 1;
    RowCodes rc;
    if (!NU) rc.load(L, mc, sp0 - 1, s1 + 3);
    // the codes of plane pl into slot c: built from the row's descriptor where it has one, else read (stretched
    // levels always read them: their width tables already fill the registers)
    auto loadCodes = [&](int c, int pl) {
        const uint32_t d = NU ? 0u : rc.at(pl);
        if (d >> 31)
        {
#pragma unroll
            for (int k = 0; k < K; k++) C[c][k] = row_pair_code(d, fa[k] == 1, fa[k] == B.nf - 2, fa[k] + 1 == B.nf - 2);
        }
        else
        {
            const uint8_t *pc = planeC(pl);
#pragma unroll
            for (int k = 0; k < K; k++) C[c][k] = pairC(pc, off[k]);
        }
    };
    // slots: plane sp0 + d -> d & 3 (d & 1 for O)
    if (!ZERO)
    {
#pragma unroll
        for (int d = -1; d <= 1; d++)
        {
            const float *px = plane(xin, sp0 + d);
#pragma unroll
            for (int k = 0; k < K; k++) X[(d + 4) & 3][k] = pair(px, off[k]);
            if (PRO)
            {
                const float *pc = planeXC(sp0 + d);
#pragma unroll
                for (int k = 0; k < K; k++) XC[(d + 4) & 3][k] = pc[offC[k]];
            }
        }
        if (haloRow)
        {
            const float *px = plane(xin, sp0);
#pragma unroll
            for (int k = 0; k < K; k++) O[0][k] = pair(px, offO[k]);
            if (PRO)
            {
                const float *pc = planeXC(sp0);
#pragma unroll
                for (int k = 0; k < K; k++) OC[0][k] = pc[offCO[k]];
            }
        }
    }
#pragma unroll
    for (int d = PRO ? -1 : 0; d <= 1; d++)
    {
        const float *pb = plane(b, sp0 + d);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (d >= 0) Bv[d][k] = pair(pb, off[k]);
        loadCodes((d + 4) & 3, sp0 + d);
    }
    if (PRO)
    {
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            fold(X[3][k], XC[3][k], C[3][k] & cm[k]);
            fold(X[0][k], XC[0][k], C[0][k] & cm[k]);
        }
    }
    __syncthreads();
    const int wUp = min(wave + 1, TM + 1), wDn = max(wave - 1, 0);

    // One plane step at sp = sp0 + 4g + U: first colour of plane sp (element E of every pair of this row), second
    // colour (and store) of plane sp - 1.
    auto step = [&](int sp, auto phase, auto colour) {
        constexpr int U = decltype(phase)::value;
        constexpr int E = decltype(colour)::value;
        constexpr int Sm2 = (U + 2) & 3, Sm1 = (U + 3) & 3, S0 = U & 3, S1 = (U + 1) & 3, S2 = (U + 2) & 3;
        constexpr int O0 = U & 1, O1 = (U + 1) & 1;
        // prefetch
        {
            const float *px = plane(xin, sp + 2);
            const float *pb = plane(b, sp + 2);
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                if (!ZERO) X[S2][k] = pair(px, off[k]);
                Bv[S2][k] = pair(pb, off[k]);
            }
            loadCodes(S2, sp + 2);
            if (!ZERO && haloRow)
            {
                const float *po = plane(xin, sp + 1);
#pragma unroll
                for (int k = 0; k < K; k++) O[O1][k] = pair(po, offO[k]);
            }
            if (PRO)
            {
                const float *pc2 = planeXC(sp + 2);
#pragma unroll
                for (int k = 0; k < K; k++) XC[S2][k] = pc2[offC[k]];
                if (haloRow)
                {
                    const float *pc1 = planeXC(sp + 1);
#pragma unroll
                    for (int k = 0; k < K; k++) OC[O1][k] = pc1[offCO[k]];
                }
            }
        }
        // publish this row's old x of plane sp for the neighbouring rows and pairs (zero past the row end)
        float(*ob)[2][kSlots] = old[sp & 1];
        if (!ZERO)
        {
            if (ragged)
            {
#pragma unroll
                for (int k = 0; k < K; k++)
                {
                    ob[wave][0][1 + lane + 64 * k] = okA[k] ? X[S0][k].x : 0.f;
                    ob[wave][1][1 + lane + 64 * k] = okB[k] ? X[S0][k].y : 0.f;
                }
            }
            else
            {
#pragma unroll
                for (int k = 0; k < K; k++)
                {
                    ob[wave][0][1 + lane + 64 * k] = X[S0][k].x;
                    ob[wave][1][1 + lane + 64 * k] = X[S0][k].y;
                }
            }
        }
        // The one barrier of the plane step: old[sp & 1] is complete, and so is red[(sp - 1) & 1], written in the
        // step before; a wave past it knows that every wave has finished the step before, whose readers of
        // old[(sp + 1) & 1] (its first colour) and of red[sp & 1] (its second colour, at plane sp - 2) are done.
        __syncthreads();
        if (PRO)
        {
#pragma unroll
            for (int k = 0; k < K; k++) fold(X[S1][k], XC[S1][k], C[S1][k] & cm[k]);
        }
        float rnew[K];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const unsigned code = ((C[S0][k] & cm[k]) >> (8 * E)) & 255u;
            const float bv = E ? Bv[S0][k].y : Bv[S0][k].x;
            float nb[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};   // s+1, s-1, m+1, m-1, f+1, f-1
            if (!ZERO)
            {
                const int p = 1 + lane + 64 * k;
                const float own = E ? X[S0][k].x : X[S0][k].y;                 // the pair's other element
                const float nbr = E ? ob[wave][0][p + 1] : ob[wave][1][p - 1];   // next pair's a / previous b
                float up = ob[wUp][E][p], dn = ob[wDn][E][p];   // rows m+1, m-1 at plane sp, element E
                const float outer = E ? O[O0][k].y : O[O0][k].x;
                if (wave == TM + 1) up = PRO ? ((code & 4u) ? outer + OC[O0][k] : 0.f) : outer;
                if (wave == 0) dn = PRO ? ((code & 8u) ? outer + OC[O0][k] : 0.f) : outer;
                nb[0] = E ? X[S1][k].y : X[S1][k].x;
                nb[1] = E ? X[Sm1][k].y : X[Sm1][k].x;
                nb[2] = up;
                nb[3] = dn;
                nb[4] = E ? nbr : own;
                nb[5] = E ? own : nbr;
            }
            float v;
            if (NU)
                v = (code & 64u) ? gs_stretched(L.c16, (int64_t)plane_off(sp) + off[k] + E, code, bv, nb, tablesS(sp),
                                                tM, tF[k][E])
                                 : 0.f;
            else
            {
                // (no leading 0 + : an iterate is never -0, so the sum is the same)
                float acc = 0.f;
                if (!ZERO)
                {
                    acc = nb[0];
                    acc += nb[1];
                    acc += nb[2];
                    acc += nb[3];
                    acc += nb[4];
                    acc += nb[5];
                }
                v = gs(acc, bv, code);
                if (code & 128u) v = gsDirichlet(acc, bv, sp, off[k] + (uint32_t)E);
            }
            rnew[k] = v;
            R[S0][k] = v;
        }
#pragma unroll
        for (int k = 0; k < K; k++) red[sp & 1][wave][fa[k] + E] = rnew[k];
        const int q = sp - 1;
        if (!haloRow && q >= s0 && q <= s1)
        {
            // second colour at plane q: the colour flips with the plane, so there the second-colour element is E
            float *po = xout + (size_t)q * (size_t)B.sS;
            const float(*rq)[kCols] = red[q & 1];
            double bx = 0.0;   // <b, x_out> of this lane's cells of the row
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                const unsigned cc = C[Sm1][k] & cm[k];
                const unsigned code = (cc >> (8 * E)) & 255u;
                const int f = fa[k] + E;
                const float bv = E ? Bv[Sm1][k].y : Bv[Sm1][k].x;
                float out;
                if (NU)
                {
                    const float nb[6] = {R[S0][k], R[Sm2][k], rq[wave + 1][f], rq[wave - 1][f],
                                         E ? rq[wave][f + 1] : R[Sm1][k], E ? R[Sm1][k] : rq[wave][f - 1]};
                    out = (code & 64u) ? gs_stretched(L.c16, (int64_t)q * B.sS + off[k] + E, code, bv, nb, tablesS(q),
                                                      tM, tF[k][E])
                                       : 0.f;
                }
                else
                {
                    float acc = R[S0][k];
                    acc += R[Sm2][k];
                    acc += rq[wave + 1][f];
                    acc += rq[wave - 1][f];
                    acc += E ? rq[wave][f + 1] : R[Sm1][k];
                    acc += E ? R[Sm1][k] : rq[wave][f - 1];
                    out = gs(acc, bv, code);
                    if (code & 128u) out = gsDirichlet(acc, bv, q, off[k] + (uint32_t)E);
                }
                const float va = E ? R[Sm1][k] : out, vb = E ? out : R[Sm1][k];
                if (dot)
                {
                    bx += (cc & 64u) ? (double)Bv[Sm1][k].x * (double)va : 0.0;
                    bx += (cc & (64u << 8)) ? (double)Bv[Sm1][k].y * (double)vb : 0.0;
                }
                if (okB[k])
                    *reinterpret_cast<float2 *>(po + off[k]) = make_float2(va, vb);
                else if (okA[k])
                    po[off[k]] = va;
            }
            if (dot)
            {
                const double rowDot = wave_sum(bx);
                if (lane == 0 && rowIn) dot[(size_t)(q - 1) * (size_t)(B.nm - 2) + (size_t)(m - 1)] = rowDot;
            }
        }
    };
    // the first-colour element at plane sp is (c1 - sp - m - 1) & 1: constant per wave at the first plane, then
    // alternating
    auto march = [&](auto colour0) {
        constexpr int E0 = decltype(colour0)::value;
        for (int sp = sp0; sp <= s1 + 1; sp += 4)
        {
            step(sp, std::integral_constant<int, 0>(), std::integral_constant<int, E0>());
            if (sp + 1 > s1 + 1) break;
            step(sp + 1, std::integral_constant<int, 1>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 2 > s1 + 1) break;
            step(sp + 2, std::integral_constant<int, 2>(), std::integral_constant<int, E0>());
            if (sp + 3 > s1 + 1) break;
            step(sp + 3, std::integral_constant<int, 3>(), std::integral_constant<int, E0 ^ 1>());
        }
    };
    if ((c1 - sp0 - m - 1) & 1)
        march(std::integral_constant<int, 1>());
    else
        march(std::integral_constant<int, 0>());
}

// ---- two sweeps in one march (whole-row waves, 3D uniform levels) ----------------------------------------------
//
// Both post-smoothing sweeps of a level as one launch (the first folding the coarse correction in, as
// k_rbgs_row<..., PRO> does): the second sweep marches three planes behind the first. A plane step sp computes the
// first sweep's first colour at plane sp (A1), its second colour at sp - 1 (A2), the second sweep's first colour at
// sp - 2 (B1) and its second colour at sp - 3 (B2, stored). A block is TM output rows and three halo rows each side
// (waves 0 and NW - 1 load their outer row, as the single sweep's halo waves do): A1 runs on every wave, A2 on all
// but the outermost, B1 on all but the two outermost, B2 on the TM output rows. Each stage's values pass to the next
// through LDS, double-buffered by plane parity, with the one barrier of the step; each stage's update is the single
// sweep's, from the same values in the same order, so x is bit for bit that of the two launches.
// b and codes stay in registers from their load (two planes ahead) to their last use by B2 (slots by plane mod 8).
template <int TM, int K>
__device__ __forceinline__ void rbgs2_row_body(const Level32 &L, const float *__restrict__ xin, float *__restrict__ xout,
                                               const float *__restrict__ b, int c1, int P, int tilesM,
                                               const SolveState *st, const Prolong &pro,
                                               int lb)
{
    if (st->done) return;
    constexpr int NW = TM + 6;
    const Box &B = L.B;
    constexpr int kCols = 128 * K + 2;
    constexpr int kSlots = 64 * K + 2;
    __shared__ float old[2][NW][2][kSlots];   // old x (plus the correction) of plane sp by parity, row, element, pair+1
    __shared__ float a1s[2][NW][kCols];       // first sweep, first colour, by column
    __shared__ float a2s[2][NW][kCols];       // first sweep, second colour
    __shared__ float b1s[2][NW][kCols];       // second sweep, first colour
    __shared__ float2 tab[128];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tm = lb % tilesM, chunk = lb / tilesM;
    const int m = tm * TM + wave - 2;   // m0 - 3 + wave with m0 = 1 + tm * TM
    const bool outerW = wave == 0 || wave == NW - 1;
    const bool doA2 = wave >= 1 && wave <= NW - 2, doB1 = wave >= 2 && wave <= NW - 3;
    const bool doB2 = wave >= 3 && wave <= NW - 4;
    const bool rowIn = m >= 1 && m <= B.nm - 2;
    const bool ragged = B.nf - 2 < 128 * K;
    const int mc = min(max(m, 0), B.nm - 1);
    const int mo = min(max(wave == 0 ? m - 1 : m + 1, 0), B.nm - 1);
    const int s0 = 1 + chunk * P, s1 = min(s0 + P - 1, B.ns - 2);
    const uint32_t sM = (uint32_t)B.sM;
    const int fcMax = ((B.nf - 2) & 1) ? B.nf - 2 : B.nf - 3;
    uint32_t off[K], offO[K], offC[K], offCO[K];
    bool okA[K], okB[K];
    unsigned cm[K];
    int fa[K];
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        fa[k] = 1 + 2 * (lane + 64 * k);
        okA[k] = rowIn && fa[k] <= B.nf - 2;
        okB[k] = rowIn && fa[k] + 1 <= B.nf - 2;
        cm[k] = (okA[k] ? 0xFFu : 0u) | (okB[k] ? 0xFF00u : 0u);
        const int fc = min(fa[k], fcMax);
        off[k] = (uint32_t)mc * sM + (uint32_t)fc;
        offO[k] = (uint32_t)mo * sM + (uint32_t)fc;
        const int Fc = min(lane + 64 * k + 1, pro.nf - 1);
        offC[k] = (uint32_t)((mc + 1) >> 1) * (uint32_t)pro.sM + (uint32_t)Fc;
        offCO[k] = (uint32_t)((mo + 1) >> 1) * (uint32_t)pro.sM + (uint32_t)Fc;
    }
    auto planeXC = [&](int pl) {
        const int sc = (min(max(pl, 0), B.ns - 1) + 1) >> 1;
        return pro.xc + (size_t)min(sc, pro.ns - 1) * (size_t)pro.sS;
    };
    for (int i = threadIdx.x; i < 2 * NW * kCols; i += 64 * NW)
    {
        (&a1s[0][0][0])[i] = 0.f;
        (&a2s[0][0][0])[i] = 0.f;
        (&b1s[0][0][0])[i] = 0.f;
    }
    for (int i = threadIdx.x; i < 2 * NW * 2 * kSlots; i += 64 * NW) (&old[0][0][0][0])[i] = 0.f;
    if (threadIdx.x < 128)
    {
        const unsigned c = threadIdx.x;
        tab[c] = make_float2((c & 64u) ? inv_count(c) : 0.f, __uint_as_float((c & 64u) ? ~0u : 0u));
    }
    auto plane = [&](const float *base, int pl) { return base + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto planeC = [&](int pl) { return L.c8 + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto pair = [](const float *row, uint32_t o) { return *reinterpret_cast<const float2 *>(row + o); };
    auto pairC = [](const uint8_t *row, uint32_t o) { return (unsigned)*reinterpret_cast<const uint16_t *>(row + o); };
    auto fold = [](float2 &x, float xc, unsigned cc) {
        x.x += (cc & 64u) ? xc : 0.f;
        x.y += (cc & (64u << 8)) ? xc : 0.f;
    };
    // acc - b / cu as one fma, as the single sweep's contracted form: here each plane's b feeds two stages, and a
    // product b / cu shared by both would otherwise be rounded before the subtraction
    auto gs = [&](float acc, float bv, unsigned code) {
        const float2 t = tab[code & 127u];
        return __uint_as_float(__float_as_uint(__builtin_fmaf(-bv, L.icu, acc) * t.x) & __float_as_uint(t.y));
    };
    auto gsDirichlet = [&](float acc, float bv, int pl, uint32_t o) {
        const unsigned c16 = L.c16[(size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS + o];
        float d = 0.f;
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned stt = (c16 >> (2 * q)) & 3u;
            d += stt == 3u ? 2.f : (stt != 0u ? 1.f : 0.f);
        }
        return d != 0.f ? __builtin_fmaf(-bv, L.icu, acc) / d : 0.f;
    };
    // update of element E of a pair from its six neighbours (s+1, s-1, m+1, m-1, f+1, f-1), as the single sweep
    auto update = [&](const float (&nb)[6], float bv, unsigned code, int pl, uint32_t o) {
        float acc = nb[0];
        acc += nb[1];
        acc += nb[2];
        acc += nb[3];
        acc += nb[4];
        acc += nb[5];
        float v = gs(acc, bv, code);
        if (code & 128u) v = gsDirichlet(acc, bv, pl, o);
        return v;
    };
    // slots: X, XC by plane mod 4 (sp - 1 .. sp + 2), O, OC by plane mod 2, b and codes by plane mod 8 (sp + 2 ..
    // sp - 3), the stages' own values by plane mod 4
    float2 X[4][K], O[2][K], Bv[8][K];
    unsigned C[8][K];
    float XC[4][K], OC[2][K];
    float R1[4][K], R2[4][K], RB[4][K];
#pragma unroll
    for (int k = 0; k < K; k++)
    {
#pragma unroll
        for (int j = 0; j < 8; j++)
        {
            Bv[j][k] = make_float2(0.f, 0.f);
            C[j][k] = 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
        {
            X[j][k] = make_float2(0.f, 0.f);
            XC[j][k] = R1[j][k] = R2[j][k] = RB[j][k] = 0.f;
        }
        O[0][k] = O[1][k] = make_float2(0.f, 0.f);
        OC[0][k] = OC[1][k] = 0.f;
    }
    const int sp0 = s0 - 3;
    // plane sp0 + d -> slot d & 3 (d & 7 for b and codes, d & 1 for O)
#pragma unroll
    for (int d = -1; d <= 1; d++)
    {
        const float *px = plane(xin, sp0 + d);
        const float *pc = planeXC(sp0 + d);
        const float *pb = plane(b, sp0 + d);
        const uint8_t *pcc = planeC(sp0 + d);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            X[(d + 4) & 3][k] = pair(px, off[k]);
            XC[(d + 4) & 3][k] = pc[offC[k]];
            C[(d + 8) & 7][k] = pairC(pcc, off[k]);
            if (d >= 0) Bv[d][k] = pair(pb, off[k]);
        }
    }
    if (outerW)
    {
        const float *px = plane(xin, sp0);
        const float *pc = planeXC(sp0);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            O[0][k] = pair(px, offO[k]);
            OC[0][k] = pc[offCO[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        fold(X[3][k], XC[3][k], C[7][k] & cm[k]);
        fold(X[0][k], XC[0][k], C[0][k] & cm[k]);
    }
    __syncthreads();
    const int wUp = min(wave + 1, NW - 1), wDn = max(wave - 1, 0);

    auto step = [&](int sp, auto phase, auto colour) {
        constexpr int U = decltype(phase)::value;
        constexpr int E = decltype(colour)::value;
        constexpr int Xm1 = (U + 3) & 3, X0 = U & 3, X1 = (U + 1) & 3, X2 = (U + 2) & 3;
        constexpr int B0 = U & 7, Bm1 = (U + 7) & 7, Bm2 = (U + 6) & 7, Bm3 = (U + 5) & 7, B2s = (U + 2) & 7;
        constexpr int O0 = U & 1, O1 = (U + 1) & 1;
        // R slots: plane p -> (p - sp0) & 3
        constexpr int R0 = U & 3, Rm1 = (U + 3) & 3, Rm2 = (U + 2) & 3, Rm3 = (U + 1) & 3, Rm4 = U & 3;
        // prefetch: plane sp + 2 (x, b, codes, coarse x); the outer row of plane sp + 1
        {
            const float *px = plane(xin, sp + 2);
            const float *pb = plane(b, sp + 2);
            const uint8_t *pcc = planeC(sp + 2);
            const float *pc2 = planeXC(sp + 2);
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                X[X2][k] = pair(px, off[k]);
                Bv[B2s][k] = pair(pb, off[k]);
                C[B2s][k] = pairC(pcc, off[k]);
                XC[X2][k] = pc2[offC[k]];
            }
            if (outerW)
            {
                const float *po = plane(xin, sp + 1);
                const float *pc1 = planeXC(sp + 1);
#pragma unroll
                for (int k = 0; k < K; k++)
                {
                    O[O1][k] = pair(po, offO[k]);
                    OC[O1][k] = pc1[offCO[k]];
                }
            }
        }
        // publish this row's old x (with the correction) of plane sp
        float(*ob)[2][kSlots] = old[sp & 1];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            ob[wave][0][1 + lane + 64 * k] = (!ragged || okA[k]) ? X[X0][k].x : 0.f;
            ob[wave][1][1 + lane + 64 * k] = (!ragged || okB[k]) ? X[X0][k].y : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) fold(X[X1][k], XC[X1][k], C[(U + 1) & 7][k] & cm[k]);
        // A1: first sweep, first colour, plane sp (every wave)
        {
            float v[K];
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                const unsigned code = ((C[B0][k] & cm[k]) >> (8 * E)) & 255u;
                const int p = 1 + lane + 64 * k;
                const float own = E ? X[X0][k].x : X[X0][k].y;
                const float nbr = E ? ob[wave][0][p + 1] : ob[wave][1][p - 1];
                float up = ob[wUp][E][p], dn = ob[wDn][E][p];
                const float outer = E ? O[O0][k].y : O[O0][k].x;
                if (wave == NW - 1) up = (code & 4u) ? outer + OC[O0][k] : 0.f;
                if (wave == 0) dn = (code & 8u) ? outer + OC[O0][k] : 0.f;
                const float nb[6] = {E ? X[X1][k].y : X[X1][k].x, E ? X[Xm1][k].y : X[Xm1][k].x, up, dn,
                                     E ? nbr : own, E ? own : nbr};
                v[k] = update(nb, E ? Bv[B0][k].y : Bv[B0][k].x, code, sp, off[k] + (uint32_t)E);
                R1[R0][k] = v[k];
            }
#pragma unroll
            for (int k = 0; k < K; k++) a1s[sp & 1][wave][fa[k] + E] = v[k];
        }
        // A2: first sweep, second colour, plane sp - 1 (element E there)
        if (doA2)
        {
            const int q = sp - 1;
            const float(*rq)[kCols] = a1s[q & 1];
            float v[K];
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                const unsigned code = ((C[Bm1][k] & cm[k]) >> (8 * E)) & 255u;
                const int f = fa[k] + E;
                const float nb[6] = {R1[R0][k], R1[Rm2][k], rq[wave + 1][f], rq[wave - 1][f],
                                     E ? rq[wave][f + 1] : R1[Rm1][k], E ? R1[Rm1][k] : rq[wave][f - 1]};
                v[k] = update(nb, E ? Bv[Bm1][k].y : Bv[Bm1][k].x, code, q, off[k] + (uint32_t)E);
                R2[Rm1][k] = v[k];
            }
#pragma unroll
            for (int k = 0; k < K; k++) a2s[q & 1][wave][fa[k] + E] = v[k];
        }
        // B1: second sweep, first colour, plane sp - 2 (element E there): from the first sweep's values
        if (doB1)
        {
            const int q = sp - 2;
            const float(*rq)[kCols] = a2s[q & 1];
            float v[K];
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                const unsigned code = ((C[Bm2][k] & cm[k]) >> (8 * E)) & 255u;
                const int f = fa[k] + E;
                const float nb[6] = {R2[Rm1][k], R2[Rm3][k], rq[wave + 1][f], rq[wave - 1][f],
                                     E ? rq[wave][f + 1] : R2[Rm2][k], E ? R2[Rm2][k] : rq[wave][f - 1]};
                v[k] = update(nb, E ? Bv[Bm2][k].y : Bv[Bm2][k].x, code, q, off[k] + (uint32_t)E);
                RB[Rm2][k] = v[k];
            }
#pragma unroll
            for (int k = 0; k < K; k++) b1s[q & 1][wave][fa[k] + E] = v[k];
        }
        // B2: second sweep, second colour, plane sp - 3 (element E there), stored
        const int q = sp - 3;
        if (doB2 && q >= s0 && q <= s1)
        {
            float *po = xout + (size_t)q * (size_t)B.sS;
            const float(*rq)[kCols] = b1s[q & 1];
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                const unsigned cc = C[Bm3][k] & cm[k];
                const unsigned code = (cc >> (8 * E)) & 255u;
                const int f = fa[k] + E;
                const float nb[6] = {RB[Rm2][k], RB[Rm4][k], rq[wave + 1][f], rq[wave - 1][f],
                                     E ? rq[wave][f + 1] : RB[Rm3][k], E ? RB[Rm3][k] : rq[wave][f - 1]};
                const float out = update(nb, E ? Bv[Bm3][k].y : Bv[Bm3][k].x, code, q, off[k] + (uint32_t)E);
                const float va = E ? RB[Rm3][k] : out, vb = E ? out : RB[Rm3][k];
                if (okB[k])
                    *reinterpret_cast<float2 *>(po + off[k]) = make_float2(va, vb);
                else if (okA[k])
                    po[off[k]] = va;
            }
        }
    };
    // the first-colour element at plane sp is (c1 - sp - m - 1) & 1
    auto march = [&](auto colour0) {
        constexpr int E0 = decltype(colour0)::value;
        for (int sp = sp0; sp <= s1 + 3; sp += 8)
        {
            step(sp, std::integral_constant<int, 0>(), std::integral_constant<int, E0>());
            if (sp + 1 > s1 + 3) break;
            step(sp + 1, std::integral_constant<int, 1>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 2 > s1 + 3) break;
            step(sp + 2, std::integral_constant<int, 2>(), std::integral_constant<int, E0>());
            if (sp + 3 > s1 + 3) break;
            step(sp + 3, std::integral_constant<int, 3>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 4 > s1 + 3) break;
            step(sp + 4, std::integral_constant<int, 4>(), std::integral_constant<int, E0>());
            if (sp + 5 > s1 + 3) break;
            step(sp + 5, std::integral_constant<int, 5>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 6 > s1 + 3) break;
            step(sp + 6, std::integral_constant<int, 6>(), std::integral_constant<int, E0>());
            if (sp + 7 > s1 + 3) break;
            step(sp + 7, std::integral_constant<int, 7>(), std::integral_constant<int, E0 ^ 1>());
        }
    };
    if ((c1 - sp0 - m - 1) & 1)
        march(std::integral_constant<int, 1>());
    else
        march(std::integral_constant<int, 0>());
}

// Coarse cell (S, M, F) gathers the residual b - A x of its fine children (fine indices 2C-1, 2C along each
// coarsened axis, clipped to the fine interior): bc = sum (R = P^T for the injection prolongation).
__global__ __launch_bounds__(kBlockThreads) void k_residual_restrict(Level32 Lf, const float *__restrict__ x,
                                                                     const float *__restrict__ b, Level32 Lc,
                                                                     float *__restrict__ bc, const SolveState *st)
{
    if (st->done) return;
    const Box &Bf = Lf.B, &Bc = Lc.B;
    const TileRange tr = my_tiles(Bc.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int S, M, F;
        if (!tile_cell(Bc, t, S, M, F)) continue;
        float sum = 0.f;
        const int s0 = Bf.threeD ? 2 * S - 1 : 0, s1 = Bf.threeD ? min(2 * S, Bf.ns - 2) : 0;
        const int m1 = min(2 * M, Bf.nm - 2), f1 = min(2 * F, Bf.nf - 2);
        for (int s = s0; s <= s1; s++)
            for (int m = 2 * M - 1; m <= m1; m++)
                for (int f = 2 * F - 1; f <= f1; f++)
                {
                    const int64_t idx = Bf.at(s, m, f);
                    const unsigned code = Lf.c8[idx];
                    if (!(code & 64u)) continue;
                    sum += b[idx] - apply32(Lf, s, m, f, idx, code, x);
                }
        bc[Bc.at(S, M, F)] = sum;
    }
}

// x += P xc: every fine unknown takes its parent's correction when the parent is a coarse unknown.
__global__ __launch_bounds__(kBlockThreads) void k_prolong(Level32 Lf, float *__restrict__ x, Level32 Lc,
                                                           const float *__restrict__ xc, const SolveState *st)
{
    if (st->done) return;
    const Box &Bf = Lf.B, &Bc = Lc.B;
    const TileRange tr = my_tiles(Bf.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(Bf, t, s, m, f)) continue;
        const int64_t idx = Bf.at(s, m, f);
        if (!(Lf.c8[idx] & 64u)) continue;
        const int64_t pidx = Bc.at(Bf.threeD ? (s + 1) >> 1 : 0, (m + 1) >> 1, (f + 1) >> 1);
        const float v = xc[pidx];
        x[idx] += (Lc.c8[pidx] & 64u) ? v : 0.f;
    }
}

// Coarsest level: dense FP64 pseudo-inverse times the gathered right-hand side. One wavefront per row.
__device__ __forceinline__ void coarse_solve_body(int n, const int32_t *__restrict__ order,
                                                  const double *__restrict__ M, const float *__restrict__ b,
                                                  float *__restrict__ x, const SolveState *st, int blk)
{
    if (st->done) return;
    __shared__ double rb[kMaxDenseCoarse];
    constexpr int kPer = kMaxDenseCoarse / 64;
    const int lane = threadIdx.x & 63;
    const int row = blk * (kBlockThreads / 64) + (threadIdx.x >> 6);
    // this lane's entries of the row, loaded before the gather of b so that the two loads overlap
    const double *Mr = M + (size_t)min(row, n - 1) * n;
    double mr[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) mr[j] = (row < n && lane + 64 * j < n) ? Mr[lane + 64 * j] : 0.0;
    for (int q = threadIdx.x; q < n; q += kBlockThreads) rb[q] = (double)b[order[q]];
    __syncthreads();
    if (row >= n) return;
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kPer; j++)
        if (lane + 64 * j < n) acc += mr[j] * rb[lane + 64 * j];
    acc = wave_sum(acc);
    if (lane == 0) x[order[row]] = (float)acc;
}

// ---- FP32 CG kernels -----------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlockThreads) void k_cg_rz(Level32 L, const float *__restrict__ r,
                                                         const float *__restrict__ z, const SolveState *st,
                                                         Red red)
{
    if (st->done) return;
    const Box &B = L.B;
    const TileRange tr = my_tiles(B.blocks());
    double v = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (L.c8[idx] & 64u) v += (double)r[idx] * (double)z[idx];
    }
    double a[1] = {v};
    block_sum<1>(a, red, gridDim.x, tr.lb);
}

// <r, z> over the unknowns of each interior row and plane of a 3D box, as the whole-row sweep writes it when it is
// the V-cycle's last one (one wave per row; lane l sums the cell pairs (1 + 2p, 2 + 2p), p = l, l + 64, ..., in
// that order, then the wave sums its lanes), so a level swept by another kernel gets the same partials bit for bit
// and k_rz_rows_sum the same <r, z>.
__global__ __launch_bounds__(kBlockThreads) void k_rz_rows(Level32 L, const float *__restrict__ r,
                                                           const float *__restrict__ z, const SolveState *st,
                                                           double *__restrict__ rows)
{
    if (st->done) return;
    const Box &B = L.B;
    const int lane = threadIdx.x & 63;
    const int nrm = B.nm - 2;
    const int64_t row = (int64_t)blockIdx.x * (kBlockThreads / 64) + (threadIdx.x >> 6);
    if (row >= (int64_t)(B.ns - 2) * nrm) return;   // whole waves: the row is wave-uniform
    const int q = (int)(row / nrm) + 1, m = (int)(row % nrm) + 1;
    const int64_t base = B.at(q, m, 0);
    double bx = 0.0;
    for (int p = lane; 2 * p + 1 <= B.nf - 2; p += 64)
    {
        const int64_t a = base + 1 + 2 * p;
        bx += (L.c8[a] & 64u) ? (double)r[a] * (double)z[a] : 0.0;
        if (2 * p + 2 <= B.nf - 2) bx += (L.c8[a + 1] & 64u) ? (double)r[a + 1] * (double)z[a + 1] : 0.0;
    }
    const double v = wave_sum(bx);
    if (lane == 0) rows[row] = v;
}

__global__ __launch_bounds__(kBlockThreads) void k_cg_direction(Level32 L, const float *__restrict__ z,
                                                                float *__restrict__ p, const SolveState *st)
{
    if (st->done) return;
    const Box &B = L.B;
    const int first = st->first;
    const float beta = (float)st->beta;
    const TileRange tr = my_tiles(B.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (L.c8[idx] & 64u) p[idx] = first ? z[idx] : z[idx] + beta * p[idx];
    }
}

__global__ __launch_bounds__(kBlockThreads) void k_cg_spmv(Level32 L, const float *__restrict__ p,
                                                           float *__restrict__ q, const SolveState *st,
                                                           Red red)
{
    if (st->done) return;
    const Box &B = L.B;
    const TileRange tr = my_tiles(B.blocks());
    double pq = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        const unsigned code = L.c8[idx];
        if (!(code & 64u)) continue;
        const float pc = p[idx];
        const float v = apply32(L, s, m, f, idx, code, p);
        q[idx] = v;
        pq += (double)pc * (double)v;
    }
    double a[1] = {pq};
    block_sum<1>(a, red, gridDim.x, tr.lb);
}

// p is pA on even inner iterations and pB on odd ones (the fused direction kernel ping-pongs; pass pA == pB when
// the direction is updated in place).
__global__ __launch_bounds__(kBlockThreads) void k_cg_update(Level32 L, float *__restrict__ e, float *__restrict__ r,
                                                             const float *__restrict__ pA,
                                                             const float *__restrict__ pB, const float *__restrict__ q,
                                                             const SolveState *st, Red red)
{
    if (st->done) return;
    const Box &B = L.B;
    const int firstIter = st->iters == 0;
    const float *__restrict__ p = (st->iters & 1) ? pB : pA;
    const float alpha = (float)st->alpha;
    const TileRange tr = my_tiles(B.blocks());
    double rr = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (!(L.c8[idx] & 64u)) continue;
        const float pv = p[idx];
        e[idx] = firstIter ? alpha * pv : e[idx] + alpha * pv;
        const float rv = r[idx] - alpha * q[idx];
        r[idx] = rv;
        rr += (double)rv * (double)rv;
    }
    double a[1] = {rr};
    block_sum<1>(a, red, gridDim.x, tr.lb);
}

// k_cg_update over the whole box as aligned quads, without the codes: outside the unknowns p, q and r are zero
// (never written, or written as zero), so there e and r stay zero and add nothing to <r, r>. Quad j holds elements
// 4j-1 .. 4j+2 of the pair-aligned arrays (element -1 is the allocation's leading pad).
__global__ __launch_bounds__(kBlockThreads) void k_cg_update_v(int64_t nq, float4 *__restrict__ e,
                                                               float4 *__restrict__ r,
                                                               const float4 *__restrict__ pA,
                                                               const float4 *__restrict__ pB,
                                                               const float4 *__restrict__ q, const SolveState *st,
                                                               Red red)
{
    if (st->done) return;
    const bool firstIter = st->iters == 0;
    const float alpha = (float)st->alpha;
    const float4 *__restrict__ p = (st->iters & 1) ? pB : pA;
    double rr = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * 2;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x * 2 + threadIdx.x; j < nq; j += stride)
    {
        const int64_t j2 = j + blockDim.x;
        const bool two = j2 < nq;
        const float4 pv = p[j], qv = q[j], rv0 = r[j];
        const float4 pv2 = two ? p[j2] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 qv2 = two ? q[j2] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 rv2 = two ? r[j2] : make_float4(0.f, 0.f, 0.f, 0.f);
        float4 ev = make_float4(0.f, 0.f, 0.f, 0.f), ev2 = ev;
        if (!firstIter)
        {
            ev = e[j];
            if (two) ev2 = e[j2];
        }
        auto upd = [&](float4 &ee, float4 &rr4, const float4 &pp, const float4 &qq) {
            ee.x = firstIter ? alpha * pp.x : ee.x + alpha * pp.x;
            ee.y = firstIter ? alpha * pp.y : ee.y + alpha * pp.y;
            ee.z = firstIter ? alpha * pp.z : ee.z + alpha * pp.z;
            ee.w = firstIter ? alpha * pp.w : ee.w + alpha * pp.w;
            rr4.x = rr4.x - alpha * qq.x;
            rr4.y = rr4.y - alpha * qq.y;
            rr4.z = rr4.z - alpha * qq.z;
            rr4.w = rr4.w - alpha * qq.w;
        };
        float4 r1 = rv0, r2 = rv2;
        upd(ev, r1, pv, qv);
        e[j] = ev;
        r[j] = r1;
        rr += (double)r1.x * (double)r1.x + (double)r1.y * (double)r1.y + (double)r1.z * (double)r1.z +
              (double)r1.w * (double)r1.w;
        if (two)
        {
            upd(ev2, r2, pv2, qv2);
            e[j2] = ev2;
            r[j2] = r2;
            rr += (double)r2.x * (double)r2.x + (double)r2.y * (double)r2.y + (double)r2.z * (double)r2.z +
                  (double)r2.w * (double)r2.w;
        }
    }
    double a[1] = {rr};
    block_sum<1>(a, red, gridDim.x, blockIdx.x);
}

// ---- 2.5D marching stencil kernels (3D boxes) ------------------------------------------------------------------
// A block owns a 64 (fast) x 4 (mid) column of the interior and walks a chunk of slow planes, keeping the slow-axis
// neighbours of its cells in registers: every plane of the field is read from memory once, and the in-plane
// neighbours come from lines the block's own waves have just brought into L1.
struct March
{
    int f, m, s0, s1, lb;
    bool in;
};

__device__ __forceinline__ March march_setup(const Box &B, int P, int lb)
{
    March M;
    M.lb = lb;
    const int tf = M.lb % B.tilesF, t2 = M.lb / B.tilesF, tm = t2 % B.tilesM, chunk = t2 / B.tilesM;
    M.f = 1 + tf * kBlockFast + (threadIdx.x & (kBlockFast - 1));
    M.m = 1 + tm * kBlockMid + (threadIdx.x / kBlockFast);
    M.s0 = 1 + chunk * P;
    M.s1 = min(M.s0 + P - 1, B.ns - 2);
    M.in = M.f <= B.nf - 2 && M.m <= B.nm - 2;
    return M;
}

__device__ __forceinline__ March march_setup(const Box &B, int P) { return march_setup(B, P, logical_block()); }

// (A x) of one cell in FP32 from its value and its six neighbours (slow+, slow-, mid+, mid-, fast+, fast-).
__device__ __forceinline__ float apply32v(const Level32 &L, int s, int m, int f, int64_t idx, unsigned code, float xc,
                                          const float (&nb)[6])
{
    float acc = 0.f;
    if (fast_cell(L, code))
    {
#pragma unroll
        for (int q = 0; q < 6; q++) acc += ((code >> q) & 1u) ? nb[q] : 0.f;
        return L.cu * (acc - (float)__popc(code & 63u) * xc);
    }
    const Op32 o = op32(L, s, m, f, idx, code);
#pragma unroll
    for (int q = 0; q < 6; q++) acc += ((o.coupled >> q) & 1u) ? o.c[q] * nb[q] : 0.f;
    return acc + o.diag * xc;
}

// (A x) of one cell in FP64 with the reference's exact arithmetic, from its value and neighbours (order as above):
// the same sums in the same order as apply64 / row64.
__device__ __forceinline__ double apply64v(const Box &B, const Tables64 &T, const uint8_t *ty, int s, int m, int f,
                                           int64_t idx, double xc, const double (&nb)[6], unsigned code)
{
#pragma clang fp contract(off)
    if (T.unitCoef)
    {
        if (!(code & 128u))
        {
            const double v = T.scale * 1.0;
            const double dv = T.scale * (-(double)__popc(code & 63u));
            double t = 0.0;
            if (code & 2u) t += v * nb[1];
            if (code & 8u) t += v * nb[3];
            if (code & 32u) t += v * nb[5];
            if (code & 63u) t += dv * xc;
            if (code & 16u) t += v * nb[4];
            if (code & 4u) t += v * nb[2];
            if (code & 1u) t += v * nb[0];
            return t;
        }
    }
    const Op64 o = op64(B, T, ty, s, m, f, idx);
    double t = 0.0;
    if (o.coupled & 2u) t += o.v[1] * nb[1];
    if (o.coupled & 8u) t += o.v[3] * nb[3];
    if (o.coupled & 32u) t += o.v[5] * nb[5];
    if (o.diag != 0.0) t += o.dval * xc;
    if (o.coupled & 16u) t += o.v[4] * nb[4];
    if (o.coupled & 4u) t += o.v[2] * nb[2];
    if (o.coupled & 1u) t += o.v[0] * nb[0];
    return t;
}

// (A x) of one cell of a unit-coefficient (3D uniform) grid in FP64 from its value and neighbours (order as
// apply64v), without the coefficient tables or cell types: every off-diagonal is scale, and the diagonal is
// -scale times the count of coupled and ring-FLUID neighbours plus interior Dirichlet neighbours plus twice the
// ring Dirichlet ones (the ring factor FaceGap / (Width / 2) is exactly 2 on uniform spacing). Those are small
// integers, so the entries and the column-ordered sum are op64 / row64's bit for bit.
__device__ __forceinline__ double apply64u(const Tables64 &T, int64_t idx, double xc, const double (&nb)[6],
                                           unsigned code)
{
#pragma clang fp contract(off)
    unsigned coupled = code & 63u, n;
    if (!(code & 128u))
        n = __popc(coupled);
    else
    {
        const unsigned c16 = T.c16[idx];
        n = 0;
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned st = (c16 >> (2 * q)) & 3u;
            n += st == 3u ? 2u : (st != 0u ? 1u : 0u);
        }
    }
    const double v = T.scale * 1.0;
    const double dv = T.scale * (-(double)n);
    double t = 0.0;
    if (coupled & 2u) t += v * nb[1];
    if (coupled & 8u) t += v * nb[3];
    if (coupled & 32u) t += v * nb[5];
    if (n) t += dv * xc;
    if (coupled & 16u) t += v * nb[4];
    if (coupled & 4u) t += v * nb[2];
    if (coupled & 1u) t += v * nb[0];
    return t;
}

__global__ __launch_bounds__(kBlockThreads) void k_cg_spmv_m(Level32 L, const float *__restrict__ p,
                                                             float *__restrict__ q, const SolveState *st,
                                                             Red red, int P)
{
    if (st->done) return;
    const Box &B = L.B;
    const March M = march_setup(B, P);
    double pq = 0.0;
    if (M.in)
    {
        int64_t idx = B.at(M.s0, M.m, M.f);
        float pm = p[idx - B.sS], pc = p[idx];
        for (int s = M.s0; s <= M.s1; s++, idx += B.sS)
        {
            const float pp = p[idx + B.sS];
            const unsigned code = L.c8[idx];
            if (code & 64u)
            {
                const float nb[6] = {pp, pm, p[idx + B.sM], p[idx - B.sM], p[idx + 1], p[idx - 1]};
                const float v = apply32v(L, s, M.m, M.f, idx, code, pc, nb);
                q[idx] = v;
                pq += (double)pc * (double)v;
            }
            pm = pc;
            pc = pp;
        }
    }
    double a[1] = {pq};
    block_sum<1>(a, red, gridDim.x, M.lb);
}

// The CG direction and its product in one pass (3D boxes): p_new = z + beta p_old (z on the first direction),
// q = A p_new, and <p_new, q>. p ping-pongs between pA (even inner iterations) and pB (odd), so every block reads
// only old directions. A plane's p_new is formed once per cell (and per halo cell) into LDS; the slow neighbours
// stay in registers; every load is issued a plane ahead (the cell's own values two planes ahead).
__global__ __launch_bounds__(kBlockThreads) void k_cg_dirspmv_t(Level32 L, const float *__restrict__ z,
                                                                float *__restrict__ pA, float *__restrict__ pB,
                                                                float *__restrict__ q, const SolveState *st,
                                                                Red red, int P)
{
    if (st->done) return;
    const Box &B = L.B;
    const int first = st->first;
    const float beta = (float)st->beta;
    const bool odd = st->iters & 1;
    float *__restrict__ pn = odd ? pB : pA;
    const float *__restrict__ po = odd ? pA : pB;
    __shared__ float Ps[kBlockMid + 2][kBlockFast + 2];
    const March M = march_setup(B, P);
    const int tx = threadIdx.x & (kBlockFast - 1), tyy = threadIdx.x / kBlockFast;
    const int f0 = M.f - tx, m0 = M.m - tyy;
    const int fc = min(M.f, B.nf - 1), mcl = min(M.m, B.nm - 1);
    const int64_t own = (int64_t)mcl * B.sM + fc;
    int hx = -1, hy = -1;
    const int t = threadIdx.x;
    if (t < 64) { hy = 0; hx = t + 1; }
    else if (t < 128) { hy = kBlockMid + 1; hx = t - 64 + 1; }
    else if (t < 128 + kBlockMid) { hy = t - 128 + 1; hx = 0; }
    else if (t < 128 + 2 * kBlockMid) { hy = t - 128 - kBlockMid + 1; hx = kBlockFast + 1; }
    const bool halo = hx >= 0;
    const int64_t hoff = halo ? (int64_t)min(max(m0 - 1 + hy, 0), B.nm - 1) * B.sM + min(max(f0 - 1 + hx, 0), B.nf - 1)
                              : own;
    auto PV = [&](int64_t i) { return first ? z[i] : z[i] + beta * po[i]; };
    auto plane = [&](int sl) { return (int64_t)min(max(sl, 0), B.ns - 1) * B.sS; };
    double pq = 0.0;
    float Pm = PV(plane(M.s0 - 1) + own), Pc = PV(plane(M.s0) + own), Pp = PV(plane(M.s0 + 1) + own);
    float H = halo ? PV(plane(M.s0) + hoff) : 0.f;
    unsigned cc = L.c8[plane(M.s0) + own];
    for (int sl = M.s0; sl <= M.s1; sl++)
    {
        const float Pn = PV(plane(sl + 2) + own);
        const float Hn = halo ? PV(plane(sl + 1) + hoff) : 0.f;
        const unsigned cn = L.c8[plane(sl + 1) + own];
        Ps[tyy + 1][tx + 1] = Pc;
        if (halo) Ps[hy][hx] = H;
        __syncthreads();
        const int64_t idx = plane(sl) + own;
        if (M.in && (cc & 64u))
        {
            const float nb[6] = {Pp, Pm, Ps[tyy + 2][tx + 1], Ps[tyy][tx + 1], Ps[tyy + 1][tx + 2], Ps[tyy + 1][tx]};
            const float v = apply32v(L, sl, M.m, M.f, idx, cc, Pc, nb);
            pn[idx] = Pc;
            q[idx] = v;
            pq += (double)Pc * (double)v;
        }
        __syncthreads();
        Pm = Pc;
        Pc = Pp;
        Pp = Pn;
        H = Hn;
        cc = cn;
    }
    double a[1] = {pq};
    block_sum<1>(a, red, gridDim.x, M.lb);
}

// The halo assignment of the 64 x 4 march tiles: threads 0-63 the row above the tile, 64-127 the row below,
// 128-131 the column left of it, 132-135 the column right of it (LDS position hy, hx; hx = -1: none).
__device__ __forceinline__ void tile_halo(int t, int &hy, int &hx)
{
    hx = -1;
    hy = -1;
    if (t < 64) { hy = 0; hx = t + 1; }
    else if (t < 128) { hy = kBlockMid + 1; hx = t - 64 + 1; }
    else if (t < 128 + kBlockMid) { hy = t - 128 + 1; hx = 0; }
    else if (t < 128 + 2 * kBlockMid) { hy = t - 128 - kBlockMid + 1; hx = kBlockFast + 1; }
}

// FP64 Krylov around the FP32 preconditioner, part 1 (3D boxes): p = z + beta p_old formed in FP64 and stored in
// FP32, and <p, A p> with A p computed in the reference's FP64 arithmetic from the stored FP32 p (part 2 applies
// A to the same FP32 values, so x and r stay consistent). p ping-pongs between pA (even inner iterations) and pB
// (odd); plane p staged in LDS, slow neighbours in registers, loads a plane ahead.
template <bool UNI>
__global__ __launch_bounds__(kBlockThreads) void k_cg_dir64_t(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                              const float *__restrict__ z, float *__restrict__ ring,
                                                              int64_t stride, const SolveState *st, Red red, int P)
{
#pragma clang fp contract(off)
    if (st->done) return;
    const int first = st->first;
    const double beta = st->beta;
    const int slot = st->iters % kXRing;
    float *__restrict__ pn = ring + (size_t)slot * (size_t)stride;
    const float *__restrict__ po = ring + (size_t)((slot + kXRing - 1) % kXRing) * (size_t)stride;
    __shared__ double Ps[kBlockMid + 2][kBlockFast + 2];
    const March M = march_setup(B, P);
    const int tx = threadIdx.x & (kBlockFast - 1), tyy = threadIdx.x / kBlockFast;
    const int f0 = M.f - tx, m0 = M.m - tyy;
    const int fc = min(M.f, B.nf - 1), mcl = min(M.m, B.nm - 1);
    const int64_t own = (int64_t)mcl * B.sM + fc;
    int hy, hx;
    tile_halo(threadIdx.x, hy, hx);
    const bool halo = hx >= 0;
    const int64_t hoff = halo ? (int64_t)min(max(m0 - 1 + hy, 0), B.nm - 1) * B.sM + min(max(f0 - 1 + hx, 0), B.nf - 1)
                              : own;
    auto PV = [&](int64_t i) { return (double)(float)(first ? (double)z[i] : (double)z[i] + beta * (double)po[i]); };
    auto plane = [&](int sl) { return (int64_t)min(max(sl, 0), B.ns - 1) * B.sS; };
    double pq = 0.0;
    double Pm = PV(plane(M.s0 - 1) + own), Pc = PV(plane(M.s0) + own), Pp = PV(plane(M.s0 + 1) + own);
    double H = halo ? PV(plane(M.s0) + hoff) : 0.0;
    unsigned cc = T.c8[plane(M.s0) + own];
    for (int sl = M.s0; sl <= M.s1; sl++)
    {
        const double Pn = PV(plane(sl + 2) + own);
        const double Hn = halo ? PV(plane(sl + 1) + hoff) : 0.0;
        const unsigned cn = T.c8[plane(sl + 1) + own];
        Ps[tyy + 1][tx + 1] = Pc;
        if (halo) Ps[hy][hx] = H;
        __syncthreads();
        const int64_t idx = plane(sl) + own;
        if (M.in && (cc & 64u))
        {
            const double nb[6] = {Pp, Pm, Ps[tyy + 2][tx + 1], Ps[tyy][tx + 1], Ps[tyy + 1][tx + 2], Ps[tyy + 1][tx]};
            const double v = UNI ? apply64u(T, idx, Pc, nb, cc) : apply64v(B, T, ty, sl, M.m, M.f, idx, Pc, nb, cc);
            pn[idx] = (float)Pc;
            pq += Pc * v;
        }
        __syncthreads();
        Pm = Pc;
        Pc = Pp;
        Pp = Pn;
        H = Hn;
        cc = cn;
    }
    double a[1] = {pq};
    block_sum<1>(a, red, gridDim.x, M.lb);
}

// Part 2: q = A p (FP64, from the FP32 p of part 1), x += alpha p, r -= alpha q, r32 = (float) r, and <r, r>.
template <bool UNI>
__global__ __launch_bounds__(kBlockThreads) void k_cg_update64_t(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                                 const float *__restrict__ ring, int64_t stride,
                                                                 double *__restrict__ x, double *__restrict__ r,
                                                                 float *__restrict__ r32, const double *__restrict__ b0,
                                                                 const SolveState *st, Red red, int P)
{
#pragma clang fp contract(off)
    if (st->done) return;
    const double alpha = st->alpha;
    const int slot = st->iters % kXRing;
    const bool flush = slot == kXRing - 1;
    // a pass begun at x = 0: the residual is still b before its first update, x still 0 before its first ring
    const double *__restrict__ rin = st->zeroStart && st->iters == 0 ? b0 : r;
    const bool x0 = st->zeroStart && st->iters < kXRing;
    const float *__restrict__ p = ring + (size_t)slot * (size_t)stride;
    __shared__ double Ps[kBlockMid + 2][kBlockFast + 2];
    const March M = march_setup(B, P);
    const int tx = threadIdx.x & (kBlockFast - 1), tyy = threadIdx.x / kBlockFast;
    const int f0 = M.f - tx, m0 = M.m - tyy;
    const int fc = min(M.f, B.nf - 1), mcl = min(M.m, B.nm - 1);
    const int64_t own = (int64_t)mcl * B.sM + fc;
    int hy, hx;
    tile_halo(threadIdx.x, hy, hx);
    const bool halo = hx >= 0;
    const int64_t hoff = halo ? (int64_t)min(max(m0 - 1 + hy, 0), B.nm - 1) * B.sM + min(max(f0 - 1 + hx, 0), B.nf - 1)
                              : own;
    auto plane = [&](int sl) { return (int64_t)min(max(sl, 0), B.ns - 1) * B.sS; };
    double rr = 0.0;
    double Pm = p[plane(M.s0 - 1) + own], Pc = p[plane(M.s0) + own], Pp = p[plane(M.s0 + 1) + own];
    double H = halo ? (double)p[plane(M.s0) + hoff] : 0.0;
    unsigned cc = T.c8[plane(M.s0) + own];
    double rc = rin[plane(M.s0) + own];
    for (int sl = M.s0; sl <= M.s1; sl++)
    {
        const double Pn = p[plane(sl + 2) + own];
        const double Hn = halo ? (double)p[plane(sl + 1) + hoff] : 0.0;
        const unsigned cn = T.c8[plane(sl + 1) + own];
        const double rn = rin[plane(sl + 1) + own];
        Ps[tyy + 1][tx + 1] = Pc;
        if (halo) Ps[hy][hx] = H;
        __syncthreads();
        const int64_t idx = plane(sl) + own;
        if (M.in && (cc & 64u))
        {
            const double nb[6] = {Pp, Pm, Ps[tyy + 2][tx + 1], Ps[tyy][tx + 1], Ps[tyy + 1][tx + 2], Ps[tyy + 1][tx]};
            const double q = UNI ? apply64u(T, idx, Pc, nb, cc) : apply64v(B, T, ty, sl, M.m, M.f, idx, Pc, nb, cc);
            const double rv = rc - alpha * q;
            r[idx] = rv;
            r32[idx] = (float)rv;
            rr += rv * rv;
            if (flush)
            {
                // x += alpha_j p_j over the ring's iterations, oldest first (this iteration's p last)
                double xv = x0 ? 0.0 : x[idx];
#pragma unroll
                for (int j = 0; j < kXRing; j++)
                    xv = xv + st->alphas[j] * (j == kXRing - 1 ? Pc : (double)ring[(size_t)j * (size_t)stride + idx]);
                x[idx] = xv;
            }
        }
        __syncthreads();
        Pm = Pc;
        Pc = Pp;
        Pp = Pn;
        H = Hn;
        cc = cn;
        rc = rn;
    }
    double a[1] = {rr};
    block_sum<1>(a, red, gridDim.x, M.lb);
}

// ---- FP64 Krylov passes, whole-row waves (3D, unit coefficients) ----------------------------------------------
//
// k_cg_dir64_t and k_cg_update64_t on the block layout of k_rbgs_row: one wave per row (lane l owns the cell pairs
// (1 + 2p, 2 + 2p), p = l + 64k, so a wave reads and writes whole rows), TM interior rows plus a halo row above and
// below, a march over P planes with every load issued two planes ahead (x and r of the own cells one plane ahead),
// and the neighbouring rows' and pairs' p through LDS, double-buffered by plane parity (one barrier per plane).
// Loads outside the interior are clamped into the box; apply64u only adds coupled neighbours, so what they read
// there is never used. The per-cell arithmetic is the tile kernels' (the same FP32 rounding of p, apply64u in the
// same order), so p, x and r match theirs bit for bit; only the order in which <p, A p> and <r, r> are summed
// differs (it depends on the box size alone).
// apply64u for an operand that is +0 at every neighbour that is no coupled unknown (p of the FP64 Krylov passes:
// 0 at every cell but the unknowns, and the LDS columns beyond the row ends): the same products added in the same
// order without testing the code's neighbour bits. Bit for bit the same: a skipped term is v * (+0) = +-0, and
// adding +-0 leaves a sum that is never -0 unchanged.
__device__ __forceinline__ double apply64u_z(const Tables64 &T, int64_t idx, double xc, const double (&nb)[6],
                                             unsigned code)
{
#pragma clang fp contract(off)
    unsigned n;
    if (!(code & 128u))
        n = __popc(code & 63u);
    else
    {
        const unsigned c16 = T.c16[idx];
        n = 0;
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned st = (c16 >> (2 * q)) & 3u;
            n += st == 3u ? 2u : (st != 0u ? 1u : 0u);
        }
    }
    const double v = T.scale * 1.0;
    const double dv = T.scale * (-(double)n);
    double t = 0.0;
    t += v * nb[1];
    t += v * nb[3];
    t += v * nb[5];
    t += dv * xc;
    t += v * nb[4];
    t += v * nb[2];
    t += v * nb[0];
    return t;
}

template <int TM, int K>
struct RowMarch
{
    int lane, wave, m, mc, s0, s1;
    bool halo, rowIn;
    uint32_t off[K];
    bool okA[K], okB[K];
};

template <int TM, int K>
__device__ __forceinline__ RowMarch<TM, K> row_march(const Box &B, int P, int tilesM, int lb)
{
    RowMarch<TM, K> R;
    R.lane = threadIdx.x & 63;
    R.wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tm = lb % tilesM, chunk = lb / tilesM;
    R.m = tm * TM + R.wave;   // m0 - 1 + wave with m0 = 1 + tm * TM
    R.halo = R.wave == 0 || R.wave == TM + 1;
    R.rowIn = R.m >= 1 && R.m <= B.nm - 2;
    R.s0 = 1 + chunk * P;
    R.s1 = min(R.s0 + P - 1, B.ns - 2);
    const int mc = min(max(R.m, 0), B.nm - 1);
    R.mc = mc;
    const int fcMax = ((B.nf - 2) & 1) ? B.nf - 2 : B.nf - 3;
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        const int fa = 1 + 2 * (R.lane + 64 * k);
        R.okA[k] = R.rowIn && fa <= B.nf - 2;
        R.okB[k] = R.rowIn && fa + 1 <= B.nf - 2;
        R.off[k] = (uint32_t)mc * (uint32_t)B.sM + (uint32_t)min(fa, fcMax);
    }
    return R;
}

// Part 1 (see k_cg_dir64_t): p = (float)(z + beta p_old) (z on the first iteration) at the unknowns, and <p, A p>.
template <int TM, int K>
__device__ __forceinline__ void cg_dir64_row_body(const Box &B, const Tables64 &T, const float *__restrict__ z,
                                                  float *__restrict__ ring, int64_t stride, const SolveState *st,
                                                  const Red &red, int P, int tilesM, int lb, int nblk)
{
#pragma clang fp contract(off)
    if (st->done) return;
    const int first = st->first;
    const double beta = st->beta;
    const int slot = st->iters % kXRing;
    float *__restrict__ pn = ring + (size_t)slot * (size_t)stride;
    const float *__restrict__ po = ring + (size_t)((slot + kXRing - 1) % kXRing) * (size_t)stride;
    constexpr int kSlots = 64 * K + 2;
    __shared__ float ldsP[2][TM + 2][2][kSlots];   // new p of a plane by parity, row, element and pair + 1
    const RowMarch<TM, K> R = row_march<TM, K>(B, P, tilesM, lb);
    for (int i = threadIdx.x; i < 2 * (TM + 2) * 2 * kSlots; i += 64 * (TM + 2)) (&ldsP[0][0][0][0])[i] = 0.f;
    auto plane = [&](int pl) { return (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto pair = [](const float *row, uint32_t o) { return *reinterpret_cast<const float2 *>(row + o); };
    auto pv = [&](float zv, float pov) { return (float)(first ? (double)zv : (double)zv + beta * (double)pov); };
    // per-plane values in slots by (plane - (s0 - 1)) mod 4
    float2 Zv[4][K], Po[4][K], Pn[4][K];
    unsigned C[4][K];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            Zv[j][k] = Po[j][k] = Pn[j][k] = make_float2(0.f, 0.f);
            C[j][k] = 0u;
        }
    const int sp0 = R.s0 - 1;
    RowCodes rc;
    rc.load(T.rowc, B.ns, R.mc, sp0, R.s1 + 3);
#pragma unroll
    for (int d = 0; d <= 2; d++)
    {
        const size_t pl = plane(sp0 + d);
        row_codes_into<K>(C[d], rc, sp0 + d, T.c8 + pl, R.off, R.lane, B.nf);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            Zv[d][k] = pair(z + pl, R.off[k]);
            if (!first) Po[d][k] = pair(po + pl, R.off[k]);
        }
    }
#pragma unroll
    for (int d = 0; d <= 1; d++)
#pragma unroll
        for (int k = 0; k < K; k++) Pn[d][k] = make_float2(pv(Zv[d][k].x, Po[d][k].x), pv(Zv[d][k].y, Po[d][k].y));
    __syncthreads();
    const int wUp = min(R.wave + 1, TM + 1), wDn = max(R.wave - 1, 0);
    double pq = 0.0;
    auto step = [&](int sp, auto phase) {
        constexpr int U = decltype(phase)::value;
        constexpr int Sm1 = (U + 3) & 3, S0 = U, S1 = (U + 1) & 3, S2 = (U + 2) & 3;
        {
            const size_t pl = plane(sp + 2);
            row_codes_into<K>(C[S2], rc, sp + 2, T.c8 + pl, R.off, R.lane, B.nf);
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                Zv[S2][k] = pair(z + pl, R.off[k]);
                if (!first) Po[S2][k] = pair(po + pl, R.off[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
            Pn[S1][k] = make_float2(pv(Zv[S1][k].x, Po[S1][k].x), pv(Zv[S1][k].y, Po[S1][k].y));
        float(*buf)[2][kSlots] = ldsP[sp & 1];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            // (past a short row's end the clamped loads are no cells of the row: 0 there)
            buf[R.wave][0][1 + R.lane + 64 * k] = R.okA[k] || !R.rowIn ? Pn[S0][k].x : 0.f;
            buf[R.wave][1][1 + R.lane + 64 * k] = R.okB[k] || !R.rowIn ? Pn[S0][k].y : 0.f;
        }
        __syncthreads();
        if (R.halo) return;
        const size_t pl = plane(sp);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const int q = 1 + R.lane + 64 * k;
#pragma unroll
            for (int e = 0; e < 2; e++)
            {
                const unsigned code = (e ? R.okB[k] : R.okA[k]) ? ((C[S0][k] >> (8 * e)) & 255u) : 0u;
                if (!(code & 64u)) continue;
                const float pc = e ? Pn[S0][k].y : Pn[S0][k].x;
                const double nb[6] = {(double)(e ? Pn[S1][k].y : Pn[S1][k].x), (double)(e ? Pn[Sm1][k].y : Pn[Sm1][k].x),
                                      (double)buf[wUp][e][q], (double)buf[wDn][e][q],
                                      (double)(e ? buf[R.wave][0][q + 1] : Pn[S0][k].y),
                                      (double)(e ? Pn[S0][k].x : buf[R.wave][1][q - 1])};
                const double v = apply64u_z(T, (int64_t)pl + R.off[k] + e, (double)pc, nb, code);
                pq += (double)pc * v;
            }
            // p of both cells of the pair (a non-unknown's p is 0, as every iteration writes it)
            if (R.okB[k])
                *reinterpret_cast<float2 *>(pn + pl + R.off[k]) = Pn[S0][k];
            else if (R.okA[k])
                pn[pl + R.off[k]] = Pn[S0][k].x;
        }
    };
    for (int sp = R.s0; sp <= R.s1; sp += 4)
    {
        step(sp, std::integral_constant<int, 1>());
        if (sp + 1 > R.s1) break;
        step(sp + 1, std::integral_constant<int, 2>());
        if (sp + 2 > R.s1) break;
        step(sp + 2, std::integral_constant<int, 3>());
        if (sp + 3 > R.s1) break;
        step(sp + 3, std::integral_constant<int, 0>());
    }
    double a[1] = {pq};
    block_sum<1>(a, red, nblk, lb);
}

// Part 2 (see k_cg_update64_t): q = A p, r -= alpha q, r32 = (float) r at the unknowns, and <r, r>; on the
// iteration that fills the ring of directions, x += alpha_j p_j for the ring's iterations in order.
template <int TM, int K>
__device__ __forceinline__ void cg_update64_row_body(const Box &B, const Tables64 &T, const float *__restrict__ ring,
                                                     int64_t stride, double *__restrict__ x,
                                                     double *__restrict__ r, float *__restrict__ r32,
                                                     const double *__restrict__ b0, const SolveState *st,
                                                     const Red &red, int P, int tilesM, int lb, int nblk)
{
#pragma clang fp contract(off)
    if (st->done) return;
    const double alpha = st->alpha;
    const int slot = st->iters % kXRing;
    const bool flush = slot == kXRing - 1;
    // a pass begun at x = 0: the residual is still b before its first update, x still 0 before its first ring
    const double *__restrict__ rin = st->zeroStart && st->iters == 0 ? b0 : r;
    const bool x0 = st->zeroStart && st->iters < kXRing;
    const float *__restrict__ p = ring + (size_t)slot * (size_t)stride;
    constexpr int kSlots = 64 * K + 2;
    __shared__ float ldsP[2][TM + 2][2][kSlots];
    const RowMarch<TM, K> R = row_march<TM, K>(B, P, tilesM, lb);
    for (int i = threadIdx.x; i < 2 * (TM + 2) * 2 * kSlots; i += 64 * (TM + 2)) (&ldsP[0][0][0][0])[i] = 0.f;
    auto plane = [&](int pl) { return (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto pair = [](const float *row, uint32_t o) { return *reinterpret_cast<const float2 *>(row + o); };
    auto pair64 = [](const double *row, uint32_t o) { return *reinterpret_cast<const double2 *>(row + o); };
    float2 Pv[4][K];
    unsigned C[4][K];
    double2 Rv[4][K];   // own r (interior waves), loaded two planes ahead
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int j = 0; j < 4; j++)
        {
            Pv[j][k] = make_float2(0.f, 0.f);
            C[j][k] = 0u;
            Rv[j][k] = make_double2(0.0, 0.0);
        }
    const int sp0 = R.s0 - 1;
    RowCodes rc;
    rc.load(T.rowc, B.ns, R.mc, sp0, R.s1 + 3);
#pragma unroll
    for (int d = 0; d <= 2; d++)
    {
        const size_t pl = plane(sp0 + d);
        row_codes_into<K>(C[d], rc, sp0 + d, T.c8 + pl, R.off, R.lane, B.nf);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            Pv[d][k] = pair(p + pl, R.off[k]);
        }
    }
    if (!R.halo)
    {
#pragma unroll
        for (int d = 1; d <= 2; d++)
        {
            const size_t pl = plane(sp0 + d);
#pragma unroll
            for (int k = 0; k < K; k++) Rv[d][k] = pair64(rin + pl, R.off[k]);
        }
    }
    __syncthreads();
    const int wUp = min(R.wave + 1, TM + 1), wDn = max(R.wave - 1, 0);
    double rr = 0.0;
    auto step = [&](int sp, auto phase) {
        constexpr int U = decltype(phase)::value;
        constexpr int Sm1 = (U + 3) & 3, S0 = U, S1 = (U + 1) & 3, S2 = (U + 2) & 3;
        {
            const size_t pl = plane(sp + 2);
            row_codes_into<K>(C[S2], rc, sp + 2, T.c8 + pl, R.off, R.lane, B.nf);
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                Pv[S2][k] = pair(p + pl, R.off[k]);
            }
        }
        if (!R.halo && sp + 2 <= R.s1)
        {
            const size_t pl = plane(sp + 2);
#pragma unroll
            for (int k = 0; k < K; k++) Rv[S2][k] = pair64(rin + pl, R.off[k]);
        }
        float(*buf)[2][kSlots] = ldsP[sp & 1];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            buf[R.wave][0][1 + R.lane + 64 * k] = R.okA[k] || !R.rowIn ? Pv[S0][k].x : 0.f;
            buf[R.wave][1][1 + R.lane + 64 * k] = R.okB[k] || !R.rowIn ? Pv[S0][k].y : 0.f;
        }
        __syncthreads();
        if (R.halo) return;
        const size_t pl = plane(sp);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const int q = 1 + R.lane + 64 * k;
            double ro[2] = {Rv[S0][k].x, Rv[S0][k].y};
            bool any = false;
#pragma unroll
            for (int e = 0; e < 2; e++)
            {
                const unsigned code = (e ? R.okB[k] : R.okA[k]) ? ((C[S0][k] >> (8 * e)) & 255u) : 0u;
                if (!(code & 64u)) continue;
                any = true;
                const double pc = (double)(e ? Pv[S0][k].y : Pv[S0][k].x);
                const double nb[6] = {(double)(e ? Pv[S1][k].y : Pv[S1][k].x), (double)(e ? Pv[Sm1][k].y : Pv[Sm1][k].x),
                                      (double)buf[wUp][e][q], (double)buf[wDn][e][q],
                                      (double)(e ? buf[R.wave][0][q + 1] : Pv[S0][k].y),
                                      (double)(e ? Pv[S0][k].x : buf[R.wave][1][q - 1])};
                const double qv = apply64u_z(T, (int64_t)pl + R.off[k] + e, pc, nb, code);
                ro[e] = ro[e] - alpha * qv;
                rr += ro[e] * ro[e];
            }
            if (!any) continue;
            // a pair with one unknown writes the other element back unchanged (r and r32 are 0 at non-unknowns)
            const int64_t idx = (int64_t)pl + R.off[k];
            if (R.okB[k])
            {
                *reinterpret_cast<double2 *>(r + idx) = make_double2(ro[0], ro[1]);
                *reinterpret_cast<float2 *>(r32 + idx) = make_float2((float)ro[0], (float)ro[1]);
            }
            else
            {
                r[idx] = ro[0];
                r32[idx] = (float)ro[0];
            }
            if (flush)
            {
                // x += alpha_j p_j over the ring's iterations, oldest first (this iteration's p last)
                double xo[2] = {0.0, 0.0};
                if (x0)
                    ;   // still the zero the pass began at
                else if (R.okB[k])
                {
                    const double2 xv = *reinterpret_cast<const double2 *>(x + idx);
                    xo[0] = xv.x;
                    xo[1] = xv.y;
                }
                else
                    xo[0] = x[idx];
#pragma unroll
                for (int j = 0; j < kXRing; j++)
                {
                    const double aj = st->alphas[j];
                    const float *rj = ring + (size_t)j * (size_t)stride + idx;
                    const float2 pj = j == kXRing - 1 ? Pv[S0][k]
                                                      : (R.okB[k] ? *reinterpret_cast<const float2 *>(rj)
                                                                  : make_float2(rj[0], 0.f));
#pragma unroll
                    for (int e = 0; e < 2; e++)
                    {
                        const unsigned code = (e ? R.okB[k] : R.okA[k]) ? ((C[S0][k] >> (8 * e)) & 255u) : 0u;
                        if (code & 64u) xo[e] = xo[e] + aj * (double)(e ? pj.y : pj.x);
                    }
                }
                if (R.okB[k])
                    *reinterpret_cast<double2 *>(x + idx) = make_double2(xo[0], xo[1]);
                else
                    x[idx] = xo[0];
            }
        }
    };
    for (int sp = R.s0; sp <= R.s1; sp += 4)
    {
        step(sp, std::integral_constant<int, 1>());
        if (sp + 1 > R.s1) break;
        step(sp + 1, std::integral_constant<int, 2>());
        if (sp + 2 > R.s1) break;
        step(sp + 2, std::integral_constant<int, 3>());
        if (sp + 3 > R.s1) break;
        step(sp + 3, std::integral_constant<int, 0>());
    }
    double a[1] = {rr};
    block_sum<1>(a, red, nblk, lb);
}

// Residual of every fine cell summed over each 2 x 2 x 2 family into the coarse right-hand side (k_residual_restrict_m)
// on the layout of k_rbgs_row: one wave per row (TM even, so a block's interior rows start on an odd row and hold
// whole row pairs; P even, so chunks start on an odd plane), x through LDS double-buffered by plane, every load two
// planes ahead. The sums are k_residual_restrict_m's in its order: the pair (r_a + r_b), then the two rows (the odd
// row first), then the two planes added to 0 -- bit for bit the same coarse right-hand side. The fast cells' six
// neighbours are added without testing the code's bits (x is +0 at every cell that is no unknown, and at the LDS
// columns beyond the row ends); the first row of each family reads its partner row's pair sums one plane step later.
template <int TM, int K>
__device__ __forceinline__ void restrict_row_body(const Level32 &Lf, const float *__restrict__ x,
                                                  const float *__restrict__ b, const Level32 &Lc,
                                                  float *__restrict__ bc, const SolveState *st, int P, int tilesM,
                                                  int lb)
{
    static_assert((TM & 1) == 0, "rows are summed in pairs");
    if (st->done) return;
    const Box &B = Lf.B;
    constexpr int kSlots = 64 * K + 2;
    __shared__ float ldsX[2][TM + 2][2][kSlots];   // x of a plane by parity, row, element and pair + 1
    __shared__ float ldsR[2][TM][64 * K];          // the rows' pair sums of the residual by plane parity
    const RowMarch<TM, K> R = row_march<TM, K>(B, P, tilesM, lb);
    for (int i = threadIdx.x; i < 2 * (TM + 2) * 2 * kSlots; i += 64 * (TM + 2)) (&ldsX[0][0][0][0])[i] = 0.f;
    auto plane = [&](int pl) { return (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto pair = [](const float *row, uint32_t o) { return *reinterpret_cast<const float2 *>(row + o); };
    auto pairC = [](const uint8_t *row, uint32_t o) { return (unsigned)*reinterpret_cast<const uint16_t *>(row + o); };
    unsigned cm[K];   // the pair's codes inside the row
#pragma unroll
    for (int k = 0; k < K; k++) cm[k] = (R.okA[k] ? 0xFFu : 0u) | (R.okB[k] ? 0xFF00u : 0u);
    float2 X[4][K], Bv[4][K];
    unsigned C[4][K];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            X[j][k] = Bv[j][k] = make_float2(0.f, 0.f);
            C[j][k] = 0u;
        }
    const int sp0 = R.s0 - 1;
#pragma unroll
    for (int d = 0; d <= 2; d++)
    {
        const size_t pl = plane(sp0 + d);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            X[d][k] = pair(x + pl, R.off[k]);
            if (d >= 1 && !R.halo)
            {
                Bv[d][k] = pair(b + pl, R.off[k]);
                C[d][k] = pairC(Lf.c8 + pl, R.off[k]);
            }
        }
    }
    __syncthreads();
    const int wUp = min(R.wave + 1, TM + 1), wDn = max(R.wave - 1, 0);
    const int i = R.wave - 1;   // interior row index of the block
    const bool famRow = !R.halo && !(i & 1);   // the odd row of a row pair: writes the family sums
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.f;
    // the family sums of plane q from the rows' pair sums in ldsR[q & 1]
    auto family = [&](int q) {
        if (!famRow || q < R.s0) return;
        const size_t cpl = (size_t)((q + 1) >> 1) * (size_t)Lc.B.sS + (size_t)((R.m + 1) >> 1) * (size_t)Lc.B.sM;
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const int pq = R.lane + 64 * k;
            acc[k] += ldsR[q & 1][i][pq] + ldsR[q & 1][i + 1][pq];
            if (!(q & 1) || q == R.s1)
            {
                if (R.okA[k]) bc[cpl + pq + 1] = acc[k];
                acc[k] = 0.f;
            }
        }
    };
    auto step = [&](int sp, auto phase) {
        constexpr int U = decltype(phase)::value;
        constexpr int Sm1 = (U + 3) & 3, S0 = U, S1 = (U + 1) & 3, S2 = (U + 2) & 3;
        {
            const size_t pl = plane(sp + 2);
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                X[S2][k] = pair(x + pl, R.off[k]);
                if (!R.halo)
                {
                    Bv[S2][k] = pair(b + pl, R.off[k]);
                    C[S2][k] = pairC(Lf.c8 + pl, R.off[k]);
                }
            }
        }
        float(*buf)[2][kSlots] = ldsX[sp & 1];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            buf[R.wave][0][1 + R.lane + 64 * k] = R.okA[k] || !R.rowIn ? X[S0][k].x : 0.f;
            buf[R.wave][1][1 + R.lane + 64 * k] = R.okB[k] || !R.rowIn ? X[S0][k].y : 0.f;
        }
        __syncthreads();
        family(sp - 1);
        if (R.halo) return;
        const size_t pl = plane(sp);
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const int q = 1 + R.lane + 64 * k;
            float r[2];
#pragma unroll
            for (int e = 0; e < 2; e++)
            {
                const unsigned code = ((C[S0][k] & cm[k]) >> (8 * e)) & 255u;
                const float xc = e ? X[S0][k].y : X[S0][k].x;
                const float nb[6] = {e ? X[S1][k].y : X[S1][k].x, e ? X[Sm1][k].y : X[Sm1][k].x, buf[wUp][e][q],
                                     buf[wDn][e][q], e ? buf[R.wave][0][q + 1] : X[S0][k].y,
                                     e ? X[S0][k].x : buf[R.wave][1][q - 1]};
                const float bv = e ? Bv[S0][k].y : Bv[S0][k].x;
                float v = 0.f;
                if (code & 64u)
                {
                    if (!(code & 128u))
                    {
#pragma clang fp contract(off)
                        // k_residual_restrict_m's rounding: the neighbour sum less n x in one fused step, A x
                        // rounded on its own, then b - A x
                        const float s6 = nb[0] + nb[1] + nb[2] + nb[3] + nb[4] + nb[5];
                        const float ax = Lf.cu * __builtin_fmaf(-(float)__popc(code & 63u), xc, s6);
                        v = bv - ax;
                    }
                    else
                        v = bv - apply32v(Lf, sp, R.m, 2 * q - 1 + e, (int64_t)pl + R.off[k] + e, code, xc, nb);
                }
                r[e] = v;
            }
            ldsR[sp & 1][i][R.lane + 64 * k] = r[0] + r[1];
        }
    };
    for (int sp = R.s0; sp <= R.s1; sp += 4)
    {
        step(sp, std::integral_constant<int, 1>());
        if (sp + 1 > R.s1) break;
        step(sp + 1, std::integral_constant<int, 2>());
        if (sp + 2 > R.s1) break;
        step(sp + 2, std::integral_constant<int, 3>());
        if (sp + 3 > R.s1) break;
        step(sp + 3, std::integral_constant<int, 0>());
    }
    __syncthreads();
    family(R.s1);
}

// The last pre-smoothing sweep and the restriction of its residual in one march (3D uniform levels, rows of at most
// 256 cells): k_rbgs_row's plain sweep on a block of TM rows with two halo rows each side, and two planes behind it
// k_restrict_row's residual and family sums, on the sweep's results while they are still in registers and LDS. The
// outer halo rows compute the first colour only (as the sweep's halo rows do), the inner ones both colours without
// storing them, so the residual of the block's first and last rows has its row neighbours; each chunk also sweeps the
// plane before and after it (not stored) for the residual of its first and last plane. Every value is computed by
// the same operations as in the two kernels: x and the coarse right-hand side are bit for bit theirs, for one pass
// over x, b and codes instead of two (TM and P even, so families lie within a block and a chunk).
template <int TM, int K>
__device__ __forceinline__ void rbgs_restrict_row_body(const Level32 &L, const float *__restrict__ xin,
                                                       float *__restrict__ xout, const float *__restrict__ b, int c1,
                                                       const Level32 &Lc, float *__restrict__ bc, int P, int tilesM,
                                                       const SolveState *st, int lb)
{
    static_assert((TM & 1) == 0, "rows are summed in pairs");
    if (st->done) return;
    const Box &B = L.B;
    constexpr int W = TM + 4;   // waves: rows m0 - 2 .. m0 + TM + 1
    constexpr int kCols = 128 * K + 2;
    constexpr int kSlots = 64 * K + 2;
    __shared__ float red[2][W][kCols];        // first-colour values by column f
    __shared__ float old[2][W][2][kSlots];    // old x of a plane by parity, row, element, pair + 1
    __shared__ float fin[4][W][2][kSlots];    // the swept x of a plane by slot (plane - sp0) & 3, row, element, pair + 1
    __shared__ float ldsR[2][TM][64 * K];     // the interior rows' pair sums of the residual by plane parity
    __shared__ float2 tab[128];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tm = lb % tilesM, chunk = lb / tilesM;
    const int m0 = 1 + tm * TM;
    const int m = m0 - 2 + wave;
    const bool outerRow = wave == 0 || wave == W - 1;
    const bool interior = wave >= 2 && wave <= W - 3;
    const int i = wave - 2;                       // interior row index of the block
    const bool famRow = interior && !(i & 1);     // the odd row of a row pair: writes the family sums
    const bool rowIn = m >= 1 && m <= B.nm - 2;
    const bool ragged = B.nf - 2 < 128 * K;
    const int mc = min(max(m, 0), B.nm - 1);
    const int mo = min(max(wave == 0 ? m - 1 : m + 1, 0), B.nm - 1);
    const int s0 = 1 + chunk * P, s1 = min(s0 + P - 1, B.ns - 2);
    const uint32_t sM = (uint32_t)B.sM;
    const int fcMax = ((B.nf - 2) & 1) ? B.nf - 2 : B.nf - 3;
    uint32_t off[K], offO[K];
    bool okA[K], okB[K];
    unsigned cm[K];
    int fa[K];
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        fa[k] = 1 + 2 * (lane + 64 * k);
        okA[k] = rowIn && fa[k] <= B.nf - 2;
        okB[k] = rowIn && fa[k] + 1 <= B.nf - 2;
        cm[k] = (okA[k] ? 0xFFu : 0u) | (okB[k] ? 0xFF00u : 0u);
        const int fc = min(fa[k], fcMax);
        off[k] = (uint32_t)mc * sM + (uint32_t)fc;
        offO[k] = (uint32_t)mo * sM + (uint32_t)fc;
    }
    for (int q = threadIdx.x; q < 2 * W * kCols; q += 64 * W) (&red[0][0][0])[q] = 0.f;
    for (int q = threadIdx.x; q < 2 * W * 2 * kSlots; q += 64 * W) (&old[0][0][0][0])[q] = 0.f;
    for (int q = threadIdx.x; q < 4 * W * 2 * kSlots; q += 64 * W) (&fin[0][0][0][0])[q] = 0.f;
    if (threadIdx.x < 128)
    {
        const unsigned c = threadIdx.x;
        tab[c] = make_float2((c & 64u) ? inv_count(c) : 0.f, __uint_as_float((c & 64u) ? ~0u : 0u));
    }
    auto plane = [&](const float *base, int pl) { return base + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto planeC = [&](int pl) { return L.c8 + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto plane_off = [&](int pl) { return (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto pair = [](const float *row, uint32_t o) { return *reinterpret_cast<const float2 *>(row + o); };
    auto gs = [&](float acc, float bv, unsigned code) {
        const float2 t = tab[code & 127u];
        return __uint_as_float(__float_as_uint((acc - bv * L.icu) * t.x) & __float_as_uint(t.y));
    };
    auto gsDirichlet = [&](float acc, float bv, int pl, uint32_t o) {
        const unsigned c16 = L.c16[(size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS + o];
        float d = 0.f;
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned stt = (c16 >> (2 * q)) & 3u;
            d += stt == 3u ? 2.f : (stt != 0u ? 1.f : 0.f);
        }
        return d != 0.f ? (acc - bv * L.icu) / d : 0.f;
    };
    // slots by plane mod 4 relative to sp0 (as k_rbgs_row): x, b and codes two planes ahead, the outer rows one
    float2 X[4][K], O[2][K], Bv[4][K];
    unsigned C[4][K];
    float R[4][K];    // first-colour value of the pair
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            X[j][k] = Bv[j][k] = make_float2(0.f, 0.f);
            O[j & 1][k] = make_float2(0.f, 0.f);
            R[j][k] = 0.f;
            C[j][k] = 0u;
        }
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.f;
    const int sp0 = s0 - 2;
    RowCodes rc;
    rc.load(L, mc, sp0, s1 + 4);
#pragma unroll
    for (int d = -1; d <= 1; d++)
    {
        const float *px = plane(xin, sp0 + d);
#pragma unroll
        for (int k = 0; k < K; k++) X[(d + 4) & 3][k] = pair(px, off[k]);
    }
    if (outerRow)
    {
        const float *px = plane(xin, sp0);
#pragma unroll
        for (int k = 0; k < K; k++) O[0][k] = pair(px, offO[k]);
    }
#pragma unroll
    for (int d = 0; d <= 1; d++)
    {
        const float *pb = plane(b, sp0 + d);
#pragma unroll
        for (int k = 0; k < K; k++) Bv[d][k] = pair(pb, off[k]);
        row_codes_into<K>(C[d], rc, sp0 + d, planeC(sp0 + d), off, lane, B.nf);
    }
    __syncthreads();
    const int wUp = min(wave + 1, W - 1), wDn = max(wave - 1, 0);
    // the family sums of plane q from the rows' pair sums in ldsR[q & 1]
    auto family = [&](int q) {
        if (!famRow || q < s0 || q > s1) return;
        const size_t cpl = (size_t)((q + 1) >> 1) * (size_t)Lc.B.sS + (size_t)((m + 1) >> 1) * (size_t)Lc.B.sM;
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const int pq = lane + 64 * k;
            acc[k] += ldsR[q & 1][i][pq] + ldsR[q & 1][i + 1][pq];
            if (!(q & 1) || q == s1)
            {
                if (okA[k]) bc[cpl + pq + 1] = acc[k];
                acc[k] = 0.f;
            }
        }
    };
    // One plane step at sp: first colour of plane sp, second colour of plane q = sp - 1 (stored on the interior rows
    // inside the chunk), residual of plane t = sp - 2 (interior rows), family sums of plane t - 1.
    auto step = [&](int sp, auto phase, auto colour) {
        constexpr int U = decltype(phase)::value;
        constexpr int E = decltype(colour)::value;
        constexpr int Sm2 = (U + 2) & 3, Sm1 = (U + 3) & 3, S0 = U & 3, S1 = (U + 1) & 3, S2 = (U + 2) & 3;
        constexpr int O0 = U & 1, O1 = (U + 1) & 1;
        // x two planes ahead now; b and codes of plane sp + 2 at the end of the step, once the residual of plane
        // t = sp - 2 has used that slot
        {
            const float *px = plane(xin, sp + 2);
#pragma unroll
            for (int k = 0; k < K; k++) X[S2][k] = pair(px, off[k]);
            if (outerRow)
            {
                const float *po = plane(xin, sp + 1);
#pragma unroll
                for (int k = 0; k < K; k++) O[O1][k] = pair(po, offO[k]);
            }
        }
        float(*ob)[2][kSlots] = old[sp & 1];
        if (ragged)
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                ob[wave][0][1 + lane + 64 * k] = okA[k] ? X[S0][k].x : 0.f;
                ob[wave][1][1 + lane + 64 * k] = okB[k] ? X[S0][k].y : 0.f;
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                ob[wave][0][1 + lane + 64 * k] = X[S0][k].x;
                ob[wave][1][1 + lane + 64 * k] = X[S0][k].y;
            }
        }
        // the one barrier of the step: old[sp & 1], red[(sp - 1) & 1], fin of plane sp - 2 and ldsR[(sp - 3) & 1],
        // written in this step or the one before, are complete; their readers of a step before are done (fin holds
        // four planes: a row's own residual reads its planes t - 1 .. t + 1, the last written by itself this step)
        __syncthreads();
        family(sp - 3);
        float rnew[K];
#pragma unroll
        for (int k = 0; k < K; k++)
        {
            const unsigned code = ((C[S0][k] & cm[k]) >> (8 * E)) & 255u;
            const float bv = E ? Bv[S0][k].y : Bv[S0][k].x;
            const int p = 1 + lane + 64 * k;
            const float own = E ? X[S0][k].x : X[S0][k].y;
            const float nbr = E ? ob[wave][0][p + 1] : ob[wave][1][p - 1];
            float up = ob[wUp][E][p], dn = ob[wDn][E][p];
            const float outer = E ? O[O0][k].y : O[O0][k].x;
            if (wave == W - 1) up = outer;
            if (wave == 0) dn = outer;
            // (no leading 0 + : an iterate is never -0, so the sum is the same)
            float a = E ? X[S1][k].y : X[S1][k].x;
            a += E ? X[Sm1][k].y : X[Sm1][k].x;
            a += up;
            a += dn;
            a += E ? nbr : own;
            a += E ? own : nbr;
            float v = gs(a, bv, code);
            if (code & 128u) v = gsDirichlet(a, bv, sp, off[k] + (uint32_t)E);
            rnew[k] = v;
            R[S0][k] = v;
        }
#pragma unroll
        for (int k = 0; k < K; k++) red[sp & 1][wave][fa[k] + E] = rnew[k];
        const int q = sp - 1;
        if (!outerRow && q >= s0 - 1 && q <= s1 + 1)
        {
            // second colour at plane q (there the second-colour element is E)
            float *po = xout + (size_t)q * (size_t)B.sS;
            const float(*rq)[kCols] = red[q & 1];
            const bool store = interior && q >= s0 && q <= s1;
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                const unsigned cc = C[Sm1][k] & cm[k];
                const unsigned code = (cc >> (8 * E)) & 255u;
                const int f = fa[k] + E;
                const float bv = E ? Bv[Sm1][k].y : Bv[Sm1][k].x;
                float a = R[S0][k];
                a += R[Sm2][k];
                a += rq[wave + 1][f];
                a += rq[wave - 1][f];
                a += E ? rq[wave][f + 1] : R[Sm1][k];
                a += E ? R[Sm1][k] : rq[wave][f - 1];
                float out = gs(a, bv, code);
                if (code & 128u) out = gsDirichlet(a, bv, q, off[k] + (uint32_t)E);
                const float va = E ? R[Sm1][k] : out, vb = E ? out : R[Sm1][k];
                fin[Sm1][wave][0][1 + lane + 64 * k] = va;
                fin[Sm1][wave][1][1 + lane + 64 * k] = vb;
                if (store)
                {
                    if (okB[k])
                        *reinterpret_cast<float2 *>(po + off[k]) = make_float2(va, vb);
                    else if (okA[k])
                        po[off[k]] = va;
                }
            }
        }
        const int t = sp - 2;
        if (interior && t >= s0 && t <= s1)
        {
            // residual of plane t (k_restrict_row's operations) and the row's pair sums
            const float(*ft)[2][kSlots] = fin[Sm2];
            const float(*fp)[2][kSlots] = fin[Sm1];   // plane t + 1
            const float(*fm)[2][kSlots] = fin[S1];    // plane t - 1
#pragma unroll
            for (int k = 0; k < K; k++)
            {
                const int p = 1 + lane + 64 * k;
                float r[2];
#pragma unroll
                for (int e = 0; e < 2; e++)
                {
                    const unsigned code = ((C[S2][k] & cm[k]) >> (8 * e)) & 255u;
                    const float xc = ft[wave][e][p];
                    const float nb[6] = {fp[wave][e][p], fm[wave][e][p], ft[wave + 1][e][p], ft[wave - 1][e][p],
                                         e ? ft[wave][0][p + 1] : ft[wave][1][p], e ? ft[wave][0][p] : ft[wave][1][p - 1]};
                    const float bv = e ? Bv[S2][k].y : Bv[S2][k].x;
                    float v = 0.f;
                    if (code & 64u)
                    {
                        if (!(code & 128u))
                        {
#pragma clang fp contract(off)
                            const float s6 = nb[0] + nb[1] + nb[2] + nb[3] + nb[4] + nb[5];
                            const float ax = L.cu * __builtin_fmaf(-(float)__popc(code & 63u), xc, s6);
                            v = bv - ax;
                        }
                        else
                            v = bv - apply32v(L, t, m, 2 * p - 1 + e, (int64_t)plane_off(t) + off[k] + e, code, xc,
                                              nb);
                    }
                    r[e] = v;
                }
                ldsR[t & 1][i][lane + 64 * k] = r[0] + r[1];
            }
        }
        {
            const float *pb = plane(b, sp + 2);
#pragma unroll
            for (int k = 0; k < K; k++) Bv[S2][k] = pair(pb, off[k]);
            row_codes_into<K>(C[S2], rc, sp + 2, planeC(sp + 2), off, lane, B.nf);
        }
    };
    // the first-colour element at plane sp is (c1 - sp - m - 1) & 1
    auto march = [&](auto colour0) {
        constexpr int E0 = decltype(colour0)::value;
        for (int sp = sp0; sp <= s1 + 2; sp += 4)
        {
            step(sp, std::integral_constant<int, 0>(), std::integral_constant<int, E0>());
            if (sp + 1 > s1 + 2) break;
            step(sp + 1, std::integral_constant<int, 1>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 2 > s1 + 2) break;
            step(sp + 2, std::integral_constant<int, 2>(), std::integral_constant<int, E0>());
            if (sp + 3 > s1 + 2) break;
            step(sp + 3, std::integral_constant<int, 3>(), std::integral_constant<int, E0 ^ 1>());
        }
    };
    if ((c1 - sp0 - m - 1) & 1)
        march(std::integral_constant<int, 1>());
    else
        march(std::integral_constant<int, 0>());
    __syncthreads();
    family(s1);
}

// ---- the whole pre-smoothing and the restriction in one march (whole-row waves, 3D uniform levels) -------------
//
// Both pre-smoothing sweeps of a level with rows of at most 128 cells, from x = 0, and the restriction of the residual
// as one launch: a plane step sp computes the zero-guess sweep's first colour at plane sp (from b alone: with x = 0
// its neighbours are zero) and its second colour at sp - 1, the plain sweep's first colour at sp - 2 and its second
// colour at sp - 3 (stored), the residual of plane sp - 4 and the family sums of plane sp - 5. Each stage computes
// what the kernel it replaces computes, from the same values in the same order (k_rbgs_row's zero-guess and plain
// sweeps, k_rbgs_restrict_row's residual and family sums), so x and the coarse right-hand side are bit for bit
// theirs. A block is TM output rows and four halo rows each side: the zero-guess first colour runs on every wave, each
// later stage on one row fewer each side. No x is read: b and codes once, x and the coarse b written once. Values pass between stages through LDS, double-buffered by plane parity, with the one barrier of the step;
// b and codes stay in registers from their load (two planes ahead) to the residual (slots by plane mod 8).
template <int TM>
__device__ __forceinline__ void rbgs_pre_restrict_row_body(const Level32 &L, float *__restrict__ xout,
                                                           const float *__restrict__ b, int c1, const Level32 &Lc,
                                                           float *__restrict__ bc, int P, int tilesM,
                                                           const SolveState *st, int lb)
{
    static_assert((TM & 1) == 0, "rows are summed in pairs");
    if (st->done) return;
    constexpr int W = TM + 8;   // waves: rows m0 - 4 .. m0 + TM + 3
    constexpr int kCols = 130;
    constexpr int kSlots = 66;
    const Box &B = L.B;
    __shared__ float z1s[2][W][kCols];       // zero-guess sweep, first colour, by column
    __shared__ float z2s[2][W][kCols];       // zero-guess sweep, second colour
    __shared__ float s1s[2][W][kCols];       // plain sweep, first colour
    __shared__ float fin[4][W][2][kSlots];   // the swept x of a plane by slot (plane - sp0) & 3, row, element, pair + 1
    __shared__ float ldsR[2][TM][64];        // the output rows' pair sums of the residual by plane parity
    __shared__ float2 tab[128];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int tm = lb % tilesM, chunk = lb / tilesM;
    const int m0 = 1 + tm * TM;
    const int m = m0 - 4 + wave;
    const bool doZ2 = wave >= 1 && wave <= W - 2, doS1 = wave >= 2 && wave <= W - 3;
    const bool doS2 = wave >= 3 && wave <= W - 4, interior = wave >= 4 && wave <= W - 5;
    const int i = wave - 4;                       // output row index of the block
    const bool famRow = interior && !(i & 1);     // the odd row of a row pair: writes the family sums
    const bool rowIn = m >= 1 && m <= B.nm - 2;
    const int mc = min(max(m, 0), B.nm - 1);
    const int s0 = 1 + chunk * P, s1 = min(s0 + P - 1, B.ns - 2);
    const uint32_t sM = (uint32_t)B.sM;
    const int fcMax = ((B.nf - 2) & 1) ? B.nf - 2 : B.nf - 3;
    const int fa = 1 + 2 * lane;
    const bool colA = fa <= B.nf - 2, colB = fa + 1 <= B.nf - 2;
    const bool okA = rowIn && colA, okB = rowIn && colB;
    const unsigned cm = (okA ? 0xFFu : 0u) | (okB ? 0xFF00u : 0u);
    const int fc = min(fa, fcMax);
    const uint32_t off = (uint32_t)mc * sM + (uint32_t)fc;
    for (int q = threadIdx.x; q < 2 * W * kCols; q += 64 * W)
    {
        (&z1s[0][0][0])[q] = 0.f;
        (&z2s[0][0][0])[q] = 0.f;
        (&s1s[0][0][0])[q] = 0.f;
    }
    for (int q = threadIdx.x; q < 4 * W * 2 * kSlots; q += 64 * W) (&fin[0][0][0][0])[q] = 0.f;
    if (threadIdx.x < 128)
    {
        const unsigned c = threadIdx.x;
        tab[c] = make_float2((c & 64u) ? inv_count(c) : 0.f, __uint_as_float((c & 64u) ? ~0u : 0u));
    }
    auto plane = [&](const float *base, int pl) { return base + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto planeC = [&](int pl) { return L.c8 + (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto plane_off = [&](int pl) { return (size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS; };
    auto pair = [](const float *row, uint32_t o) { return *reinterpret_cast<const float2 *>(row + o); };
    auto pairC = [](const uint8_t *row, uint32_t o) { return (unsigned)*reinterpret_cast<const uint16_t *>(row + o); };
    // acc - b / cu as one fma, the single sweep's contracted form (each plane's b feeds several stages here, and a
    // product b / cu shared by them would otherwise be rounded before the subtraction)
    auto gs = [&](float acc, float bv, unsigned code) {
        const float2 t = tab[code & 127u];
        return __uint_as_float(__float_as_uint(__builtin_fmaf(-bv, L.icu, acc) * t.x) & __float_as_uint(t.y));
    };
    auto gsDirichlet = [&](float acc, float bv, int pl, uint32_t o) {
        const unsigned c16 = L.c16[(size_t)min(max(pl, 0), B.ns - 1) * (size_t)B.sS + o];
        float d = 0.f;
#pragma unroll
        for (int q = 0; q < 6; q++)
        {
            const unsigned stt = (c16 >> (2 * q)) & 3u;
            d += stt == 3u ? 2.f : (stt != 0u ? 1.f : 0.f);
        }
        return d != 0.f ? __builtin_fmaf(-bv, L.icu, acc) / d : 0.f;
    };
    // the update of element E of a pair from its six neighbours (s+1, s-1, m+1, m-1, f+1, f-1), as the single sweep
    auto update = [&](const float (&nb)[6], float bv, unsigned code, int pl, uint32_t o) {
        float acc = nb[0];
        acc += nb[1];
        acc += nb[2];
        acc += nb[3];
        acc += nb[4];
        acc += nb[5];
        float v = gs(acc, bv, code);
        if (code & 128u) v = gsDirichlet(acc, bv, pl, o);
        return v;
    };
    // the zero-guess first colour (no neighbours: an iterate is never -0, so 0 stands for their sum)
    auto first0 = [&](float bv, unsigned code, int pl, uint32_t o) {
        float v = gs(0.f, bv, code);
        if (code & 128u) v = gsDirichlet(0.f, bv, pl, o);
        return v;
    };
    // slots: b and codes by plane mod 8 (sp + 2 .. sp - 4), the stages' own values by plane mod 4 (relative to sp0)
    float2 Bv[8];
    unsigned C[8];
    float R1[4], R2[4], RS[4];
#pragma unroll
    for (int j = 0; j < 8; j++)
    {
        Bv[j] = make_float2(0.f, 0.f);
        C[j] = 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) R1[j] = R2[j] = RS[j] = 0.f;
    float acc = 0.f;
    const int sp0 = s0 - 4;
#pragma unroll
    for (int d = 0; d <= 1; d++)
    {
        Bv[d] = pair(plane(b, sp0 + d), off);
        C[d] = pairC(planeC(sp0 + d), off);
    }
    __syncthreads();
    // the family sums of plane q from the rows' pair sums in ldsR[q & 1]
    auto family = [&](int q) {
        if (!famRow || q < s0 || q > s1) return;
        const size_t cpl = (size_t)((q + 1) >> 1) * (size_t)Lc.B.sS + (size_t)((m + 1) >> 1) * (size_t)Lc.B.sM;
        acc += ldsR[q & 1][i][lane] + ldsR[q & 1][i + 1][lane];
        if (!(q & 1) || q == s1)
        {
            if (okA) bc[cpl + lane + 1] = acc;
            acc = 0.f;
        }
    };
    auto step = [&](int sp, auto phase, auto colour) {
        constexpr int U = decltype(phase)::value;
        constexpr int E = decltype(colour)::value;
        constexpr int B0 = U & 7, Bm1 = (U + 7) & 7, Bm2 = (U + 6) & 7, Bm3 = (U + 5) & 7, Bm4 = (U + 4) & 7;
        constexpr int B2s = (U + 2) & 7;
        constexpr int R0 = U & 3, Rm1 = (U + 3) & 3, Rm2 = (U + 2) & 3, Rm3 = (U + 1) & 3, Rm4 = U & 3;
        // fin slots: plane t = sp - 4, t + 1, t - 1
        constexpr int F0 = U & 3, Fp = (U + 1) & 3, Fm = (U + 3) & 3;
        // prefetch: b and codes of plane sp + 2
        Bv[B2s] = pair(plane(b, sp + 2), off);
        C[B2s] = pairC(planeC(sp + 2), off);
        // the one barrier of the step: what the stages of the step before wrote (z1s of plane sp - 1, z2s of sp - 2,
        // s1s of sp - 3, fin of sp - 4, ldsR of sp - 5) is complete, and their readers of a step before are done
        __syncthreads();
        family(sp - 5);
        // Z1: zero-guess sweep, first colour, plane sp (element E)
        {
            const unsigned code = ((C[B0] & cm) >> (8 * E)) & 255u;
            const float v = first0(E ? Bv[B0].y : Bv[B0].x, code, sp, off + (uint32_t)E);
            R1[R0] = v;
            z1s[sp & 1][wave][fa + E] = v;
        }
        // Z2: zero-guess sweep, second colour, plane sp - 1 (element E there)
        if (doZ2)
        {
            const int q = sp - 1;
            const float(*rq)[kCols] = z1s[q & 1];
            const unsigned code = ((C[Bm1] & cm) >> (8 * E)) & 255u;
            const int f = fa + E;
            const float nb[6] = {R1[R0], R1[Rm2], rq[wave + 1][f], rq[wave - 1][f],
                                 E ? rq[wave][f + 1] : R1[Rm1], E ? R1[Rm1] : rq[wave][f - 1]};
            const float v = update(nb, E ? Bv[Bm1].y : Bv[Bm1].x, code, q, off + (uint32_t)E);
            R2[Rm1] = v;
            z2s[q & 1][wave][f] = v;
        }
        // S1: plain sweep, first colour, plane sp - 2 (element E there), from the zero-guess sweep's second colour
        if (doS1)
        {
            const int q = sp - 2;
            const float(*rq)[kCols] = z2s[q & 1];
            const unsigned code = ((C[Bm2] & cm) >> (8 * E)) & 255u;
            const int f = fa + E;
            const float nb[6] = {R2[Rm1], R2[Rm3], rq[wave + 1][f], rq[wave - 1][f],
                                 E ? rq[wave][f + 1] : R2[Rm2], E ? R2[Rm2] : rq[wave][f - 1]};
            const float v = update(nb, E ? Bv[Bm2].y : Bv[Bm2].x, code, q, off + (uint32_t)E);
            RS[Rm2] = v;
            s1s[q & 1][wave][f] = v;
        }
        // S2: plain sweep, second colour, plane sp - 3 (element E there): the swept x of that plane, kept for the
        // residual and stored on the output rows
        if (doS2)
        {
            const int q = sp - 3;
            const float(*rq)[kCols] = s1s[q & 1];
            const unsigned code = ((C[Bm3] & cm) >> (8 * E)) & 255u;
            const int f = fa + E;
            const float nb[6] = {RS[Rm2], RS[Rm4], rq[wave + 1][f], rq[wave - 1][f],
                                 E ? rq[wave][f + 1] : RS[Rm3], E ? RS[Rm3] : rq[wave][f - 1]};
            const float out = update(nb, E ? Bv[Bm3].y : Bv[Bm3].x, code, q, off + (uint32_t)E);
            const float va = E ? RS[Rm3] : out, vb = E ? out : RS[Rm3];
            fin[Fp][wave][0][1 + lane] = va;
            fin[Fp][wave][1][1 + lane] = vb;
            if (interior && q >= s0 && q <= s1)
            {
                float *po = xout + (size_t)q * (size_t)B.sS;
                if (okB)
                    *reinterpret_cast<float2 *>(po + off) = make_float2(va, vb);
                else if (okA)
                    po[off] = va;
            }
        }
        // residual of plane t = sp - 4 on the output rows (k_rbgs_restrict_row's operations) and the row's pair sums
        const int t = sp - 4;
        if (interior && t >= s0 && t <= s1)
        {
            const float(*ft)[2][kSlots] = fin[F0];
            const float(*fp)[2][kSlots] = fin[Fp];
            const float(*fm)[2][kSlots] = fin[Fm];
            const int p = 1 + lane;
            float r[2];
#pragma unroll
            for (int e = 0; e < 2; e++)
            {
                const unsigned code = ((C[Bm4] & cm) >> (8 * e)) & 255u;
                const float xc = ft[wave][e][p];
                const float nb[6] = {fp[wave][e][p], fm[wave][e][p], ft[wave + 1][e][p], ft[wave - 1][e][p],
                                     e ? ft[wave][0][p + 1] : ft[wave][1][p], e ? ft[wave][0][p] : ft[wave][1][p - 1]};
                const float bv = e ? Bv[Bm4].y : Bv[Bm4].x;
                float v = 0.f;
                if (code & 64u)
                {
                    if (!(code & 128u))
                    {
#pragma clang fp contract(off)
                        const float s6 = nb[0] + nb[1] + nb[2] + nb[3] + nb[4] + nb[5];
                        const float ax = L.cu * __builtin_fmaf(-(float)__popc(code & 63u), xc, s6);
                        v = bv - ax;
                    }
                    else
                        v = bv - apply32v(L, t, m, 2 * p - 1 + e, (int64_t)plane_off(t) + off + e, code, xc, nb);
                }
                r[e] = v;
            }
            ldsR[t & 1][i][lane] = r[0] + r[1];
        }
    };
    // the first-colour element at plane sp is (c1 - sp - m - 1) & 1
    auto march = [&](auto colour0) {
        constexpr int E0 = decltype(colour0)::value;
        for (int sp = sp0; sp <= s1 + 4; sp += 8)
        {
            step(sp, std::integral_constant<int, 0>(), std::integral_constant<int, E0>());
            if (sp + 1 > s1 + 4) break;
            step(sp + 1, std::integral_constant<int, 1>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 2 > s1 + 4) break;
            step(sp + 2, std::integral_constant<int, 2>(), std::integral_constant<int, E0>());
            if (sp + 3 > s1 + 4) break;
            step(sp + 3, std::integral_constant<int, 3>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 4 > s1 + 4) break;
            step(sp + 4, std::integral_constant<int, 4>(), std::integral_constant<int, E0>());
            if (sp + 5 > s1 + 4) break;
            step(sp + 5, std::integral_constant<int, 5>(), std::integral_constant<int, E0 ^ 1>());
            if (sp + 6 > s1 + 4) break;
            step(sp + 6, std::integral_constant<int, 6>(), std::integral_constant<int, E0>());
            if (sp + 7 > s1 + 4) break;
            step(sp + 7, std::integral_constant<int, 7>(), std::integral_constant<int, E0 ^ 1>());
        }
    };
    if ((c1 - sp0 - m - 1) & 1)
        march(std::integral_constant<int, 1>());
    else
        march(std::integral_constant<int, 0>());
    __syncthreads();
    family(s1);
}

// Residual of every fine cell, summed over each 2 x 2 x 2 family into the coarse right-hand side: fast pairs are
// neighbouring lanes (shuffle), mid pairs neighbouring rows of the block (LDS), slow pairs consecutive planes of
// the march (P even, chunks start on an odd plane). A family cut by the end of an odd interior keeps its 1 or 2
// children (the same clipping as the coarse widths).
__device__ __forceinline__ void residual_restrict_m_body(const Level32 &Lf, const float *__restrict__ x,
                                                         const float *__restrict__ b, const Level32 &Lc,
                                                         float *__restrict__ bc, const SolveState *st, int P, int lb)
{
    if (st->done) return;
    const Box &B = Lf.B;
    const March M = march_setup(B, P, lb);
    __shared__ float sh[kBlockMid][kBlockFast];
    const int tx = threadIdx.x & (kBlockFast - 1), ty = threadIdx.x / kBlockFast;
    int64_t idx = B.at(M.s0, M.in ? M.m : 1, M.in ? M.f : 1);
    float xm = 0.f, xc = 0.f;
    if (M.in)
    {
        xm = x[idx - B.sS];
        xc = x[idx];
    }
    float acc = 0.f;
    for (int s = M.s0; s <= M.s1; s++, idx += B.sS)
    {
        float r = 0.f, xp = 0.f;
        if (M.in)
        {
            xp = x[idx + B.sS];
            const unsigned code = Lf.c8[idx];
            if (code & 64u)
            {
                const float nb[6] = {xp, xm, x[idx + B.sM], x[idx - B.sM], x[idx + 1], x[idx - 1]};
                r = b[idx] - apply32v(Lf, s, M.m, M.f, idx, code, xc, nb);
            }
        }
        r += __shfl_xor(r, 1, 64);
        sh[ty][tx] = r;
        __syncthreads();
        acc += r + sh[ty ^ 1][tx];
        __syncthreads();
        if (!(s & 1) || s == M.s1)
        {
            if (M.in && !(tx & 1) && !(ty & 1))
                bc[Lc.B.at((s + 1) >> 1, (M.m + 1) >> 1, (M.f + 1) >> 1)] = acc;
            acc = 0.f;
        }
        xm = xc;
        xc = xp;
    }
}

__global__ __launch_bounds__(kBlockThreads) void k_update_residual64_m(Box B, Tables64 T,
                                                                       const uint8_t *__restrict__ ty,
                                                                       const double *__restrict__ x,
                                                                       const float *__restrict__ e,
                                                                       double *__restrict__ xout,
                                                                       const double *__restrict__ b,
                                                                       float *__restrict__ r32,
                                                                       double *__restrict__ r64, Red red,
                                                                       int P)
{
#pragma clang fp contract(off)
    const March M = march_setup(B, P);
    double rr = 0.0;
    if (M.in)
    {
        auto X = [&](int64_t i) { return e ? x[i] + (double)e[i] : x[i]; };
        int64_t idx = B.at(M.s0, M.m, M.f);
        double xm = X(idx - B.sS), xc = X(idx);
        for (int s = M.s0; s <= M.s1; s++, idx += B.sS)
        {
            const double xp = X(idx + B.sS);
            double r = 0.0;
            if (T.c8[idx] & 64u)
            {
                const double nb[6] = {xp, xm, X(idx + B.sM), X(idx - B.sM), X(idx + 1), X(idx - 1)};
                r = b[idx] - apply64v(B, T, ty, s, M.m, M.f, idx, xc, nb, T.c8[idx]);
                if (e) xout[idx] = xc;
            }
            if (r32) r32[idx] = (float)r;
            if (r64) r64[idx] = r;
            rr += r * r;
            xm = xc;
            xc = xp;
        }
    }
    double a[1] = {rr};
    block_sum<1>(a, red, gridDim.x, M.lb);
}

// As k_update_residual64_m, with the plane staged in LDS: each thread forms X = x + e of its cell (and, for 136 of
// the 256 threads, of one halo cell) once per plane, the in-plane neighbours come from LDS, and the loads of a
// plane are issued a plane ahead (the cell's own X two planes ahead). Tile positions outside the box load clamped
// in-box values that no interior cell couples to.
template <bool UNI>
__global__ __launch_bounds__(kBlockThreads) void k_update_residual64_t(Box B, Tables64 T,
                                                                       const uint8_t *__restrict__ ty,
                                                                       const double *__restrict__ x,
                                                                       const float *__restrict__ e,
                                                                       double *__restrict__ xout,
                                                                       const double *__restrict__ b,
                                                                       float *__restrict__ r32,
                                                                       double *__restrict__ r64, Red red,
                                                                       int P, int gate, int sumX)
{
#pragma clang fp contract(off)
    __shared__ double Xs[kBlockMid + 2][kBlockFast + 2];
    if (gate && !red.st->done) return;   // every block alike: no partial, no ticket
    const March M = march_setup(B, P);
    const int tx = threadIdx.x & (kBlockFast - 1), tyy = threadIdx.x / kBlockFast;
    const int f0 = M.f - tx, m0 = M.m - tyy;
    // own cell (clamped into the box) and its halo assignment
    const int fc = min(M.f, B.nf - 1), mcl = min(M.m, B.nm - 1);
    const int64_t own = (int64_t)mcl * B.sM + fc;
    int hx = -1, hy = -1;
    const int t = threadIdx.x;
    if (t < 64) { hy = 0; hx = t + 1; }
    else if (t < 128) { hy = kBlockMid + 1; hx = t - 64 + 1; }
    else if (t < 128 + kBlockMid) { hy = t - 128 + 1; hx = 0; }
    else if (t < 128 + 2 * kBlockMid) { hy = t - 128 - kBlockMid + 1; hx = kBlockFast + 1; }
    const bool halo = hx >= 0;
    const int64_t hoff = halo ? (int64_t)min(max(m0 - 1 + hy, 0), B.nm - 1) * B.sM + min(max(f0 - 1 + hx, 0), B.nf - 1)
                              : own;
    auto X = [&](int64_t i) { return e ? x[i] + (double)e[i] : x[i]; };
    auto plane = [&](int s) { return (int64_t)min(max(s, 0), B.ns - 1) * B.sS; };
    double rr = 0.0, xs = 0.0;   // <r, r>; with sumX the sum of x over the unknowns
    // registers: own X at s-1, s, s+1; halo X at s; b and code at s
    double Xm = X(plane(M.s0 - 1) + own), Xc = X(plane(M.s0) + own), Xp = X(plane(M.s0 + 1) + own);
    double H = halo ? X(plane(M.s0) + hoff) : 0.0;
    double bc = b[plane(M.s0) + own];
    unsigned cc = T.c8[plane(M.s0) + own];
    for (int s = M.s0; s <= M.s1; s++)
    {
        // prefetch: own X two planes ahead, halo, b and code one plane ahead
        const double Xn = X(plane(s + 2) + own);
        const double Hn = halo ? X(plane(s + 1) + hoff) : 0.0;
        const double bn = b[plane(s + 1) + own];
        const unsigned cn = T.c8[plane(s + 1) + own];
        Xs[tyy + 1][tx + 1] = Xc;
        if (halo) Xs[hy][hx] = H;
        __syncthreads();
        const int64_t idx = plane(s) + own;
        double r = 0.0;
        if (M.in && (cc & 64u))
        {
            const double nb[6] = {Xp, Xm, Xs[tyy + 2][tx + 1], Xs[tyy][tx + 1], Xs[tyy + 1][tx + 2], Xs[tyy + 1][tx]};
            r = bc - (UNI ? apply64u(T, idx, Xc, nb, cc) : apply64v(B, T, ty, s, M.m, M.f, idx, Xc, nb, cc));
            if (e) xout[idx] = Xc;
            xs += Xc;
        }
        if (M.in)
        {
            if (r32) r32[idx] = (float)r;
            if (r64) r64[idx] = r;
        }
        rr += r * r;
        __syncthreads();
        Xm = Xc;
        Xc = Xp;
        Xp = Xn;
        H = Hn;
        bc = bn;
        cc = cn;
    }
    if (sumX)
    {
        double a[2] = {rr, xs};
        block_sum<2>(a, red, gridDim.x, M.lb);
    }
    else
    {
        double a[1] = {rr};
        block_sum<1>(a, red, gridDim.x, M.lb);
    }
}

// UNI: the unit-coefficient row (apply64u), the one the FP64 Krylov and residual kernels use on uniform grids.
template <bool UNI>
__global__ __launch_bounds__(kBlockThreads) void k_apply64_m(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                             const double *__restrict__ x, double *__restrict__ y,
                                                             int P)
{
    const March M = march_setup(B, P);
    if (!M.in) return;
    int64_t idx = B.at(M.s0, M.m, M.f);
    double xm = x[idx - B.sS], xc = x[idx];
    for (int s = M.s0; s <= M.s1; s++, idx += B.sS)
    {
        const double xp = x[idx + B.sS];
        double v = 0.0;
        if (T.c8[idx] & 64u)
        {
            const double nb[6] = {xp, xm, x[idx + B.sM], x[idx - B.sM], x[idx + 1], x[idx - 1]};
            v = UNI ? apply64u(T, idx, xc, nb, T.c8[idx]) : apply64v(B, T, ty, s, M.m, M.f, idx, xc, nb, T.c8[idx]);
        }
        y[idx] = v;
        xm = xc;
        xc = xp;
    }
}

__global__ __launch_bounds__(kBlockThreads) void k_cg64_spmv_m(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                               const double *__restrict__ p, double *__restrict__ q,
                                                               const SolveState *st, Red red, int P)
{
    if (st->done) return;
    const March M = march_setup(B, P);
    double pq = 0.0;
    if (M.in)
    {
        int64_t idx = B.at(M.s0, M.m, M.f);
        double pm = p[idx - B.sS], pc = p[idx];
        for (int s = M.s0; s <= M.s1; s++, idx += B.sS)
        {
            const double pp = p[idx + B.sS];
            double v = 0.0;
            if (T.c8[idx] & 64u)
            {
                const double nb[6] = {pp, pm, p[idx + B.sM], p[idx - B.sM], p[idx + 1], p[idx - 1]};
                v = apply64v(B, T, ty, s, M.m, M.f, idx, pc, nb, T.c8[idx]);
                pq += pc * v;
            }
            q[idx] = v;
            pm = pc;
            pc = pp;
        }
    }
    double a[1] = {pq};
    block_sum<1>(a, red, gridDim.x, M.lb);
}

// ---- the inner iteration's kernels, for one system and batched over several -----------------------------------
//
// A batched launch runs one kernel for several systems at once. System q owns the logical blocks [start[q],
// start[q + 1]) of the launch (numbered after the XCD remap over the whole launch), and logical block g of the launch
// runs block g - start[q] of the system's own launch with the system's own arguments: the same work in the same
// order, so a system solved in a batch gives the bits it gives alone. Reductions stay per system (its partials, its
// ticket counter, its block count).

struct RowSweepArgs
{
    Level32 L;
    const float *xin;
    float *xout;
    const float *b;
    int c1, P, tilesM;
    const SolveState *st;
    double *dot;
    Prolong pro;
};

struct RestrictRowArgs
{
    Level32 Lf;
    const float *x;
    const float *b;
    Level32 Lc;
    float *bc;
    const SolveState *st;
    int P, tilesM;
};

struct SweepRestrictArgs
{
    Level32 L;
    const float *xin;
    float *xout;
    const float *b;
    int c1;
    Level32 Lc;
    float *bc;
    int P, tilesM;
    const SolveState *st;
};

struct RestrictMarchArgs
{
    Level32 Lf;
    const float *x;
    const float *b;
    Level32 Lc;
    float *bc;
    const SolveState *st;
    int P;
};

struct Dir64RowArgs
{
    Box B;
    Tables64 T;
    const float *z;
    float *ring;
    int64_t stride;
    const SolveState *st;
    Red red;
    int P, tilesM;
};

struct Update64RowArgs
{
    Box B;
    Tables64 T;
    const float *ring;
    int64_t stride;
    double *x, *r;
    float *r32;
    const double *b0;
    const SolveState *st;
    Red red;
    int P, tilesM;
};

struct CoarseArgs
{
    int n;
    const int32_t *order;
    const double *M;
    const float *b;
    float *x;
    const SolveState *st;
};

struct RzSumArgs
{
    SolveState *st;
    const double *rows;
    int nrows, seg;
    double *blockSums;
};

struct FinalizeArgs
{
    SolveState *st;
    double *partials;
    int nblk, nv, op;
};

struct BatchBlock
{
    int q, lb, nb;
};

// This block's system and its logical block in that system's launch (start: nsys + 1 increasing block offsets,
// start[0] = 0 and start[nsys] = the launch's grid).
__device__ __forceinline__ BatchBlock batch_block(const int *__restrict__ start, int nsys)
{
    const int g = logical_block();
    int lo = 0, hi = nsys - 1;   // the last system whose first block is at most g
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return BatchBlock{lo, g - start[lo], start[lo + 1] - start[lo]};
}

template <int TM, int K, bool ZERO, bool PRO, bool NU>
__global__ __launch_bounds__(64 * (TM + 2)) void k_rbgs_row(Level32 L, const float *__restrict__ xin,
                                                            float *__restrict__ xout, const float *__restrict__ b,
                                                            int c1, int P, int tilesM, const SolveState *st,
                                                            double *__restrict__ dot, Prolong pro)
{
    rbgs_row_body<TM, K, ZERO, PRO, NU>(L, xin, xout, b, c1, P, tilesM, st, dot, pro, logical_block());
}

template <int TM, int K, bool ZERO, bool PRO, bool NU>
__global__ __launch_bounds__(64 * (TM + 2)) void kb_rbgs_row(const RowSweepArgs *__restrict__ args,
                                                             const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const RowSweepArgs a = args[bb.q];
    rbgs_row_body<TM, K, ZERO, PRO, NU>(a.L, a.xin, a.xout, a.b, a.c1, a.P, a.tilesM, a.st, a.dot, a.pro, bb.lb);
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 6)) void k_rbgs2_row(Level32 L, const float *__restrict__ xin,
                                                             float *__restrict__ xout, const float *__restrict__ b,
                                                             int c1, int P, int tilesM, const SolveState *st,
                                                             Prolong pro)
{
    rbgs2_row_body<TM, K>(L, xin, xout, b, c1, P, tilesM, st, pro, logical_block());
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 6)) void kb_rbgs2_row(const RowSweepArgs *__restrict__ args,
                                                              const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const RowSweepArgs a = args[bb.q];
    rbgs2_row_body<TM, K>(a.L, a.xin, a.xout, a.b, a.c1, a.P, a.tilesM, a.st, a.pro, bb.lb);
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 2)) void k_restrict_row(Level32 Lf, const float *__restrict__ x,
                                                                const float *__restrict__ b, Level32 Lc,
                                                                float *__restrict__ bc, const SolveState *st, int P,
                                                                int tilesM)
{
    restrict_row_body<TM, K>(Lf, x, b, Lc, bc, st, P, tilesM, logical_block());
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 2)) void kb_restrict_row(const RestrictRowArgs *__restrict__ args,
                                                                 const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const RestrictRowArgs a = args[bb.q];
    restrict_row_body<TM, K>(a.Lf, a.x, a.b, a.Lc, a.bc, a.st, a.P, a.tilesM, bb.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_residual_restrict_m(Level32 Lf, const float *__restrict__ x,
                                                                       const float *__restrict__ b, Level32 Lc,
                                                                       float *__restrict__ bc, const SolveState *st,
                                                                       int P)
{
    residual_restrict_m_body(Lf, x, b, Lc, bc, st, P, logical_block());
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 4)) void k_rbgs_restrict_row(Level32 L, const float *__restrict__ xin,
                                                                     float *__restrict__ xout,
                                                                     const float *__restrict__ b, int c1, Level32 Lc,
                                                                     float *__restrict__ bc, int P, int tilesM,
                                                                     const SolveState *st)
{
    rbgs_restrict_row_body<TM, K>(L, xin, xout, b, c1, Lc, bc, P, tilesM, st, logical_block());
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 4)) void kb_rbgs_restrict_row(const SweepRestrictArgs *__restrict__ args,
                                                                      const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const SweepRestrictArgs a = args[bb.q];
    rbgs_restrict_row_body<TM, K>(a.L, a.xin, a.xout, a.b, a.c1, a.Lc, a.bc, a.P, a.tilesM, a.st, bb.lb);
}

template <int TM>
__global__ __launch_bounds__(64 * (TM + 8)) void k_rbgs_pre_restrict_row(Level32 L, float *__restrict__ xout,
                                                                         const float *__restrict__ b, int c1,
                                                                         Level32 Lc, float *__restrict__ bc, int P,
                                                                         int tilesM, const SolveState *st)
{
    rbgs_pre_restrict_row_body<TM>(L, xout, b, c1, Lc, bc, P, tilesM, st, logical_block());
}

template <int TM>
__global__ __launch_bounds__(64 * (TM + 8)) void kb_rbgs_pre_restrict_row(const SweepRestrictArgs *__restrict__ args,
                                                                          const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const SweepRestrictArgs a = args[bb.q];
    rbgs_pre_restrict_row_body<TM>(a.L, a.xout, a.b, a.c1, a.Lc, a.bc, a.P, a.tilesM, a.st, bb.lb);
}

__global__ __launch_bounds__(kBlockThreads) void kb_residual_restrict_m(const RestrictMarchArgs *__restrict__ args,
                                                                        const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const RestrictMarchArgs a = args[bb.q];
    residual_restrict_m_body(a.Lf, a.x, a.b, a.Lc, a.bc, a.st, a.P, bb.lb);
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 2)) void k_cg_dir64_row(Box B, Tables64 T, const float *__restrict__ z,
                                                                float *__restrict__ ring, int64_t stride,
                                                                const SolveState *st, Red red, int P, int tilesM)
{
    cg_dir64_row_body<TM, K>(B, T, z, ring, stride, st, red, P, tilesM, logical_block(), gridDim.x);
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 2)) void kb_cg_dir64_row(const Dir64RowArgs *__restrict__ args,
                                                                 const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const Dir64RowArgs a = args[bb.q];
    cg_dir64_row_body<TM, K>(a.B, a.T, a.z, a.ring, a.stride, a.st, a.red, a.P, a.tilesM, bb.lb, bb.nb);
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 2), 4) void k_cg_update64_row(Box B, Tables64 T,
                                                                   const float *__restrict__ ring, int64_t stride,
                                                                   double *__restrict__ x, double *__restrict__ r,
                                                                   float *__restrict__ r32,
                                                                   const double *__restrict__ b0, const SolveState *st,
                                                                   Red red, int P, int tilesM)
{
    cg_update64_row_body<TM, K>(B, T, ring, stride, x, r, r32, b0, st, red, P, tilesM, logical_block(), gridDim.x);
}

template <int TM, int K>
__global__ __launch_bounds__(64 * (TM + 2), 4) void kb_cg_update64_row(const Update64RowArgs *__restrict__ args,
                                                                    const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const Update64RowArgs a = args[bb.q];
    cg_update64_row_body<TM, K>(a.B, a.T, a.ring, a.stride, a.x, a.r, a.r32, a.b0, a.st, a.red, a.P, a.tilesM, bb.lb,
                                bb.nb);
}

__global__ __launch_bounds__(kBlockThreads) void k_coarse_solve(int n, const int32_t *__restrict__ order,
                                                                const double *__restrict__ M,
                                                                const float *__restrict__ b, float *__restrict__ x,
                                                                const SolveState *st)
{
    coarse_solve_body(n, order, M, b, x, st, blockIdx.x);
}

__global__ __launch_bounds__(kBlockThreads) void kb_coarse_solve(const CoarseArgs *__restrict__ args,
                                                                 const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const CoarseArgs a = args[bb.q];
    coarse_solve_body(a.n, a.order, a.M, a.b, a.x, a.st, bb.lb);
}

__global__ __launch_bounds__(kBlockThreads) void kb_rz_rows_sum(const RzSumArgs *__restrict__ args,
                                                                const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const RzSumArgs a = args[bb.q];
    rz_rows_sum_body(a.st, a.rows, a.nrows, a.seg, a.blockSums, bb.lb, bb.nb);
}

__global__ __launch_bounds__(1024) void kb_finalize(const FinalizeArgs *__restrict__ args,
                                                    const int *__restrict__ start, int nsys)
{
    const BatchBlock bb = batch_block(start, nsys);
    const FinalizeArgs a = args[bb.q];
    fin_apply(a.st, a.partials, a.nblk, a.nv, a.op, false);
}

__global__ void k_inner_setup(SolveState *st, double reduction, int maxIters)
{
    const double r2 = st->sums[1];
    st->thr2 = reduction * reduction * r2;
    st->iters = 0;
    st->first = 1;
    st->rz = 0.0;
    st->maxIters = maxIters;
    st->done = (r2 == 0.0 || maxIters <= 0) ? 1 : 0;
}

__global__ void k_state_init(SolveState *st, double thr2, int maxIters)
{
    st->thr2 = thr2;
    st->iters = 0;
    st->first = 1;
    st->maxIters = maxIters;
    st->done = maxIters <= 0 ? 2 : 0;
    st->accepted = 0;
    st->skip = 0;
    st->zeroStart = 0;
}

// The FP64 Krylov solve's set-up from the true residual of the guess (sums[0] = |b|^2, sums[1] = |r|^2), on the
// device so that the host needs no round trip first: target = tol * max(|b|, ref), the recursion aims at
// aimFactor * target; done at once when the guess is the answer (b = 0, |b| <= tol * ref) or already meets the
// target.
__global__ void k_krylov_init(SolveState *st, double tol, double ref, double aimFactor, int maxIters, int residualIsB)
{
    const double rr = residualIsB ? st->sums[0] : st->sums[1];
    const double bn = sqrt(st->sums[0]), rn = sqrt(rr);
    const double target = tol * fmax(bn, ref), aim = aimFactor * target;
    st->rr0 = rr;
    st->acc2 = target * target;
    st->thr2 = aim * aim;
    st->iters = 0;
    st->first = 1;
    st->maxIters = maxIters;
    st->accepted = 0;
    st->skip = (bn == 0.0 || (ref > 0.0 && bn <= tol * ref)) ? 1 : 0;
    st->done = (st->skip || rn < target) ? 1 : (maxIters <= 0 ? 2 : 0);
    st->zeroStart = residualIsB;
}

inline int grid_of(const Box &B) { return std::min(B.blocks(), kMaxGrid); }

// Blocks of a tile-looping kernel resident at once (blocks per CU at its occupancy x CUs, at most kMaxGrid): more
// blocks than that run in two batches, the second looping as long as the first on a partly idle GPU.
static int resident_grid(const void *kernel)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kBlockThreads, 0) != hipSuccess || per < 1 ||
        cus < 1)
        return kMaxGrid;
    return std::min(kMaxGrid, per * cus);
}

// Planes per marching chunk: enough chunks for about 4096 blocks, at least 2 planes, even (the restriction pairs
// planes), at most 64.
inline int march_planes(const Box &B)
{
    const int cols = B.tilesF * B.tilesM;
    const int chunks = std::max(1, (4096 + cols - 1) / cols);
    int P = (B.ns - 2 + chunks - 1) / chunks;
    P = std::max(2, std::min(64, P));
    return (P + 1) & ~1;
}
inline int march_grid(const Box &B, int P) { return B.tilesF * B.tilesM * ((B.ns - 2 + P - 1) / P); }
inline int grid_1d(int64_t n) { return (int)std::min<int64_t>((n + kBlockThreads - 1) / kBlockThreads, kMaxGrid); }

// Grids of at most kFuseBlocks blocks finalise their reduction in the kernel (no extra launch); larger grids are
// followed by a k_finalize launch (the in-kernel ticket wait would lengthen every block).
constexpr int kFuseBlocks = 512;

// ---- launch recording (batched solves) -------------------------------------------------------------------------

namespace
{
thread_local LaunchRecorder *tRecorder = nullptr;

// With a recorder set on this thread: the launch appended to it (the batched kernel, the system's grid and block,
// its arguments) instead of made. Returns whether it was recorded.
template <class A>
bool record(void (*fn)(const A *, const int *, int), int grid, int threads, const A &a, int kind = REC_FIXED,
            int tm = 0, int k = 0, int flags = 0)
{
    LaunchRecorder *r = tRecorder;
    if (!r) return false;
    LaunchRecord q;
    q.fn = reinterpret_cast<const void *>(fn);
    q.grid = grid;
    q.threads = threads;
    q.phase = r->phase;
    q.level = r->level;
    q.seq = r->seq++;
    q.kind = kind;
    q.tm = tm;
    q.k = k;
    q.flags = flags;
    q.args.assign(reinterpret_cast<const unsigned char *>(&a), reinterpret_cast<const unsigned char *>(&a) + sizeof(A));
    r->recs.push_back(std::move(q));
    return true;
}

// variant bits of the whole-row sweep: 1 ZERO, 2 PRO, 4 NU
template <int TM, int K>
const void *row_sweep_variant(int flags)
{
    switch (flags)
    {
    case 0: return reinterpret_cast<const void *>(kb_rbgs_row<TM, K, false, false, false>);
    case 1: return reinterpret_cast<const void *>(kb_rbgs_row<TM, K, true, false, false>);
    case 2: return reinterpret_cast<const void *>(kb_rbgs_row<TM, K, false, true, false>);
    case 4: return reinterpret_cast<const void *>(kb_rbgs_row<TM, K, false, false, true>);
    case 5: return reinterpret_cast<const void *>(kb_rbgs_row<TM, K, true, false, true>);
    case 6: return reinterpret_cast<const void *>(kb_rbgs_row<TM, K, false, true, true>);
    default: return nullptr;
    }
}

// The batched whole-row sweep of that shape and variant (the shapes launch_rbgs instantiates), or null.
const void *row_sweep_batched(int tm, int k, int flags)
{
    if (tm == 6 && k == 1) return row_sweep_variant<6, 1>(flags);
    if (tm == 14 && k == 1) return row_sweep_variant<14, 1>(flags);
    if (tm == 6 && k == 2) return row_sweep_variant<6, 2>(flags);
    if (tm == 14 && k == 2) return row_sweep_variant<14, 2>(flags);
    if (tm == 10 && k == 2) return row_sweep_variant<10, 2>(flags);
    if (tm == 6 && k == 4) return row_sweep_variant<6, 4>(flags);
    return nullptr;
}

thread_local hipEvent_t tEvStart = nullptr, tEvStop = nullptr;

template <int TM, int K, bool ZERO, bool PRO, bool NU>
void row_sweep(hipStream_t s, int grid, int threads, const RowSweepArgs &a)
{
    if (record(kb_rbgs_row<TM, K, ZERO, PRO, NU>, grid, threads, a, REC_ROW_SWEEP, TM, K,
               (ZERO ? 1 : 0) | (PRO ? 2 : 0) | (NU ? 4 : 0)))
        return;
    if (tEvStart)
    {
        hipExtLaunchKernelGGL(k_rbgs_row<TM, K, ZERO, PRO, NU>, dim3(grid), dim3(threads), 0, s, tEvStart, tEvStop, 0,
                              a.L, a.xin, a.xout, a.b, a.c1, a.P, a.tilesM, a.st, a.dot, a.pro);
        tEvStart = tEvStop = nullptr;
        return;
    }
    k_rbgs_row<TM, K, ZERO, PRO, NU><<<grid, threads, 0, s>>>(a.L, a.xin, a.xout, a.b, a.c1, a.P, a.tilesM, a.st,
                                                              a.dot, a.pro);
}

template <int TM, int K>
void row_sweep2(hipStream_t s, int grid, const RowSweepArgs &a)
{
    if (record(kb_rbgs2_row<TM, K>, grid, 64 * (TM + 6), a, REC_POST2, TM, K)) return;
    k_rbgs2_row<TM, K><<<grid, 64 * (TM + 6), 0, s>>>(a.L, a.xin, a.xout, a.b, a.c1, a.P, a.tilesM, a.st, a.pro);
}

template <int TM, int K>
void sweep_restrict(hipStream_t s, int grid, const SweepRestrictArgs &a)
{
    if (record(kb_rbgs_restrict_row<TM, K>, grid, 64 * (TM + 4), a, REC_SWEEP_RESTRICT, TM, K)) return;
    k_rbgs_restrict_row<TM, K><<<grid, 64 * (TM + 4), 0, s>>>(a.L, a.xin, a.xout, a.b, a.c1, a.Lc, a.bc, a.P,
                                                                 a.tilesM, a.st);
}

template <int TM>
void sweep_pre_restrict(hipStream_t s, int grid, const SweepRestrictArgs &a)
{
    if (record(kb_rbgs_pre_restrict_row<TM>, grid, 64 * (TM + 8), a, REC_PRE2, TM, 1)) return;
    k_rbgs_pre_restrict_row<TM><<<grid, 64 * (TM + 8), 0, s>>>(a.L, a.xout, a.b, a.c1, a.Lc, a.bc, a.P, a.tilesM,
                                                                 a.st);
}

// The fused sweep and restriction of that shape: launched (or recorded), false for a shape it has no form for.
bool sweep_restrict_shape(hipStream_t s, int tm, int k, int grid, const SweepRestrictArgs &a)
{
    if (k == 1 && tm == 6) sweep_restrict<6, 1>(s, grid, a);
    else if (k == 1 && tm == 8) sweep_restrict<8, 1>(s, grid, a);
    else if (k == 1 && tm == 10) sweep_restrict<10, 1>(s, grid, a);
    else if (k == 1 && tm == 12) sweep_restrict<12, 1>(s, grid, a);
    else if (k == 2 && tm == 6) sweep_restrict<6, 2>(s, grid, a);
    else if (k == 2 && tm == 8) sweep_restrict<8, 2>(s, grid, a);
    else return false;
    return true;
}

const void *sweep_restrict_batched(int tm, int k)
{
    if (k == 1 && tm == 6) return reinterpret_cast<const void *>(kb_rbgs_restrict_row<6, 1>);
    if (k == 1 && tm == 8) return reinterpret_cast<const void *>(kb_rbgs_restrict_row<8, 1>);
    if (k == 1 && tm == 10) return reinterpret_cast<const void *>(kb_rbgs_restrict_row<10, 1>);
    if (k == 1 && tm == 12) return reinterpret_cast<const void *>(kb_rbgs_restrict_row<12, 1>);
    if (k == 2 && tm == 6) return reinterpret_cast<const void *>(kb_rbgs_restrict_row<6, 2>);
    if (k == 2 && tm == 8) return reinterpret_cast<const void *>(kb_rbgs_restrict_row<8, 2>);
    return nullptr;
}

template <int TM, int K>
void restrict_row(hipStream_t s, int grid, int threads, const RestrictRowArgs &a)
{
    if (record(kb_restrict_row<TM, K>, grid, threads, a, REC_RESTRICT_ROW, TM, K)) return;
    k_restrict_row<TM, K><<<grid, threads, 0, s>>>(a.Lf, a.x, a.b, a.Lc, a.bc, a.st, a.P, a.tilesM);
}

template <int TM, int K>
void dir64_row(hipStream_t s, int grid, int threads, const Dir64RowArgs &a)
{
    if (record(kb_cg_dir64_row<TM, K>, grid, threads, a)) return;
    k_cg_dir64_row<TM, K><<<grid, threads, 0, s>>>(a.B, a.T, a.z, a.ring, a.stride, a.st, a.red, a.P, a.tilesM);
}

template <int TM, int K>
void update64_row(hipStream_t s, int grid, int threads, const Update64RowArgs &a)
{
    if (record(kb_cg_update64_row<TM, K>, grid, threads, a)) return;
    k_cg_update64_row<TM, K><<<grid, threads, 0, s>>>(a.B, a.T, a.ring, a.stride, a.x, a.r, a.r32, a.b0, a.st,
                                                      a.red, a.P, a.tilesM);
}
}   // namespace


Red red_of(double *partials, SolveState *st, FinOp op, int blocks)
{
    return Red{partials, st, (int)op, blocks <= kFuseBlocks ? 1 : 0};
}

// After a launch that reduced into red over nblk blocks: the separate finalisation when the kernel did not fuse it.
hipError_t finish(hipStream_t s, const Red &red, int nblk, int nv)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || red.fused) return e;
    if (record(kb_finalize, 1, fin_threads(nblk), FinalizeArgs{red.st, red.partials, nblk, nv, red.op}))
        return hipSuccess;
    k_finalize<<<1, fin_threads(nblk), 0, s>>>(red.st, red.partials, nblk, nv, red.op);
    return hipGetLastError();
}

#define LAUNCH_OK()                                    \
    do                                                 \
    {                                                  \
        hipError_t e_ = hipGetLastError();             \
        if (e_ != hipSuccess) return e_;               \
    } while (0)

}  // namespace

void set_launch_recorder(LaunchRecorder *r) { tRecorder = r; }

void set_launch_events(hipEvent_t start, hipEvent_t stop)
{
    tEvStart = start;
    tEvStop = stop;
}

bool launch_events_pending() { return tEvStart != nullptr; }

void record_slot(int phase, int level)
{
    if (!tRecorder) return;
    tRecorder->phase = phase;
    tRecorder->level = level;
    tRecorder->seq = 0;
}

void retile_rows(const LaunchRecord &r, std::vector<int> &tms)
{
    tms.clear();
    if (r.kind == REC_SWEEP_RESTRICT)
    {
        rbgs_restrict_rows(r.k, tms);
        return;
    }
    if (r.kind != REC_ROW_SWEEP) return;
    // 14-row tiles need 1024-thread blocks, whose register limit the prolongation-folding K = 2 sweep exceeds (10
    // rows there on uniform levels); K = 4 rows only in 6-row tiles
    if (r.k == 4) tms = {6};
    else if (r.k == 2 && (r.flags & 2)) tms = (r.flags & 4) ? std::vector<int>{6} : std::vector<int>{6, 10};
    else tms = {6, 14};
}

bool retile_record(LaunchRecord &r, int tm, int p)
{
    const int P = std::max(1, std::min(64, p));
    const int Pe = std::max(2, P + (P & 1));
    switch (r.kind)
    {
    case REC_ROW_SWEEP:
    {
        RowSweepArgs a;
        std::memcpy(&a, r.args.data(), sizeof a);
        const int TM = tm > 0 ? tm : r.tm;
        std::vector<int> ok;
        retile_rows(r, ok);
        const void *fn = row_sweep_batched(TM, r.k, r.flags);
        if (!fn || std::find(ok.begin(), ok.end(), TM) == ok.end()) return false;
        const Box &B = a.L.B;
        a.P = P;
        a.tilesM = std::max(1, (B.nm - 2 + TM - 1) / TM);
        r.grid = a.tilesM * std::max(1, (B.ns - 2 + P - 1) / P);
        r.threads = 64 * (TM + 2);
        r.fn = fn;
        r.tm = TM;
        std::memcpy(r.args.data(), &a, sizeof a);
        return true;
    }
    case REC_SWEEP_RESTRICT:
    {
        SweepRestrictArgs a;
        std::memcpy(&a, r.args.data(), sizeof a);
        const int TM = tm > 0 ? tm : r.tm;
        const void *fn = sweep_restrict_batched(TM, r.k);
        if (!fn) return false;
        a.P = Pe;
        a.tilesM = std::max(1, (a.L.B.nm - 2 + TM - 1) / TM);
        r.grid = a.tilesM * std::max(1, (a.L.B.ns - 2 + Pe - 1) / Pe);
        r.threads = 64 * (TM + 4);
        r.fn = fn;
        r.tm = TM;
        std::memcpy(r.args.data(), &a, sizeof a);
        return true;
    }
    case REC_POST2:
    {
        RowSweepArgs a;
        std::memcpy(&a, r.args.data(), sizeof a);
        a.P = P;
        r.grid = a.tilesM * std::max(1, (a.L.B.ns - 2 + P - 1) / P);
        std::memcpy(r.args.data(), &a, sizeof a);
        return true;
    }
    case REC_PRE2:
    {
        SweepRestrictArgs a;
        std::memcpy(&a, r.args.data(), sizeof a);
        a.P = Pe;
        r.grid = a.tilesM * std::max(1, (a.L.B.ns - 2 + Pe - 1) / Pe);
        std::memcpy(r.args.data(), &a, sizeof a);
        return true;
    }
    case REC_RESTRICT_ROW:
    {
        RestrictRowArgs a;
        std::memcpy(&a, r.args.data(), sizeof a);
        a.P = Pe;
        r.grid = a.tilesM * std::max(1, (a.Lf.B.ns - 2 + Pe - 1) / Pe);
        std::memcpy(r.args.data(), &a, sizeof a);
        return true;
    }
    case REC_RESTRICT_MARCH:
    {
        RestrictMarchArgs a;
        std::memcpy(&a, r.args.data(), sizeof a);
        a.P = Pe;
        r.grid = march_grid(a.Lf.B, Pe);
        std::memcpy(r.args.data(), &a, sizeof a);
        return true;
    }
    default: return false;
    }
}

hipError_t launch_batched(hipStream_t s, const void *fn, int grid, int threads, const void *args, const int *start,
                          int nsys)
{
    void *kargs[] = {(void *)&args, (void *)&start, (void *)&nsys};
    return hipLaunchKernel(fn, dim3(grid), dim3(threads), kargs, 0, s);
}

hipError_t launch_scatter64(hipStream_t s, int64_t n, const int32_t *order, const double *src, double *box,
                            SolveState *st, int sumSlot, double *partials)
{
    if (n <= 0) return hipSuccess;
    const int G = grid_1d(n);
    const FinOp op = sumSlot == 0 ? FIN_SUM0 : (sumSlot == 1 ? FIN_SUM1 : FIN_SUM2);
    const Red red = red_of(partials, st, op, G);
    k_scatter64<<<G, kBlockThreads, 0, s>>>(n, order, src, box, red, sumSlot >= 0);
    if (sumSlot < 0) return hipGetLastError();
    return finish(s, red, G, 1);
}

hipError_t launch_scatter_rhs64(hipStream_t s, int64_t n, const int32_t *order, const double *src, double *b64,
                                double *r64, float *r32, double *x64, SolveState *st, double *partials)
{
    if (n <= 0) return hipSuccess;
    const int G = grid_1d(n);
    const Red red = red_of(partials, st, FIN_SUM0, G);
    k_scatter_rhs64<<<G, kBlockThreads, 0, s>>>(n, order, src, b64, r64, r32, x64, red);
    return finish(s, red, G, 1);
}

hipError_t launch_gather64(hipStream_t s, int64_t n, const int32_t *order, const double *box, double *dst,
                           const SolveState *st, int meanSlot, int gate)
{
    if (n <= 0) return hipSuccess;
    k_gather64<<<grid_1d(n), kBlockThreads, 0, s>>>(n, order, box, dst, st, meanSlot, gate);
    return hipGetLastError();
}

hipError_t launch_sum64(hipStream_t s, int64_t n, const int32_t *order, const double *box, SolveState *st, int slot,
                        double *partials, int gate)
{
    if (n <= 0) return hipSuccess;
    const int G = grid_1d(n);
    const FinOp op = gate ? FIN_SUM2_ACC : (slot == 0 ? FIN_SUM0 : (slot == 1 ? FIN_SUM1 : FIN_SUM2));
    const Red red = red_of(partials, st, op, G);
    k_sum64<<<G, kBlockThreads, 0, s>>>(n, order, box, red, gate);
    return finish(s, red, G, 1);
}

hipError_t launch_update_residual64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty,
                                    const double *x, const float *e, double *xout, const double *b, float *r32,
                                    double *r64, SolveState *st, double *partials, int gate, int sumX)
{
    if (B.threeD)
    {
        sumX = gate && sumX ? 1 : 0;
        const FinOp op = gate ? (sumX ? FIN_TRUE_RES_SUM : FIN_TRUE_RES) : FIN_SUM1;
        const int P = march_planes(B), G = march_grid(B, P);
        const Red red = red_of(partials, st, op, G);
        if (T.unitCoef)
            k_update_residual64_t<true><<<G, kBlockThreads, 0, s>>>(B, T, ty, x, e, xout, b, r32, r64, red, P, gate,
                                                                    sumX);
        else
            k_update_residual64_t<false><<<G, kBlockThreads, 0, s>>>(B, T, ty, x, e, xout, b, r32, r64, red, P,
                                                                     gate, sumX);
        return finish(s, red, G, sumX ? 2 : 1);
    }
    sumX = gate && sumX ? 1 : 0;
    const FinOp op = gate ? (sumX ? FIN_TRUE_RES_SUM : FIN_TRUE_RES) : FIN_SUM1;
    static const int resident = resident_grid(reinterpret_cast<const void *>(k_update_residual64));
    const int G = std::min(B.blocks(), resident);
    const Red red = red_of(partials, st, op, G);
    k_update_residual64<<<G, kBlockThreads, 0, s>>>(B, T, ty, x, e, xout, b, r32, r64, red, gate, sumX);
    return finish(s, red, G, sumX ? 2 : 1);
}

hipError_t launch_apply64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *x,
                          double *y)
{
    if (B.threeD)
    {
        const int P = march_planes(B);
        if (T.unitCoef)
            k_apply64_m<true><<<march_grid(B, P), kBlockThreads, 0, s>>>(B, T, ty, x, y, P);
        else
            k_apply64_m<false><<<march_grid(B, P), kBlockThreads, 0, s>>>(B, T, ty, x, y, P);
        return hipGetLastError();
    }
    k_apply64<<<grid_of(B), kBlockThreads, 0, s>>>(B, T, ty, x, y);
    return hipGetLastError();
}

hipError_t launch_cg64_init(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *r,
                            double *p, SolveState *st, double *partials)
{
    const int G = grid_of(B);
    const Red red = red_of(partials, st, FIN_CG64_INIT, G);
    k_cg64_init<<<G, kBlockThreads, 0, s>>>(B, T, ty, r, p, red);
    return finish(s, red, G, 1);
}

hipError_t launch_cg64_spmv(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *p,
                            double *q, SolveState *st, double *partials)
{
    if (B.threeD)
    {
        const int P = march_planes(B), G = march_grid(B, P);
        const Red red = red_of(partials, st, FIN_CG64_PQ, G);
        k_cg64_spmv_m<<<G, kBlockThreads, 0, s>>>(B, T, ty, p, q, st, red, P);
        return finish(s, red, G, 1);
    }
    const int G = grid_of(B);
    const Red red = red_of(partials, st, FIN_CG64_PQ, G);
    k_cg64_spmv<<<G, kBlockThreads, 0, s>>>(B, T, ty, p, q, st, red);
    return finish(s, red, G, 1);
}

hipError_t launch_cg64_update(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, double *x,
                              double *r, const double *p, const double *q, SolveState *st, double *partials)
{
    const int G = grid_of(B);
    const Red red = red_of(partials, st, FIN_CG64_UPDATE, G);
    k_cg64_update<<<G, kBlockThreads, 0, s>>>(B, T, ty, x, r, p, q, st, red);
    return finish(s, red, G, 2);
}

hipError_t launch_cg64_direction(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty,
                                 const double *r, double *p, SolveState *st)
{
    k_cg64_direction<<<grid_of(B), kBlockThreads, 0, s>>>(B, T, ty, r, p, st);
    return hipGetLastError();
}

// The whole-row sweep reads and writes cell pairs (2p + 1, 2p + 2) with single aligned accesses: the arrays must
// have element 1 of every row pair-aligned (odd float offsets 8-byte aligned, odd code offsets 2-byte aligned).
bool rbgs_pairs_aligned(const Level32 &L, const float *xin, const float *xout, const float *b)
{
    auto oddAligned = [](const void *p, uintptr_t a) { return p == nullptr || (((uintptr_t)p + (a >> 1)) % a) == 0; };
    return oddAligned(xin, 8) && oddAligned(xout, 8) && oddAligned(b, 8) && oddAligned(L.c8, 2) &&
           (L.B.sM % 16) == 0 && (L.B.sS % 16) == 0;
}

int rbgs_family(const Level32 &L, const float *xin, const float *xout, const float *b)
{
    const Box &B = L.B;
    // NPPE_RBGS_MARCH=1 keeps 3D levels on the marching sweep (the tests compare it with the whole-row sweep)
    const char *march = std::getenv("NPPE_RBGS_MARCH");
    if (march && std::atoi(march) != 0 && B.threeD) return RBGS_MARCH;
    if (B.threeD && B.nf - 2 <= 512 && rbgs_pairs_aligned(L, xin, xout, b)) return RBGS_ROW;
    return B.threeD ? RBGS_MARCH : RBGS_2D;
}

int rbgs_row_tile(const Box &B, int tileRows, bool folding, bool uniform)
{
    const int K = B.nf - 2 <= 128 ? 1 : (B.nf - 2 <= 256 ? 2 : 4);
    if (K == 2 && folding && uniform) return 10;
    return (tileRows >= 10 && (K == 1 || (K == 2 && !folding))) ? 14 : 6;
}

int64_t rbgs_blocks(const Box &B, int family, int tileRows, int planesPerChunk, bool folding, bool uniform)
{
    const int P = std::max(1, planesPerChunk);
    const int64_t chunks = B.threeD ? std::max(1, (B.ns - 2 + P - 1) / P) : 1;
    if (family == RBGS_ROW)
    {
        const int TM = rbgs_row_tile(B, tileRows, folding, uniform);
        return (int64_t)std::max(1, (B.nm - 2 + TM - 1) / TM) * chunks;
    }
    if (family == RBGS_MARCH)
    {
        const int TM = tileRows <= 4 ? 4 : (tileRows <= 6 ? 6 : 8);
        return (int64_t)std::max(1, (B.nf - 2 + kPW - 1) / kPW) * std::max(1, (B.nm - 2 + TM - 1) / TM) * chunks;
    }
    const int TM = tileRows == 8 ? 8 : 16;
    return (int64_t)std::max(1, (B.nf - 2 + kTF - 1) / kTF) * std::max(1, (B.nm - 2 + TM - 1) / TM);
}

void rbgs_tile_candidates(int family, std::vector<int> &tileRows)
{
    if (family == RBGS_ROW) tileRows = {6, 14};
    else if (family == RBGS_MARCH) tileRows = {4, 6, 8};
    else tileRows = {8, 16};
}

// Tile rows of the two-sweep march: three halo rows each side of 16 waves, rows of at most 128 cells.
constexpr int kRbgs2Rows = 10;

bool rbgs2_level(const Level32 &L, const float *xin, const float *xout, const float *b)
{
    const char *env = std::getenv("NPPE_FUSE_POST");
    if (env && std::atoi(env) == 0) return false;
    const char *tiles = std::getenv("NPPE_TILE_PASSES");
    if (tiles && std::atoi(tiles) != 0) return false;
    return L.B.threeD && L.uniform && L.B.nf - 2 <= 128 && rbgs_family(L, xin, xout, b) == RBGS_ROW;
}

int64_t rbgs2_blocks(const Box &B, int planesPerChunk)
{
    const int P = std::max(1, planesPerChunk);
    return (int64_t)((B.nm - 2 + kRbgs2Rows - 1) / kRbgs2Rows) * ((B.ns - 2 + P - 1) / P);
}

hipError_t launch_rbgs2(hipStream_t s, const Level32 &L, const float *xin, float *xout, const float *b, int c1,
                        int planesPerChunk, SolveState *st, const Level32 &Lc, const float *xc)
{
    const Box &B = L.B;
    if (!rbgs2_level(L, xin, xout, b) || xin == xout || !xc) return hipErrorInvalidValue;
    const int P = std::max(1, planesPerChunk);
    const int tilesM = (B.nm - 2 + kRbgs2Rows - 1) / kRbgs2Rows;
    const int grid = tilesM * ((B.ns - 2 + P - 1) / P);
    const Prolong pro{xc, Lc.B.sS, (int)Lc.B.sM, Lc.B.ns, Lc.B.nf};
    const RowSweepArgs ra{L, xin, xout, b, c1, P, tilesM, st, nullptr, pro};
    row_sweep2<kRbgs2Rows, 1>(s, grid, ra);
    return hipGetLastError();
}

hipError_t launch_rbgs(hipStream_t s, const Level32 &L, const float *xin, float *xout, const float *b, int c1,
                       int zero, int tileRows, int planesPerChunk, SolveState *st, double *rzPartials,
                       const Level32 *prolongFrom, const float *prolongXc)
{
    if (prolongFrom && (zero || (rbgs_family(L, xin, xout, b) != RBGS_ROW && rbgs_family(L, xin, xout, b) != RBGS_2D)))
    {
        // the correction as its own pass, then the plain sweep
        hipError_t e = launch_prolong(s, L, const_cast<float *>(xin), *prolongFrom, prolongXc, st);
        if (e != hipSuccess) return e;
        return launch_rbgs(s, L, xin, xout, b, c1, zero, tileRows, planesPerChunk, st, rzPartials);
    }
    if (rzPartials && rbgs_family(L, xin, xout, b) != RBGS_ROW)
    {
        hipError_t e = launch_rbgs(s, L, xin, xout, b, c1, zero, tileRows, planesPerChunk, st, nullptr);
        if (e != hipSuccess) return e;
        if (!L.B.threeD) return launch_cg_rz(s, L, b, xout, st, rzPartials);
        // 3D: the row partials of the whole-row sweep and their two-stage sum (the same <r, z> bit for bit)
        const int64_t nrows = (int64_t)(L.B.ns - 2) * (L.B.nm - 2);
        const int G = (int)((nrows + kBlockThreads / 64 - 1) / (kBlockThreads / 64));
        k_rz_rows<<<G, kBlockThreads, 0, s>>>(L, b, xout, st, rzPartials);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        return launch_rz_rows_finalize(s, L.B, st, rzPartials);
    }
    const Box &B = L.B;
    if (rbgs_family(L, xin, xout, b) == RBGS_ROW)
    {
        const int K = B.nf - 2 <= 128 ? 1 : (B.nf - 2 <= 256 ? 2 : 4);
        // 14-row tiles for K <= 2 when asked; the prolongation-folding K = 2 sweep in 10-row tiles on uniform levels
        // (1024-thread blocks cap its registers at 128, which it exceeds; its 157 fit 768-thread blocks, and at 256^3
        // it took 61 us there against 69 us in 6-row tiles)
        const int TM = rbgs_row_tile(B, tileRows, prolongFrom != nullptr, L.uniform != 0);
        const int tilesM = std::max(1, (B.nm - 2 + TM - 1) / TM);
        const int P = std::max(1, planesPerChunk);
        const int chunks = std::max(1, (B.ns - 2 + P - 1) / P);
        const int grid = tilesM * chunks;
        const int threads = 64 * (TM + 2);
        Prolong pro{nullptr, 0, 0, 1, 1};
        if (prolongFrom) pro = Prolong{prolongXc, prolongFrom->B.sS, (int)prolongFrom->B.sM, prolongFrom->B.ns,
                                       prolongFrom->B.nf};
        const RowSweepArgs ra{L, xin, xout, b, c1, P, tilesM, st, rzPartials, pro};
#define NPPE_ROW_NU(TMv, Kv, NUv)                                                                                  \
    (zero ? row_sweep<TMv, Kv, true, false, NUv>(s, grid, threads, ra)                                             \
          : (prolongFrom ? row_sweep<TMv, Kv, false, true, NUv>(s, grid, threads, ra)                              \
                         : row_sweep<TMv, Kv, false, false, NUv>(s, grid, threads, ra)))
#define NPPE_ROW(TMv, Kv) (L.uniform ? NPPE_ROW_NU(TMv, Kv, false) : NPPE_ROW_NU(TMv, Kv, true))
        if (TM == 10)
            NPPE_ROW(10, 2);
        else if (TM == 14)
        {
            if (K == 1) NPPE_ROW(14, 1);
            else NPPE_ROW(14, 2);
        }
        else if (K == 1) NPPE_ROW(6, 1);
        else if (K == 2) NPPE_ROW(6, 2);
        else NPPE_ROW(6, 4);
#undef NPPE_ROW
#undef NPPE_ROW_NU
        return hipGetLastError();
    }
    if (B.threeD)
    {
        const int TM = tileRows <= 4 ? 4 : (tileRows <= 6 ? 6 : 8);
        const int tilesF = std::max(1, (B.nf - 2 + kPW - 1) / kPW);
        const int tilesM = std::max(1, (B.nm - 2 + TM - 1) / TM);
        const int P = std::max(1, planesPerChunk);
        const int chunks = std::max(1, (B.ns - 2 + P - 1) / P);
        const int grid = tilesF * tilesM * chunks;
#define NPPE_RBGS_MARCH(TMv)                                                                                       \
    (L.uniform ? (k_rbgs_march<TMv, true><<<grid, 64 * (TMv + 2), 0, s>>>(L, xin, xout, b, c1, zero, P, tilesF,    \
                                                                          tilesM, st),                             \
                  0)                                                                                               \
               : (k_rbgs_march<TMv, false><<<grid, 64 * (TMv + 2), 0, s>>>(L, xin, xout, b, c1, zero, P, tilesF,   \
                                                                           tilesM, st),                            \
                  0))
        if (TM == 4) NPPE_RBGS_MARCH(4);
        else if (TM == 6) NPPE_RBGS_MARCH(6);
        else NPPE_RBGS_MARCH(8);
#undef NPPE_RBGS_MARCH
        return hipGetLastError();
    }
    const int tilesF = std::max(1, (B.nf - 2 + kTF - 1) / kTF);
    const int TM = tileRows == 8 ? 8 : 16;
    const int tilesM = std::max(1, (B.nm - 2 + TM - 1) / TM);
    const int grid = tilesF * tilesM;
    // 2D: the coarse correction folded into the sweep's loads
    const bool fold = prolongFrom && prolongXc && !B.threeD && !zero;
    const Prolong pro{fold ? prolongXc : nullptr, fold ? prolongFrom->B.sS : 0, fold ? (int)prolongFrom->B.sM : 0,
                      fold ? prolongFrom->B.ns : 0, fold ? prolongFrom->B.nf : 0};
    if (TM == 8)
    {
        if (fold) k_rbgs<8, true><<<grid, kBlockThreads, 0, s>>>(L, xin, xout, b, c1, zero, tilesF, st, pro);
        else k_rbgs<8, false><<<grid, kBlockThreads, 0, s>>>(L, xin, xout, b, c1, zero, tilesF, st, pro);
    }
    else
    {
        if (fold) k_rbgs<16, true><<<grid, kBlockThreads, 0, s>>>(L, xin, xout, b, c1, zero, tilesF, st, pro);
        else k_rbgs<16, false><<<grid, kBlockThreads, 0, s>>>(L, xin, xout, b, c1, zero, tilesF, st, pro);
    }
    return hipGetLastError();
}

// The whole-row marches of the FP64 Krylov passes and the restriction: 3D boxes with uniform spacing (unit
// coefficients), rows of at most 256 cells, and every array with element 1 of each row pair-aligned. Tile rows and
// the chunk length follow from the box alone (never from timing), so the order of the partial sums is a function
// of the box.
struct RowPlan
{
    bool ok;
    int K, TM, P, tilesM, G;
};

RowPlan row_plan(const Box &B, const uint8_t *c8, bool eligible,
                 std::initializer_list<std::pair<const void *, uintptr_t>> arrays, int tm, int blocksPerCU,
                 bool evenP = false)
{
    RowPlan R{};
    auto oddAligned = [](const void *p, uintptr_t a) { return p == nullptr || (((uintptr_t)p + (a >> 1)) % a) == 0; };
    // NPPE_TILE_PASSES=1 runs the tile kernels instead (the tests compare the two)
    const char *tiles = std::getenv("NPPE_TILE_PASSES");
    R.ok = eligible && !(tiles && std::atoi(tiles) != 0) && B.threeD && B.nf - 2 <= 256 && (B.sM % 16) == 0 &&
           (B.sS % 16) == 0 && oddAligned(c8, 2);
    for (const auto &a : arrays) R.ok = R.ok && oddAligned(a.first, a.second);
    if (!R.ok) return R;
    constexpr int kCUs = 256;
    R.K = B.nf - 2 <= 128 ? 1 : 2;
    R.TM = tm;
    R.tilesM = std::max(1, (B.nm - 2 + tm - 1) / tm);
    const int planes = B.ns - 2;
    const int chunksMax = std::max(1, kCUs * blocksPerCU / R.tilesM);
    // at least kMinPlanes planes per chunk (as many as the box has): a shorter chunk spends most of its loads on the
    // planes it reads ahead, which costs little on a small box alone but much when many systems share a launch
    constexpr int kMinPlanes = 8;
    R.P = std::max(std::min(kMinPlanes, planes), (planes + chunksMax - 1) / chunksMax);
    R.P = std::max(1, R.P);
    if (evenP) R.P += R.P & 1;
    R.G = R.tilesM * ((planes + R.P - 1) / R.P);
    return R;
}

hipError_t launch_residual_restrict(hipStream_t s, const Level32 &Lf, const float *x, const float *b,
                                    const Level32 &Lc, float *bc, const SolveState *st)
{
    if (Lf.B.threeD)
    {
        // 6-row blocks, two per CU (the same time at 256^3, a little less on the small levels than 14-row blocks)
        const RowPlan rp = row_plan(Lf.B, Lf.c8, Lf.uniform != 0, {{x, 8}, {b, 8}}, 6, 2, true);
        if (rp.ok)
        {
            const RestrictRowArgs a{Lf, x, b, Lc, bc, st, rp.P, rp.tilesM};
            if (rp.K == 1)
                restrict_row<6, 1>(s, rp.G, 64 * 8, a);
            else
                restrict_row<6, 2>(s, rp.G, 64 * 8, a);
            return hipGetLastError();
        }
        const int P = march_planes(Lf.B), G = march_grid(Lf.B, P);
        if (record(kb_residual_restrict_m, G, kBlockThreads, RestrictMarchArgs{Lf, x, b, Lc, bc, st, P},
                   REC_RESTRICT_MARCH))
            return hipSuccess;
        k_residual_restrict_m<<<G, kBlockThreads, 0, s>>>(Lf, x, b, Lc, bc, st, P);
        return hipGetLastError();
    }
    static const int resident = resident_grid(reinterpret_cast<const void *>(k_residual_restrict));
    k_residual_restrict<<<std::min(Lc.B.blocks(), resident), kBlockThreads, 0, s>>>(Lf, x, b, Lc, bc, st);
    return hipGetLastError();
}

void rbgs_restrict_rows(int k, std::vector<int> &tms)
{
    // (two pairs per lane: larger tiles exceed the registers of their blocks)
    if (k == 1) tms = {6, 8, 10, 12};
    else if (k == 2) tms = {6, 8};
    else tms.clear();
}

bool rbgs_restrict_fusable(const Level32 &Lf, const float *xin, const float *xout, const float *b, const float *bc)
{
    const char *env = std::getenv("NPPE_FUSE_RESTRICT");
    if (env && std::atoi(env) == 0) return false;
    const char *tiles = std::getenv("NPPE_TILE_PASSES");
    if (tiles && std::atoi(tiles) != 0) return false;
    return Lf.B.threeD && Lf.uniform && Lf.B.nf - 2 <= 256 && bc && rbgs_family(Lf, xin, xout, b) == RBGS_ROW;
}

hipError_t launch_rbgs_restrict(hipStream_t s, const Level32 &Lf, const float *xin, float *xout, const float *b,
                                int c1, const Level32 &Lc, float *bc, int tileRows, int planesPerChunk,
                                const SolveState *st)
{
    const Box &B = Lf.B;
    const int K = B.nf - 2 <= 128 ? 1 : 2;
    std::vector<int> tms;
    rbgs_restrict_rows(K, tms);
    int TM = tms.front();
    for (int t : tms)
        if (t <= tileRows) TM = t;
    int P = std::max(2, planesPerChunk);
    P += P & 1;
    const int tilesM = std::max(1, (B.nm - 2 + TM - 1) / TM);
    const int grid = tilesM * std::max(1, (B.ns - 2 + P - 1) / P);
    const SweepRestrictArgs a{Lf, xin, xout, b, c1, Lc, bc, P, tilesM, st};
    if (!sweep_restrict_shape(s, TM, K, grid, a)) return hipErrorInvalidValue;
    return hipGetLastError();
}

// Output rows of the whole pre-smoothing march: four halo rows each side in 16 waves (even: rows are summed in pairs).
constexpr int kPreRows = 8;

bool rbgs_pre_restrict_level(const Level32 &Lf, const float *xin, const float *xout, const float *b, const float *bc)
{
    const char *env = std::getenv("NPPE_FUSE_PRE");
    if (env && std::atoi(env) == 0) return false;
    return Lf.B.nf - 2 <= 128 && rbgs_restrict_fusable(Lf, xin, xout, b, bc);
}

int64_t rbgs_pre_restrict_blocks(const Box &B, int planesPerChunk)
{
    int P = std::max(2, planesPerChunk);
    P += P & 1;
    return (int64_t)((B.nm - 2 + kPreRows - 1) / kPreRows) * ((B.ns - 2 + P - 1) / P);
}

hipError_t launch_rbgs_pre_restrict(hipStream_t s, const Level32 &Lf, float *xout, const float *b, int c1,
                                    const Level32 &Lc, float *bc, int planesPerChunk, const SolveState *st)
{
    const Box &B = Lf.B;
    if (!B.threeD || !Lf.uniform || B.nf - 2 > 128 || !bc || !xout || !b) return hipErrorInvalidValue;
    int P = std::max(2, planesPerChunk);
    P += P & 1;
    const int tilesM = (B.nm - 2 + kPreRows - 1) / kPreRows;
    const int grid = tilesM * ((B.ns - 2 + P - 1) / P);
    const SweepRestrictArgs a{Lf, nullptr, xout, b, c1, Lc, bc, P, tilesM, st};
    sweep_pre_restrict<kPreRows>(s, grid, a);
    return hipGetLastError();
}

int64_t rbgs_restrict_blocks(const Box &B, int tileRows, int planesPerChunk)
{
    const int K = B.nf - 2 <= 128 ? 1 : 2;
    std::vector<int> tms;
    rbgs_restrict_rows(K, tms);
    if (tms.empty()) return 0;
    int TM = tms.front();
    for (int t : tms)
        if (t <= tileRows) TM = t;
    int P = std::max(2, planesPerChunk);
    P += P & 1;
    return (int64_t)std::max(1, (B.nm - 2 + TM - 1) / TM) * std::max(1, (B.ns - 2 + P - 1) / P);
}

hipError_t launch_prolong(hipStream_t s, const Level32 &Lf, float *x, const Level32 &Lc, const float *xc,
                          const SolveState *st)
{
    k_prolong<<<grid_of(Lf.B), kBlockThreads, 0, s>>>(Lf, x, Lc, xc, st);
    return hipGetLastError();
}

hipError_t launch_coarse_solve(hipStream_t s, int n, const int32_t *order, const double *M, const float *b, float *x,
                               const SolveState *st)
{
    if (n <= 0) return hipSuccess;
    if (n > kMaxDenseCoarse) return hipErrorInvalidValue;
    const int rowsPerBlock = kBlockThreads / 64, grid = (n + rowsPerBlock - 1) / rowsPerBlock;
    if (record(kb_coarse_solve, grid, kBlockThreads, CoarseArgs{n, order, M, b, x, st})) return hipSuccess;
    k_coarse_solve<<<grid, kBlockThreads, 0, s>>>(n, order, M, b, x, st);
    return hipGetLastError();
}

hipError_t launch_rz_rows_finalize(hipStream_t s, const Box &B, SolveState *st, double *partials)
{
    const int nrows = (B.ns - 2) * (B.nm - 2);
    const int nblk = std::min(kFuseBlocks, std::max(1, (nrows + kBlockThreads - 1) / kBlockThreads));
    const int seg = (nrows + nblk - 1) / nblk;
    // the block sums follow the row partials (the partials buffer holds at least twice the row count)
    const RzSumArgs a{st, partials, nrows, seg, partials + ((nrows + 1) & ~1)};
    if (record(kb_rz_rows_sum, nblk, kBlockThreads, a)) return hipSuccess;
    k_rz_rows_sum<<<nblk, kBlockThreads, 0, s>>>(a.st, a.rows, a.nrows, a.seg, a.blockSums);
    return hipGetLastError();
}

hipError_t launch_cg_rz(hipStream_t s, const Level32 &L, const float *r, const float *z, SolveState *st,
                        double *partials)
{
    const int G = grid_of(L.B);
    const Red red = red_of(partials, st, FIN_RZ, G);
    k_cg_rz<<<G, kBlockThreads, 0, s>>>(L, r, z, st, red);
    return finish(s, red, G, 1);
}

hipError_t launch_cg_direction(hipStream_t s, const Level32 &L, const float *z, float *p, SolveState *st)
{
    k_cg_direction<<<grid_of(L.B), kBlockThreads, 0, s>>>(L, z, p, st);
    return hipGetLastError();
}

hipError_t launch_cg_spmv(hipStream_t s, const Level32 &L, const float *p, float *q, SolveState *st,
                          double *partials)
{
    if (L.B.threeD)
    {
        const int P = march_planes(L.B), G = march_grid(L.B, P);
        const Red red = red_of(partials, st, FIN_PQ, G);
        k_cg_spmv_m<<<G, kBlockThreads, 0, s>>>(L, p, q, st, red, P);
        return finish(s, red, G, 1);
    }
    const int G = grid_of(L.B);
    const Red red = red_of(partials, st, FIN_PQ, G);
    k_cg_spmv<<<G, kBlockThreads, 0, s>>>(L, p, q, st, red);
    return finish(s, red, G, 1);
}

bool cg_fused_direction(const Level32 &L) { return L.B.threeD; }

hipError_t launch_cg_dirspmv(hipStream_t s, const Level32 &L, const float *z, float *pA, float *pB, float *q,
                             SolveState *st, double *partials)
{
    const int P = march_planes(L.B), G = march_grid(L.B, P);
    const Red red = red_of(partials, st, FIN_PQ, G);
    k_cg_dirspmv_t<<<G, kBlockThreads, 0, s>>>(L, z, pA, pB, q, st, red, P);
    return finish(s, red, G, 1);
}

hipError_t launch_cg_dir64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const float *z,
                           float *ring, int64_t stride, SolveState *st, double *partials)
{
    // 6-row blocks, two per CU (measured at 256^3: 57 us against 61 us for one 14-row block per CU)
    if (!B.threeD)
    {
        static const int resident = resident_grid(reinterpret_cast<const void *>(k_cg_dir64_g));
        const int G = std::min(B.blocks(), resident);
        const Red red = red_of(partials, st, FIN_PQ, G);
        k_cg_dir64_g<<<G, kBlockThreads, 0, s>>>(B, T, ty, z, ring, stride, st, red);
        return finish(s, red, G, 1);
    }
    const RowPlan rp = row_plan(B, T.c8, T.unitCoef, {{z, 8}, {ring, 8}}, 6, 2);
    if (rp.ok)
    {
        const Red red = red_of(partials, st, FIN_PQ, rp.G);
        const Dir64RowArgs a{B, T, z, ring, stride, st, red, rp.P, rp.tilesM};
        if (rp.K == 1)
            dir64_row<6, 1>(s, rp.G, 64 * 8, a);
        else
            dir64_row<6, 2>(s, rp.G, 64 * 8, a);
        return finish(s, red, rp.G, 1);
    }
    const int P = march_planes(B), G = march_grid(B, P);
    const Red red = red_of(partials, st, FIN_PQ, G);
    if (T.unitCoef)
        k_cg_dir64_t<true><<<G, kBlockThreads, 0, s>>>(B, T, ty, z, ring, stride, st, red, P);
    else
        k_cg_dir64_t<false><<<G, kBlockThreads, 0, s>>>(B, T, ty, z, ring, stride, st, red, P);
    return finish(s, red, G, 1);
}

hipError_t launch_cg_update64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const float *ring,
                              int64_t stride, double *x, double *r, float *r32, const double *b, SolveState *st,
                              double *partials)
{
    if (!B.threeD)
    {
        static const int resident = resident_grid(reinterpret_cast<const void *>(k_cg_update64_g));
        const int G = std::min(B.blocks(), resident);
        const Red red = red_of(partials, st, FIN_RR, G);
        k_cg_update64_g<<<G, kBlockThreads, 0, s>>>(B, T, ty, ring, stride, x, r, r32, b, st, red);
        return finish(s, red, G, 1);
    }
    const RowPlan rp = row_plan(B, T.c8, T.unitCoef, {{ring, 8}, {x, 16}, {r, 16}, {r32, 8}, {b, 16}}, 6, 2);
    if (rp.ok)
    {
        const Red red = red_of(partials, st, FIN_RR, rp.G);
        const Update64RowArgs a{B, T, ring, stride, x, r, r32, b, st, red, rp.P, rp.tilesM};
        if (rp.K == 1)
            update64_row<6, 1>(s, rp.G, 64 * 8, a);
        else
            update64_row<6, 2>(s, rp.G, 64 * 8, a);
        return finish(s, red, rp.G, 1);
    }
    const int P = march_planes(B), G = march_grid(B, P);
    const Red red = red_of(partials, st, FIN_RR, G);
    if (T.unitCoef)
        k_cg_update64_t<true><<<G, kBlockThreads, 0, s>>>(B, T, ty, ring, stride, x, r, r32, b, st, red, P);
    else
        k_cg_update64_t<false><<<G, kBlockThreads, 0, s>>>(B, T, ty, ring, stride, x, r, r32, b, st, red, P);
    return finish(s, red, G, 1);
}

hipError_t launch_flush_x(hipStream_t s, const Box &B, const Tables64 &T, const float *ring, int64_t stride, double *x,
                          const SolveState *st)
{
    // quads covering elements -1 .. cells - 1 (the allocations extend at least 15 elements past the box)
    const int64_t nq = (B.cells + 1 + 3) / 4;
    const int G = grid_1d(nq);
    k_flush_x<<<G, kBlockThreads, 0, s>>>(nq, T.c8, ring, stride, x, st);
    return hipGetLastError();
}

hipError_t launch_cg_update(hipStream_t s, const Level32 &L, float *e, float *r, const float *pA, const float *pB,
                            const float *q, SolveState *st, double *partials)
{
    auto quad = [](const float *p) { return ((uintptr_t)(p - 1) & 15u) == 0; };
    if (quad(e) && quad(r) && quad(pA) && quad(pB) && quad(q))
    {
        // quads covering elements -1 .. cells-1 (the allocation extends at least 15 elements past the box)
        const int64_t nq = (L.B.cells + 1 + 3) / 4;
        const int G = (int)std::min<int64_t>((nq + 2 * kBlockThreads - 1) / (2 * kBlockThreads), kMaxGrid);
        const Red red = red_of(partials, st, FIN_RR, G);
        k_cg_update_v<<<G, kBlockThreads, 0, s>>>(nq, reinterpret_cast<float4 *>(e - 1),
                                                   reinterpret_cast<float4 *>(r - 1),
                                                   reinterpret_cast<const float4 *>(pA - 1),
                                                   reinterpret_cast<const float4 *>(pB - 1),
                                                   reinterpret_cast<const float4 *>(q - 1), st, red);
        return finish(s, red, G, 1);
    }
    const int G = grid_of(L.B);
    const Red red = red_of(partials, st, FIN_RR, G);
    k_cg_update<<<G, kBlockThreads, 0, s>>>(L, e, r, pA, pB, q, st, red);
    return finish(s, red, G, 1);
}

hipError_t launch_inner_setup(hipStream_t s, SolveState *st, double reduction, int maxIters)
{
    k_inner_setup<<<1, 1, 0, s>>>(st, reduction, maxIters);
    return hipGetLastError();
}

hipError_t launch_state_init(hipStream_t s, SolveState *st, double thr2, int maxIters)
{
    k_state_init<<<1, 1, 0, s>>>(st, thr2, maxIters);
    return hipGetLastError();
}

hipError_t launch_krylov_init(hipStream_t s, SolveState *st, double tol, double ref, double aimFactor, int maxIters,
                              int residualIsB)
{
    k_krylov_init<<<1, 1, 0, s>>>(st, tol, ref, aimFactor, maxIters, residualIsB);
    return hipGetLastError();
}

}  // namespace nppe
