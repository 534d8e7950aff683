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
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace nppe
{
namespace
{

constexpr uint8_t kFluid = 0, kEmpty = 2, kPressure = 4;

__device__ __forceinline__ bool is_dirichlet(uint8_t t) { return t == kEmpty || t == kPressure; }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}


// Block sum of NV per-thread values, written by thread 0 to partials[a * nblk + blk] (no atomics: the sums over
// blocks are taken in index order by k_finalize, so every reduction is deterministic).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *partials, int nblk, int blk)
{
    constexpr int kWaves = kBlockThreads / 64;
    __shared__ double sh[NV][kWaves];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int a = 0; a < NV; a++)
    {
        const double x = wave_sum(v[a]);
        if (lane == 0) sh[a][w] = x;
    }
    __syncthreads();
    if (tid == 0)
    {
#pragma unroll
        for (int a = 0; a < NV; a++)
        {
            double x = 0.0;
#pragma unroll
            for (int q = 0; q < kWaves; q++) x += sh[a][q];
            partials[(size_t)a * nblk + blk] = x;
        }
    }
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
// coupled neighbour (coupled bit q set), and the diagonal (negative, or 0 for an isolated cell).
struct Op32
{
    float c[6];
    float diag;
    unsigned coupled;
};

__device__ __forceinline__ Op32 op32(const Box &B, const Tables32 &T, const uint8_t *ty, int s, int m, int f,
                                     int64_t idx)
{
    Op32 o;
    const float wS = T.wS[s], wM = T.wM[m], wF = T.wF[f];
    const float aS = wM * wF, aM = wS * wF, aF = wS * wM;
    float ring[6];
    o.c[0] = aS * T.gS[s];
    o.c[1] = aS * T.gS[s - (B.threeD ? 1 : 0)];
    o.c[2] = aM * T.gM[m];
    o.c[3] = aM * T.gM[m - 1];
    o.c[4] = aF * T.gF[f];
    o.c[5] = aF * T.gF[f - 1];
    ring[0] = ring[1] = aS * T.hS[s];
    ring[2] = ring[3] = aM * T.hM[m];
    ring[4] = ring[5] = aF * T.hF[f];
    const int64_t off[6] = {B.sS, -B.sS, B.sM, -B.sM, 1, -1};
    const bool onRing[6] = {s + 1 == B.ns - 1, s - 1 == 0, m + 1 == B.nm - 1, m - 1 == 0, f + 1 == B.nf - 1,
                            f - 1 == 0};
    float diag = 0.f;
    unsigned cp = 0;
#pragma unroll
    for (int q = 0; q < 6; q++)
    {
        if (q < 2 && !B.threeD) continue;
        const uint8_t t = ty[idx + off[q]];
        if (t == kFluid)
        {
            diag -= o.c[q];
            if (!onRing[q]) cp |= 1u << q;
        }
        else if (is_dirichlet(t))
            diag -= onRing[q] ? ring[q] : o.c[q];
    }
    o.diag = diag;
    o.coupled = cp;
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

// ---- FP64 operator with the reference's exact arithmetic ----------------------------------------------------
struct Op64
{
    double v[6];   // scale * c[q]: the off-diagonal values
    double dval;   // scale * diag (the diagonal entry; present iff diag != 0)
    double diag;
    unsigned coupled;
};

__device__ __forceinline__ Op64 op64(const Box &B, const Tables64 &T, const uint8_t *ty, int s, int m, int f,
                                     int64_t idx)
{
#pragma clang fp contract(off)
    Op64 o;
    double c[6];
    if (T.unitCoef)
    {
#pragma unroll
        for (int q = 0; q < 6; q++) c[q] = 1.0;
    }
    else if (B.threeD)
    {
        // FaceCoefs3D (PressureSolver3D.cpp:47-63): ay = Width(X,j)*Width(Z,k), ax = Width(Y,i)*Width(Z,k),
        // az = Width(X,j)*Width(Y,i); slow = Y, mid = X, fast = Z.
        const double ay = T.wM[m] * T.wF[f];
        const double ax = T.wS[s] * T.wF[f];
        const double az = T.wM[m] * T.wS[s];
        c[0] = ay / T.gS[s];
        c[1] = ay / T.gS[s - 1];
        c[2] = ax / T.gM[m];
        c[3] = ax / T.gM[m - 1];
        c[4] = az / T.gF[f];
        c[5] = az / T.gF[f - 1];
    }
    else
    {
        // PressureSolver2D.cpp:125-128; mid = Y (i), fast = X (j).
        c[0] = c[1] = 0.0;
        c[2] = T.wF[f] / T.gM[m];
        c[3] = T.wF[f] / T.gM[m - 1];
        c[4] = T.wM[m] / T.gF[f];
        c[5] = T.wM[m] / T.gF[f - 1];
    }
    const int64_t off[6] = {B.sS, -B.sS, B.sM, -B.sM, 1, -1};
    const bool onRing[6] = {s + 1 == B.ns - 1, s - 1 == 0, m + 1 == B.nm - 1, m - 1 == 0, f + 1 == B.nf - 1,
                            f - 1 == 0};
    uint8_t t[6];
#pragma unroll
    for (int q = 0; q < 6; q++) t[q] = (q < 2 && !B.threeD) ? (uint8_t)1 : ty[idx + off[q]];
    double diag = 0.0;
    unsigned cp = 0;
#pragma unroll
    for (int q = 0; q < 6; q++)
        if (t[q] == kFluid)
        {
            diag -= c[q];
            if (!onRing[q]) cp |= 1u << q;
        }
#pragma unroll
    for (int q = 0; q < 6; q++)
        if (is_dirichlet(t[q]))
        {
            double rf = 1.0;
            if (onRing[q])
            {
                // DirichletRingFactor (PressureSettings.h:307-310): FaceGap on the neighbour's side / half width.
                const double *g = q < 2 ? T.gS : (q < 4 ? T.gM : T.gF);
                const double *w = q < 2 ? T.wS : (q < 4 ? T.wM : T.wF);
                const int n = q < 2 ? s : (q < 4 ? m : f);
                rf = g[(q & 1) ? n - 1 : n] / (0.5 * w[n]);
            }
            diag -= c[q] * rf;
        }
#pragma unroll
    for (int q = 0; q < 6; q++) o.v[q] = T.scale * c[q];
    o.diag = diag;
    o.dval = T.scale * diag;
    o.coupled = cp;
    return o;
}

// Row q of A times x, summed in ascending column order from 0.0 as the CSR product does
// (Eigen SparseDenseProduct, row-major: tmp += value * x(col)). X(offset) fetches x at idx + offset.
template <class X>
__device__ __forceinline__ double row64(const Box &B, const Op64 &o, int64_t idx, X xat)
{
#pragma clang fp contract(off)
    double t = 0.0;
    if (o.coupled & 2u) t += o.v[1] * xat(-B.sS);
    if (o.coupled & 8u) t += o.v[3] * xat(-B.sM);
    if (o.coupled & 32u) t += o.v[5] * xat(-1);
    if (o.diag != 0.0) t += o.dval * xat(0);
    if (o.coupled & 16u) t += o.v[4] * xat(1);
    if (o.coupled & 4u) t += o.v[2] * xat(B.sM);
    if (o.coupled & 1u) t += o.v[0] * xat(B.sS);
    (void)idx;
    return t;
}

// ---- scalar finalisation -----------------------------------------------------------------------------------

// What k_finalize does with the block partials of the kernel before it.
enum FinOp
{
    FIN_SUM0 = 0,     // sums[0] = t0
    FIN_SUM1,         // sums[1] = t0
    FIN_SUM2,         // sums[2] = t0
    FIN_RZ,           // beta = t0 / rz (0 on the first direction); rz = t0
    FIN_PQ,           // alpha = rz / t0; first = 0
    FIN_RR,           // rr = t0; iters++; done when rr < thr2 (1) or at the cap (2)
    FIN_CG64_INIT,    // rz = t0; iters = 0
    FIN_CG64_PQ,      // alpha = rz / t0
    FIN_CG64_UPDATE   // rr = t0; converged -> done = 1 (not counted, as Eigen); else beta = t1 / rz, rz = t1, iters++
};

__device__ __forceinline__ bool fin_in_loop(int op) { return op >= FIN_RZ; }

// One block: sums the partials of each value over the nblk blocks in index order, then applies op.
__global__ __launch_bounds__(kBlockThreads) void k_finalize(SolveState *st, const double *__restrict__ partials,
                                                            int nblk, int nv, int op)
{
    if (fin_in_loop(op) && st->done) return;
    constexpr int kWaves = kBlockThreads / 64;
    __shared__ double sh[2][kWaves];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int a = 0; a < nv; a++)
    {
        double x = 0.0;
        for (int q = tid; q < nblk; q += kBlockThreads) x += partials[(size_t)a * nblk + q];
        x = wave_sum(x);
        if (lane == 0) sh[a][w] = x;
    }
    __syncthreads();
    if (tid != 0) return;
    double t[2] = {0.0, 0.0};
    for (int a = 0; a < nv; a++)
        for (int q = 0; q < kWaves; q++) t[a] += sh[a][q];
    switch (op)
    {
    case FIN_SUM0: st->sums[0] = t[0]; break;
    case FIN_SUM1: st->sums[1] = t[0]; break;
    case FIN_SUM2: st->sums[2] = t[0]; break;
    case FIN_RZ:
        st->beta = (st->first || st->rz == 0.0) ? 0.0 : t[0] / st->rz;
        st->rz = t[0];
        break;
    case FIN_PQ:
        st->alpha = t[0] != 0.0 ? st->rz / t[0] : 0.0;
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

// ---- FP64 kernels --------------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlockThreads) void k_scatter64(int64_t n, const int32_t *__restrict__ order,
                                                             const double *__restrict__ src, double *__restrict__ box,
                                                             double *partials, int withSum)
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
    block_sum<1>(a, partials, gridDim.x, blockIdx.x);
}

__global__ __launch_bounds__(kBlockThreads) void k_gather64(int64_t n, const int32_t *__restrict__ order,
                                                            const double *__restrict__ box, double *__restrict__ dst,
                                                            const SolveState *st, int meanSlot)
{
    const double mean = meanSlot >= 0 ? st->sums[meanSlot] / (double)n : 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlockThreads)
        dst[q] = box[order[q]] - mean;
}

__global__ __launch_bounds__(kBlockThreads) void k_sum64(int64_t n, const int32_t *__restrict__ order,
                                                         const double *__restrict__ box, double *partials)
{
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlockThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlockThreads)
        acc += box[order[q]];
    double a[1] = {acc};
    block_sum<1>(a, partials, gridDim.x, blockIdx.x);
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
                                                                     double *partials)
{
#pragma clang fp contract(off)
    const TileRange tr = my_tiles(B.blocks());
    double rr = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        double r = 0.0;
        if (ty[idx] == kFluid)
        {
            const Op64 o = op64(B, T, ty, s, m, f, idx);
            double ax;
            if (e)
            {
                ax = row64(B, o, idx, [&](int64_t d) { return x[idx + d] + (double)e[idx + d]; });
                xout[idx] = x[idx] + (double)e[idx];
            }
            else
            {
                ax = row64(B, o, idx, [&](int64_t d) { return x[idx + d]; });
            }
            r = b[idx] - ax;
        }
        if (r32) r32[idx] = (float)r;
        if (r64) r64[idx] = r;
        rr += r * r;
    }
    double a[1] = {rr};
    block_sum<1>(a, partials, gridDim.x, tr.lb);
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
        if (ty[idx] == kFluid)
        {
            const Op64 o = op64(B, T, ty, s, m, f, idx);
            v = row64(B, o, idx, [&](int64_t d) { return x[idx + d]; });
        }
        y[idx] = v;
    }
}

// Jacobi: Eigen's DiagonalPreconditioner takes 1/value of the diagonal entry, or 1 where the row has none.
__device__ __forceinline__ double inv_diag64(const Op64 &o) { return o.diag != 0.0 ? 1.0 / o.dval : 1.0; }

__global__ __launch_bounds__(kBlockThreads) void k_cg64_init(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                             const double *__restrict__ r, double *__restrict__ p,
                                                             double *partials)
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
            const Op64 o = op64(B, T, ty, s, m, f, idx);
            pv = inv_diag64(o) * r[idx];
            rz += r[idx] * pv;
        }
        p[idx] = pv;
    }
    double a[1] = {rz};
    block_sum<1>(a, partials, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg64_spmv(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                             const double *__restrict__ p, double *__restrict__ q,
                                                             const SolveState *st, double *partials)
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
            const Op64 o = op64(B, T, ty, s, m, f, idx);
            v = row64(B, o, idx, [&](int64_t d) { return p[idx + d]; });
            pq += p[idx] * v;
        }
        q[idx] = v;
    }
    double a[1] = {pq};
    block_sum<1>(a, partials, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg64_update(Box B, Tables64 T, const uint8_t *__restrict__ ty,
                                                               double *__restrict__ x, double *__restrict__ r,
                                                               const double *__restrict__ p,
                                                               const double *__restrict__ q, const SolveState *st,
                                                               double *partials)
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
        const Op64 o = op64(B, T, ty, s, m, f, idx);
        rr += rv * rv;
        rz += rv * (inv_diag64(o) * rv);
    }
    double a[2] = {rr, rz};
    block_sum<2>(a, partials, gridDim.x, tr.lb);
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
        const Op64 o = op64(B, T, ty, s, m, f, idx);
        p[idx] = inv_diag64(o) * r[idx] + beta * p[idx];
    }
}

// ---- FP32 multigrid kernels ----------------------------------------------------------------------------------

// One colour of a red-black Gauss-Seidel sweep: x = (b - sum_coupled c x_nb) / diag on cells with
// (s + m + f) % 2 == color. zero: the sweep starts from x = 0 (the neighbours are not read).
__global__ __launch_bounds__(kBlockThreads) void k_smooth(Box B, Tables32 T, const uint8_t *__restrict__ ty,
                                                          float *__restrict__ x, const float *__restrict__ b,
                                                          int color, int zero, const SolveState *st)
{
    if (st->done) return;
    const TileRange tr = my_tiles(B.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        if (((s + m + f) & 1) != color) continue;
        const int64_t idx = B.at(s, m, f);
        if (ty[idx] != kFluid) continue;
        const Op32 o = op32(B, T, ty, s, m, f, idx);
        const float acc = zero ? 0.f : offsum32(B, o, x, idx);
        x[idx] = o.diag != 0.f ? (b[idx] - acc) / o.diag : 0.f;
    }
}

// Coarse cell (S, M, F) gathers the residual b - A x of its fine children (fine indices 2C-1, 2C along each
// coarsened axis, clipped to the fine interior): bc = sum (R = P^T for the injection prolongation).
__global__ __launch_bounds__(kBlockThreads) void k_residual_restrict(Box Bf, Tables32 Tf,
                                                                     const uint8_t *__restrict__ tyf,
                                                                     const float *__restrict__ x,
                                                                     const float *__restrict__ b, Box Bc,
                                                                     float *__restrict__ bc, const SolveState *st)
{
    if (st->done) return;
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
                    if (tyf[idx] != kFluid) continue;
                    const Op32 o = op32(Bf, Tf, tyf, s, m, f, idx);
                    sum += b[idx] - (offsum32(Bf, o, x, idx) + o.diag * x[idx]);
                }
        bc[Bc.at(S, M, F)] = sum;
    }
}

// x += P xc: every fine FLUID cell takes its parent's correction when the parent is a coarse unknown.
__global__ __launch_bounds__(kBlockThreads) void k_prolong(Box Bf, const uint8_t *__restrict__ tyf,
                                                           float *__restrict__ x, Box Bc,
                                                           const uint8_t *__restrict__ tyc,
                                                           const float *__restrict__ xc, const SolveState *st)
{
    if (st->done) return;
    const TileRange tr = my_tiles(Bf.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(Bf, t, s, m, f)) continue;
        const int64_t idx = Bf.at(s, m, f);
        if (tyf[idx] != kFluid) continue;
        const int64_t pidx = Bc.at(Bf.threeD ? (s + 1) >> 1 : 0, (m + 1) >> 1, (f + 1) >> 1);
        const float v = xc[pidx];
        x[idx] += tyc[pidx] == kFluid ? v : 0.f;
    }
}

// Coarsest level: dense FP64 pseudo-inverse times the gathered right-hand side. One wavefront per row.
__global__ __launch_bounds__(kBlockThreads) void k_coarse_solve(int n, const int32_t *__restrict__ order,
                                                                const double *__restrict__ M,
                                                                const float *__restrict__ b, float *__restrict__ x,
                                                                const SolveState *st)
{
    if (st->done) return;
    __shared__ double rb[kMaxDenseCoarse];
    for (int q = threadIdx.x; q < n; q += kBlockThreads) rb[q] = (double)b[order[q]];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (kBlockThreads / 64) + (threadIdx.x >> 6);
    if (row >= n) return;
    const double *Mr = M + (size_t)row * n;
    double acc = 0.0;
    for (int q = lane; q < n; q += 64) acc += Mr[q] * rb[q];
    acc = wave_sum(acc);
    if (lane == 0) x[order[row]] = (float)acc;
}

// ---- FP32 CG kernels -----------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlockThreads) void k_cg_rz(Box B, const uint8_t *__restrict__ ty,
                                                         const float *__restrict__ r, const float *__restrict__ z,
                                                         const SolveState *st, double *partials)
{
    if (st->done) return;
    const TileRange tr = my_tiles(B.blocks());
    double v = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (ty[idx] == kFluid) v += (double)r[idx] * (double)z[idx];
    }
    double a[1] = {v};
    block_sum<1>(a, partials, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg_direction(Box B, const uint8_t *__restrict__ ty,
                                                                const float *__restrict__ z, float *__restrict__ p,
                                                                const SolveState *st)
{
    if (st->done) return;
    const int first = st->first;
    const float beta = (float)st->beta;
    const TileRange tr = my_tiles(B.blocks());
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (ty[idx] == kFluid) p[idx] = first ? z[idx] : z[idx] + beta * p[idx];
    }
}

__global__ __launch_bounds__(kBlockThreads) void k_cg_spmv(Box B, Tables32 T, const uint8_t *__restrict__ ty,
                                                           const float *__restrict__ p, float *__restrict__ q,
                                                           const SolveState *st, double *partials)
{
    if (st->done) return;
    const TileRange tr = my_tiles(B.blocks());
    double pq = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (ty[idx] != kFluid) continue;
        const Op32 o = op32(B, T, ty, s, m, f, idx);
        const float pc = p[idx];
        const float v = offsum32(B, o, p, idx) + o.diag * pc;
        q[idx] = v;
        pq += (double)pc * (double)v;
    }
    double a[1] = {pq};
    block_sum<1>(a, partials, gridDim.x, tr.lb);
}

__global__ __launch_bounds__(kBlockThreads) void k_cg_update(Box B, const uint8_t *__restrict__ ty,
                                                             float *__restrict__ e, float *__restrict__ r,
                                                             const float *__restrict__ p, const float *__restrict__ q,
                                                             const SolveState *st, double *partials)
{
    if (st->done) return;
    const int firstIter = st->iters == 0;
    const float alpha = (float)st->alpha;
    const TileRange tr = my_tiles(B.blocks());
    double rr = 0.0;
    for (int t = tr.begin; t < tr.end; t++)
    {
        int s, m, f;
        if (!tile_cell(B, t, s, m, f)) continue;
        const int64_t idx = B.at(s, m, f);
        if (ty[idx] != kFluid) continue;
        const float pv = p[idx];
        e[idx] = firstIter ? alpha * pv : e[idx] + alpha * pv;
        const float rv = r[idx] - alpha * q[idx];
        r[idx] = rv;
        rr += (double)rv * (double)rv;
    }
    double a[1] = {rr};
    block_sum<1>(a, partials, gridDim.x, tr.lb);
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
}

inline int grid_of(const Box &B) { return std::min(B.blocks(), kMaxGrid); }
inline int grid_1d(int64_t n) { return (int)std::min<int64_t>((n + kBlockThreads - 1) / kBlockThreads, kMaxGrid); }

hipError_t finalize(hipStream_t s, SolveState *st, const double *partials, int nblk, int nv, FinOp op)
{
    k_finalize<<<1, kBlockThreads, 0, s>>>(st, partials, nblk, nv, (int)op);
    return hipGetLastError();
}

#define LAUNCH_OK()                                    \
    do                                                 \
    {                                                  \
        hipError_t e_ = hipGetLastError();             \
        if (e_ != hipSuccess) return e_;               \
    } while (0)

}  // namespace

hipError_t launch_scatter64(hipStream_t s, int64_t n, const int32_t *order, const double *src, double *box,
                            SolveState *st, int sumSlot, double *partials)
{
    if (n <= 0) return hipSuccess;
    const int G = grid_1d(n);
    k_scatter64<<<G, kBlockThreads, 0, s>>>(n, order, src, box, partials, sumSlot >= 0);
    LAUNCH_OK();
    if (sumSlot < 0) return hipSuccess;
    return finalize(s, st, partials, G, 1, sumSlot == 0 ? FIN_SUM0 : (sumSlot == 1 ? FIN_SUM1 : FIN_SUM2));
}

hipError_t launch_gather64(hipStream_t s, int64_t n, const int32_t *order, const double *box, double *dst,
                           const SolveState *st, int meanSlot)
{
    if (n <= 0) return hipSuccess;
    k_gather64<<<grid_1d(n), kBlockThreads, 0, s>>>(n, order, box, dst, st, meanSlot);
    return hipGetLastError();
}

hipError_t launch_sum64(hipStream_t s, int64_t n, const int32_t *order, const double *box, SolveState *st, int slot,
                        double *partials)
{
    if (n <= 0) return hipSuccess;
    const int G = grid_1d(n);
    k_sum64<<<G, kBlockThreads, 0, s>>>(n, order, box, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 1, slot == 0 ? FIN_SUM0 : (slot == 1 ? FIN_SUM1 : FIN_SUM2));
}

hipError_t launch_update_residual64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty,
                                    const double *x, const float *e, double *xout, const double *b, float *r32,
                                    double *r64, SolveState *st, double *partials)
{
    const int G = grid_of(B);
    k_update_residual64<<<G, kBlockThreads, 0, s>>>(B, T, ty, x, e, xout, b, r32, r64, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 1, FIN_SUM1);
}

hipError_t launch_apply64(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *x,
                          double *y)
{
    k_apply64<<<grid_of(B), kBlockThreads, 0, s>>>(B, T, ty, x, y);
    return hipGetLastError();
}

hipError_t launch_cg64_init(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *r,
                            double *p, SolveState *st, double *partials)
{
    const int G = grid_of(B);
    k_cg64_init<<<G, kBlockThreads, 0, s>>>(B, T, ty, r, p, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 1, FIN_CG64_INIT);
}

hipError_t launch_cg64_spmv(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, const double *p,
                            double *q, SolveState *st, double *partials)
{
    const int G = grid_of(B);
    k_cg64_spmv<<<G, kBlockThreads, 0, s>>>(B, T, ty, p, q, st, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 1, FIN_CG64_PQ);
}

hipError_t launch_cg64_update(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty, double *x,
                              double *r, const double *p, const double *q, SolveState *st, double *partials)
{
    const int G = grid_of(B);
    k_cg64_update<<<G, kBlockThreads, 0, s>>>(B, T, ty, x, r, p, q, st, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 2, FIN_CG64_UPDATE);
}

hipError_t launch_cg64_direction(hipStream_t s, const Box &B, const Tables64 &T, const uint8_t *ty,
                                 const double *r, double *p, SolveState *st)
{
    k_cg64_direction<<<grid_of(B), kBlockThreads, 0, s>>>(B, T, ty, r, p, st);
    return hipGetLastError();
}

hipError_t launch_smooth(hipStream_t s, const Box &B, const Tables32 &T, const uint8_t *ty, float *x, const float *b,
                         int color, int zero, const SolveState *st)
{
    k_smooth<<<grid_of(B), kBlockThreads, 0, s>>>(B, T, ty, x, b, color, zero, st);
    return hipGetLastError();
}

hipError_t launch_residual_restrict(hipStream_t s, const Box &Bf, const Tables32 &Tf, const uint8_t *tyf,
                                    const float *x, const float *b, const Box &Bc, float *bc, const SolveState *st)
{
    k_residual_restrict<<<grid_of(Bc), kBlockThreads, 0, s>>>(Bf, Tf, tyf, x, b, Bc, bc, st);
    return hipGetLastError();
}

hipError_t launch_prolong(hipStream_t s, const Box &Bf, const uint8_t *tyf, float *x, const Box &Bc,
                          const uint8_t *tyc, const float *xc, const SolveState *st)
{
    k_prolong<<<grid_of(Bf), kBlockThreads, 0, s>>>(Bf, tyf, x, Bc, tyc, xc, st);
    return hipGetLastError();
}

hipError_t launch_coarse_solve(hipStream_t s, int n, const int32_t *order, const double *M, const float *b, float *x,
                               const SolveState *st)
{
    if (n <= 0) return hipSuccess;
    if (n > kMaxDenseCoarse) return hipErrorInvalidValue;
    const int rowsPerBlock = kBlockThreads / 64;
    k_coarse_solve<<<(n + rowsPerBlock - 1) / rowsPerBlock, kBlockThreads, 0, s>>>(n, order, M, b, x, st);
    return hipGetLastError();
}

hipError_t launch_cg_rz(hipStream_t s, const Box &B, const uint8_t *ty, const float *r, const float *z, SolveState *st,
                        double *partials)
{
    const int G = grid_of(B);
    k_cg_rz<<<G, kBlockThreads, 0, s>>>(B, ty, r, z, st, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 1, FIN_RZ);
}

hipError_t launch_cg_direction(hipStream_t s, const Box &B, const uint8_t *ty, const float *z, float *p,
                               SolveState *st)
{
    k_cg_direction<<<grid_of(B), kBlockThreads, 0, s>>>(B, ty, z, p, st);
    return hipGetLastError();
}

hipError_t launch_cg_spmv(hipStream_t s, const Box &B, const Tables32 &T, const uint8_t *ty, const float *p, float *q,
                          SolveState *st, double *partials)
{
    const int G = grid_of(B);
    k_cg_spmv<<<G, kBlockThreads, 0, s>>>(B, T, ty, p, q, st, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 1, FIN_PQ);
}

hipError_t launch_cg_update(hipStream_t s, const Box &B, const uint8_t *ty, float *e, float *r, const float *p,
                            const float *q, SolveState *st, double *partials)
{
    const int G = grid_of(B);
    k_cg_update<<<G, kBlockThreads, 0, s>>>(B, ty, e, r, p, q, st, partials);
    LAUNCH_OK();
    return finalize(s, st, partials, G, 1, FIN_RR);
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

}  // namespace nppe
