// The row walk of the MAC-field kernels (fields.hip) and of the allocation-free entry points (noalloc.hip): a
// wavefront walks one interior row of the box along the fast axis (k in 3D, j in 2D: the contiguous index of every
// MAC array), 64 cells per step, and a ballot over the row's stencil codes numbers the unknowns in IDP order as it
// goes (the row's first number comes from a per-row start count built at commit). Four rows per block; consecutive
// rows stay on one XCD.
#pragma once

#include <hip/hip_runtime.h>

#include "fields.h"

namespace nppe
{
namespace rows
{

constexpr int kRowsPerBlock = 4;   // one wavefront per row
constexpr int kThreads = 64 * kRowsPerBlock;
// Hardware workgroup ids go round-robin to the eight XCDs; give XCD x the x-th contiguous share of the logical
// blocks so that neighbouring rows are swept by one XCD.
__device__ __forceinline__ int xcd_block()
{
    const int b = blockIdx.x, nb = gridDim.x;
    const int xcd = b & 7, r = b >> 3, per = nb >> 3, extra = nb & 7;
    return xcd < extra ? xcd * (per + 1) + r : extra * (per + 1) + (xcd - extra) * per + r;
}

// The interior row a wave walks: (i, j) in 3D with cells along k, i in 2D with cells along j. Returns false for a
// wave past the last row (it still reaches every barrier of its block).
struct Row
{
    int i, j;          // 3D: the row's i, j; 2D: i (j unused)
    int nfast;         // interior cells along the row
    int64_t pitched;   // pitched box index of the row's cell 0
    bool active;
};

template <int DIM>
__device__ __forceinline__ Row interior_row(const FieldGeom &G, int r)
{
    Row R;
    R.active = r < G.rows;
    if (!R.active) r = 0;
    if (DIM == 3)
    {
        R.i = 1 + r / (G.Nx - 2);
        R.j = 1 + r % (G.Nx - 2);
        R.nfast = G.Nz - 2;
        R.pitched = G.B.at(R.i, R.j, 0);
    }
    else
    {
        R.i = 1 + r;
        R.j = 0;
        R.nfast = G.Nx - 2;
        R.pitched = G.B.at(0, R.i, 0);
    }
    return R;
}

// Walks the row in 64-cell chunks; for each unknown cell f (1-based along the fast axis) calls
// cell(f, id, code) with id its IDP number.
template <int DIM, class F>
__device__ __forceinline__ void walk_unknowns(const FieldGeom &G, const Row &R, int r, F &&cell)
{
    if (!R.active) return;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    int64_t base = G.rowStart[r];
    for (int f0 = 1; f0 <= R.nfast; f0 += 64)
    {
        const int f = f0 + lane;
        const unsigned code = f <= R.nfast ? G.c8[R.pitched + f] : 0u;
        const bool unk = (code & 64u) != 0u;
        const uint64_t m = __ballot(unk);
        if (unk) cell(f, base + __popcll(m & below), code);
        base += __popcll(m);
    }
}

// Sum of v over the block (256 threads); the result is valid on thread 0.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[kRowsPerBlock];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return threadIdx.x == 0 ? (red[0] + red[1]) + (red[2] + red[3]) : 0.0;
}

}  // namespace rows
}  // namespace nppe
