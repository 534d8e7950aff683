// Layout kernels for gfx950: a new cell layout of the committed box (FLIP's moving liquid) made into the multigrid
// hierarchy on the device -- every coarse level's types by the 2:1 coarsening rule, every level's stencil codes,
// the unknown counts that decide whether the hierarchy keeps its shape, and the fine level's IDP numbering. The
// coarsening rule is the host construction's (hierarchy.cpp coarsen), the codes and the numbering
// (PressureSolver3D.cpp:133-139) are built by these same kernels at commit, so a handle updated here solves exactly
// as one committed with the new layout.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "layout.h"

namespace nppe
{
namespace
{

constexpr uint8_t kFluid = 0, kSolid = 1, kEmpty = 2, kPressure = 4;

__device__ __forceinline__ int rank_of(uint8_t t)
{
    return (t == kEmpty || t == kPressure) ? 2 : (t == kFluid ? 1 : 0);
}

// Children of coarse index C along an axis with fine interior extent n (fine ring at 0 and n + 1); mc = coarse
// interior extent.
__device__ __forceinline__ void children(int C, int mc, int n, int &lo, int &hi)
{
    if (C == 0) { lo = hi = 0; return; }
    if (C == mc + 1) { lo = hi = n + 1; return; }
    lo = 2 * C - 1;
    hi = min(2 * C, n);
}

__global__ __launch_bounds__(256) void k_coarsen_types(Box Bf, const uint8_t *__restrict__ tf, Box Bc,
                                                       uint8_t *__restrict__ tc)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)Bc.ns * Bc.nm * Bc.nf;
    if (q >= total) return;
    const int F = (int)(q % Bc.nf), M = (int)((q / Bc.nf) % Bc.nm), S = (int)(q / ((int64_t)Bc.nf * Bc.nm));
    const int nsi = Bf.ns - 2, nmi = Bf.nm - 2, nfi = Bf.nf - 2;
    const int msc = (nsi + 1) / 2, mmc = (nmi + 1) / 2, mfc = (nfi + 1) / 2;
    int slo = 0, shi = 0, mlo, mhi, flo, fhi;
    if (Bf.threeD) children(S, msc, nsi, slo, shi);
    children(M, mmc, nmi, mlo, mhi);
    children(F, mfc, nfi, flo, fhi);
    int best = 0;
    for (int s = slo; s <= shi; s++)
        for (int m = mlo; m <= mhi; m++)
            for (int f = flo; f <= fhi; f++) best = max(best, rank_of(tf[Bf.at(s, m, f)]));
    tc[Bc.at(S, M, F)] = best == 2 ? kEmpty : (best == 1 ? kFluid : kSolid);
}

__global__ __launch_bounds__(256) void k_stencil_codes(Box B, const uint8_t *__restrict__ t, uint8_t *__restrict__ c8,
                                                       uint16_t *__restrict__ c16)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int si = B.threeD ? B.ns - 2 : 1, mi = B.nm - 2, fi = B.nf - 2;
    if (q >= (int64_t)si * mi * fi) return;
    const int f = 1 + (int)(q % fi), m = 1 + (int)((q / fi) % mi), s = (B.threeD ? 1 : 0) + (int)(q / ((int64_t)fi * mi));
    const int64_t c = B.at(s, m, f);
    unsigned code8 = 0u, code16 = 0u;
    if (t[c] == kFluid)
    {
        code8 = 1u << 6;
        code16 = 1u << 12;
        const int ds[6] = {1, -1, 0, 0, 0, 0}, dm[6] = {0, 0, 1, -1, 0, 0}, df[6] = {0, 0, 0, 0, 1, -1};
#pragma unroll
        for (int a = 0; a < 6; a++)
        {
            if (!B.threeD && a < 2) continue;
            const int S = s + ds[a], M = m + dm[a], F = f + df[a];
            const bool ring = (B.threeD && (S == 0 || S == B.ns - 1)) || M == 0 || M == B.nm - 1 || F == 0 ||
                              F == B.nf - 1;
            const uint8_t tn = t[B.at(S, M, F)];
            unsigned st = 0u;
            if (tn == kFluid) st = ring ? 2u : 1u;
            else if (tn == kEmpty || tn == kPressure) st = ring ? 3u : 2u;
            if (st == 1u) code8 |= 1u << a;
            if (st >= 2u) code8 |= 1u << 7;
            code16 |= st << (2 * a);
        }
    }
    c8[c] = (uint8_t)code8;
    c16[c] = (uint16_t)code16;
}

// One wave per row (s, m) of the box, ring rows included: the row's code descriptor (see launch_row_codes).
__global__ __launch_bounds__(256) void k_row_codes(Box B, const uint8_t *__restrict__ c8, uint32_t *__restrict__ rowc)
{
    const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int rows = B.ns * B.nm;
    if (row >= rows) return;
    const int s = (int)(row / B.nm), m = (int)(row % B.nm);
    const uint8_t *r = c8 + (size_t)B.at(s, m, 0);
    const int last = B.nf - 2;
    const unsigned mid = last >= 3 ? r[2] : 0u;
    bool same = true;
    for (int f = 2 + lane; f <= last - 1; f += 64) same = same && r[f] == mid;
    same = __all(same);
    if (lane == 0)
        rowc[(size_t)m * B.ns + s] = (same ? 0x80000000u : 0u) | (unsigned)r[1] | (mid << 8) | ((unsigned)r[last] << 16);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_count_codes(Box B, const uint8_t *__restrict__ c8,
                                                     unsigned long long *__restrict__ partials)
{
    __shared__ unsigned long long red[2][4];
    const int si = B.threeD ? B.ns - 2 : 1, mi = B.nm - 2, fi = B.nf - 2;
    const int64_t total = (int64_t)si * mi * fi;
    unsigned long long unk = 0, dir = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x)
    {
        const int f = 1 + (int)(q % fi), m = 1 + (int)((q / fi) % mi),
                  s = (B.threeD ? 1 : 0) + (int)(q / ((int64_t)fi * mi));
        const unsigned code = c8[B.at(s, m, f)];
        unk += (code >> 6) & 1u;
        dir += ((code >> 6) & (code >> 7)) & 1u;
    }
    unk = wave_sum(unk);
    dir = wave_sum(dir);
    if ((threadIdx.x & 63) == 0)
    {
        red[0][threadIdx.x >> 6] = unk;
        red[1][threadIdx.x >> 6] = dir;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        partials[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        partials[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

__global__ __launch_bounds__(1024) void k_finish_counts(const unsigned long long *__restrict__ partials, int blocks,
                                                        long long *__restrict__ out)
{
    __shared__ unsigned long long red[2][16];
    unsigned long long a = 0, b = 0;
    for (int q = threadIdx.x; q < blocks; q += 1024)
    {
        a += partials[2 * q];
        b += partials[2 * q + 1];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if ((threadIdx.x & 63) == 0)
    {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long sa = 0, sb = 0;
        for (int q = 0; q < 16; q++)
        {
            sa += red[0][q];
            sb += red[1][q];
        }
        out[0] = (long long)sa;
        out[1] = (long long)sb;
    }
}

// The interior row r of the fine level (3D rows (i, j) with cells along k, 2D rows i with cells along j).
__device__ __forceinline__ int64_t row_base(const FieldGeom &G, int r, int &nfast)
{
    if (G.dim == 3)
    {
        nfast = G.Nz - 2;
        return G.B.at(1 + r / (G.Nx - 2), 1 + r % (G.Nx - 2), 0);
    }
    nfast = G.Nx - 2;
    return G.B.at(0, 1 + r, 0);
}

// One wave per interior row: its unknowns into rowStart[r] (made a prefix by k_scan_rows).
__global__ __launch_bounds__(256) void k_row_counts(FieldGeom G, int32_t *__restrict__ rowStart)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= G.rows) return;
    const int lane = threadIdx.x & 63;
    int nfast;
    const int64_t base = row_base(G, r, nfast);
    int count = 0;
    for (int f0 = 1; f0 <= nfast; f0 += 64)
    {
        const int f = f0 + lane;
        const bool unk = f <= nfast && ((G.c8[base + f] >> 6) & 1u);
        count += __popcll(__ballot(unk));
    }
    if (lane == 0) rowStart[r] = count;
}

// Exclusive prefix sums of rowStart[0 .. rows) in place, in one block; *total = the sum.
__global__ __launch_bounds__(1024) void k_scan_rows(int32_t *__restrict__ rowStart, int rows,
                                                    long long *__restrict__ total)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int seg = (rows + 1023) / 1024, lo = min(rows, t * seg), hi = min(rows, lo + seg);
    long long s = 0;
    for (int q = lo; q < hi; q++) s += rowStart[q];
    part[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the 1024 segment sums
    for (int o = 1; o < 1024; o <<= 1)
    {
        const long long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - s;   // exclusive start of this segment
    for (int q = lo; q < hi; q++)
    {
        const int c = rowStart[q];
        rowStart[q] = (int32_t)run;
        run += c;
    }
    if (t == 1023) *total = part[1023];
}

// order[id] = pitched box index of unknown id, walking each interior row in IDP order.
__global__ __launch_bounds__(256) void k_write_order(FieldGeom G, int32_t *__restrict__ order)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= G.rows) return;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    int nfast;
    const int64_t base = row_base(G, r, nfast);
    int64_t id = G.rowStart[r];
    for (int f0 = 1; f0 <= nfast; f0 += 64)
    {
        const int f = f0 + lane;
        const bool unk = f <= nfast && ((G.c8[base + f] >> 6) & 1u);
        const uint64_t m = __ballot(unk);
        if (unk) order[id + __popcll(m & below)] = (int32_t)(base + f);
        id += __popcll(m);
    }
}

unsigned grid_for(int64_t n, int threads) { return (unsigned)std::max<int64_t>(1, (n + threads - 1) / threads); }

}  // namespace

hipError_t launch_coarsen_types(hipStream_t s, const Box &Bf, const uint8_t *tf, const Box &Bc, uint8_t *tc)
{
    k_coarsen_types<<<grid_for((int64_t)Bc.ns * Bc.nm * Bc.nf, 256), 256, 0, s>>>(Bf, tf, Bc, tc);
    return hipGetLastError();
}

hipError_t launch_stencil_codes(hipStream_t s, const Box &B, const uint8_t *t, uint8_t *c8, uint16_t *c16)
{
    const int64_t interior = (int64_t)(B.threeD ? B.ns - 2 : 1) * (B.nm - 2) * (B.nf - 2);
    k_stencil_codes<<<grid_for(interior, 256), 256, 0, s>>>(B, t, c8, c16);
    return hipGetLastError();
}

hipError_t launch_row_codes(hipStream_t s, const Box &B, const uint8_t *c8, uint32_t *rowc)
{
    k_row_codes<<<grid_for((int64_t)B.ns * B.nm * 64, 256), 256, 0, s>>>(B, c8, rowc);
    return hipGetLastError();
}

hipError_t launch_count_codes(hipStream_t s, const Box &B, const uint8_t *c8, unsigned long long *partials,
                              int *blocks)
{
    const int64_t interior = (int64_t)(B.threeD ? B.ns - 2 : 1) * (B.nm - 2) * (B.nf - 2);
    const int nb = (int)std::min<int64_t>(kMaxCountBlocks, grid_for(interior, 256));
    *blocks = nb;
    k_count_codes<<<nb, 256, 0, s>>>(B, c8, partials);
    return hipGetLastError();
}

hipError_t launch_finish_counts(hipStream_t s, const unsigned long long *partials, int blocks, long long *out)
{
    k_finish_counts<<<1, 1024, 0, s>>>(partials, blocks, out);
    return hipGetLastError();
}

hipError_t launch_number_rows(hipStream_t s, const FieldGeom &G, int32_t *rowStart, long long *total)
{
    if (G.rows <= 0) return hipSuccess;
    k_row_counts<<<grid_for(G.rows, 4), 256, 0, s>>>(G, rowStart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_scan_rows<<<1, 1024, 0, s>>>(rowStart, G.rows, total);
    return hipGetLastError();
}

hipError_t launch_write_order(hipStream_t s, const FieldGeom &G, int32_t *order)
{
    if (G.rows <= 0) return hipSuccess;
    k_write_order<<<grid_for(G.rows, 4), 256, 0, s>>>(G, order);
    return hipGetLastError();
}

}  // namespace nppe
