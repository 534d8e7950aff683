// The coarsest multigrid level's dense pseudo-inverse, built on the device (at commit and after a layout update,
// so an update reads nothing back): the level's operator B = -A in the sigma-scaled volume form of the FP32 levels,
// its null modes (the constant on every connected component without a Dirichlet face) shifted onto the diagonal
// block, B inverted by Gauss-Jordan elimination, the shift removed again and the sign flipped: M = A^+ on the
// unknowns, null modes projected out. k_coarse_solve multiplies by M in every V-cycle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "handle.h"
#include "kernels.h"

namespace nppe
{
namespace
{

constexpr int kThreads = 1024;
constexpr uint8_t kCFluid = 0, kCEmpty = 2, kCPressure = 4;

__device__ __forceinline__ double block_sum_1024(double v, double *sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kThreads / 64; q++) t += sh[q];
    return t;
}

// One workgroup. Numbers the level's unknowns (interior FLUID cells in index order: a block-wide scan over
// contiguous chunks of the interior), assembles B (row-major, n x n), finds the components by label propagation,
// shifts the singular ones (B += dmean / |c| on the component's block, dmean the mean diagonal) and records the
// shift to remove after the inversion (shift[r] = 1 / (dmean |c|) on a singular component, else 0).
// state[0] = the unknowns found (the caller checks it against the count it expects).
__global__ __launch_bounds__(kThreads) void k_coarse_assemble(Box B, const uint8_t *__restrict__ ty,
                                                              const double *__restrict__ wS,
                                                              const double *__restrict__ wM,
                                                              const double *__restrict__ wF, double sigma, int n,
                                                              int32_t *__restrict__ id, int32_t *__restrict__ order,
                                                              double *__restrict__ M, int32_t *__restrict__ lab,
                                                              int32_t *__restrict__ aux, double *__restrict__ shift,
                                                              int32_t *__restrict__ state)
{
#pragma clang fp contract(off)
    __shared__ int counts[kThreads];
    __shared__ int changed;
    __shared__ double sh[kThreads / 64];
    __shared__ double dmeanS;
    const int tid = threadIdx.x;
    const int s0 = B.threeD ? 1 : 0, s1 = B.threeD ? B.ns - 2 : 0;
    const int nm = B.nm - 2, nf = B.nf - 2;
    const int64_t total = (int64_t)(s1 - s0 + 1) * nm * nf;
    for (int64_t c = tid; c < B.cells; c += kThreads) id[c] = -1;
    // interior cells in index order, a contiguous chunk per thread
    const int64_t chunk = (total + kThreads - 1) / kThreads;
    const int64_t t0 = tid * chunk, t1 = t0 + chunk < total ? t0 + chunk : total;
    auto cell = [&](int64_t t) {
        const int s = s0 + (int)(t / ((int64_t)nm * nf));
        const int r = (int)(t % ((int64_t)nm * nf));
        return B.at(s, 1 + r / nf, 1 + r % nf);
    };
    int mine = 0;
    for (int64_t t = t0; t < t1; t++) mine += ty[cell(t)] == kCFluid ? 1 : 0;
    counts[tid] = mine;
    __syncthreads();
    if (tid == 0)
    {
        int run = 0;
        for (int q = 0; q < kThreads; q++)
        {
            const int v = counts[q];
            counts[q] = run;
            run += v;
        }
        state[0] = run;
    }
    __syncthreads();
    if (state[0] != n) return;   // every thread sees the same count
    {
        int k = counts[tid];
        for (int64_t t = t0; t < t1; t++)
        {
            const int64_t c = cell(t);
            if (ty[c] == kCFluid)
            {
                id[c] = k;
                order[k] = (int32_t)c;
                k++;
            }
        }
    }
    for (int64_t e = tid; e < (int64_t)n * n; e += kThreads) M[e] = 0.0;
    __syncthreads();
    int32_t *touch = aux, *size = aux + n, *anyTouch = aux + 2 * n;
    // rows: FLUID neighbour -> diagonal c (off-diagonal -c when an interior unknown, else a Dirichlet face);
    // EMPTY/PRESSURE -> diagonal c, or the ring coefficient on the domain ring
    for (int r = tid; r < n; r += kThreads)
    {
        const int64_t c = order[r];
        const int s = (int)(c / B.sS), m = (int)((c % B.sS) / B.sM), f = (int)(c % B.sM);
        const double aS = wM[m] * wF[f], aM = wS[s] * wF[f], aF = wS[s] * wM[m];
        auto gap = [](const double *w, int q) { return 0.5 * (w[q] + w[q + 1]); };
        double cq[6] = {0, 0, 0, 0, 0, 0}, ring[6] = {0, 0, 0, 0, 0, 0};
        if (B.threeD)
        {
            cq[0] = aS * sigma / gap(wS, s);
            cq[1] = aS * sigma / gap(wS, s - 1);
            ring[0] = ring[1] = aS * 2.0 * sigma / wS[s];
        }
        cq[2] = aM * sigma / gap(wM, m);
        cq[3] = aM * sigma / gap(wM, m - 1);
        cq[4] = aF * sigma / gap(wF, f);
        cq[5] = aF * sigma / gap(wF, f - 1);
        ring[2] = ring[3] = aM * 2.0 * sigma / wM[m];
        ring[4] = ring[5] = aF * 2.0 * sigma / wF[f];
        const int ds[6] = {1, -1, 0, 0, 0, 0}, dm[6] = {0, 0, 1, -1, 0, 0}, df[6] = {0, 0, 0, 0, 1, -1};
        double diag = 0.0;
        int tch = 0;
        for (int q = 0; q < 6; q++)
        {
            if (q < 2 && !B.threeD) continue;
            const int S = s + ds[q], Mm = m + dm[q], F = f + df[q];
            const bool onRing = (B.threeD && (S == 0 || S == B.ns - 1)) || Mm == 0 || Mm == B.nm - 1 || F == 0 ||
                                F == B.nf - 1;
            const int64_t nc = B.at(S, Mm, F);
            const uint8_t t = ty[nc];
            if (t == kCFluid)
            {
                diag += cq[q];
                const int col = id[nc];
                if (!onRing && col >= 0) M[(size_t)r * n + col] -= cq[q];
                else tch = 1;
            }
            else if (t == kCEmpty || t == kCPressure)
            {
                diag += onRing ? ring[q] : cq[q];
                tch = 1;
            }
        }
        M[(size_t)r * n + r] += diag;
        touch[r] = tch;
        lab[r] = r;
    }
    __syncthreads();
    // components: the smallest unknown index reachable through off-diagonal couplings
    for (;;)
    {
        if (tid == 0) changed = 0;
        __syncthreads();
        for (int r = tid; r < n; r += kThreads)
        {
            const int64_t c = order[r];
            const int s = (int)(c / B.sS), m = (int)((c % B.sS) / B.sM), f = (int)(c % B.sM);
            int l = lab[r];
            const int64_t nb[6] = {B.threeD ? B.at(s + 1, m, f) : -1, B.threeD ? B.at(s - 1, m, f) : -1,
                                   B.at(s, m + 1, f), B.at(s, m - 1, f), B.at(s, m, f + 1), B.at(s, m, f - 1)};
            for (int q = 0; q < 6; q++)
            {
                if (nb[q] < 0) continue;
                const int col = id[nb[q]];
                if (col >= 0 && M[(size_t)r * n + col] != 0.0)
                {
                    const int lc = lab[col];
                    if (lc < l) l = lc;
                }
            }
            if (l < lab[r])
            {
                lab[r] = l;
                changed = 1;
            }
        }
        __syncthreads();
        const int again = changed;
        __syncthreads();
        if (!again) break;
    }
    for (int r = tid; r < n; r += kThreads)
    {
        size[r] = 0;
        anyTouch[r] = 0;
    }
    __syncthreads();
    for (int r = tid; r < n; r += kThreads)
    {
        atomicAdd(&size[lab[r]], 1);
        if (touch[r]) atomicOr(&anyTouch[lab[r]], 1);
    }
    double d = 0.0;
    for (int r = tid; r < n; r += kThreads) d += M[(size_t)r * n + r];
    __syncthreads();
    const double dsum = block_sum_1024(d, sh);
    if (tid == 0)
    {
        double dm = n > 0 ? dsum / n : 1.0;
        dmeanS = dm > 0.0 ? dm : 1.0;
    }
    __syncthreads();
    const double dmean = dmeanS;
    for (int r = tid; r < n; r += kThreads)
    {
        const int L = lab[r];
        shift[r] = anyTouch[L] ? 0.0 : 1.0 / dmean / (double)size[L];
    }
    for (int64_t e = tid; e < (int64_t)n * n; e += kThreads)
    {
        const int a = (int)(e / n), b = (int)(e % n);
        const int L = lab[a];
        if (!anyTouch[L] && lab[b] == L) M[e] += dmean / (double)size[L];
    }
}

// One Gauss-Jordan step on pivot k from `in` into `out` (both n x n, row-major): the pivot row scaled by 1/p, every
// other row reduced by it, column k made the inverse's. A non-positive pivot (the shifted operator is not positive
// definite) sets state[1].
__global__ __launch_bounds__(256) void k_gauss_jordan(const double *__restrict__ in, double *__restrict__ out, int n,
                                                     int k, int32_t *__restrict__ state)
{
    const double p = in[(size_t)k * n + k];
    const double ip = 1.0 / p;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)n * n;
         e += (int64_t)gridDim.x * blockDim.x)
    {
        const int i = (int)(e / n), j = (int)(e % n);
        double v;
        if (i == k) v = j == k ? ip : in[e] * ip;
        else if (j == k) v = -in[e] * ip;
        else v = in[e] - in[(size_t)i * n + k] * (in[(size_t)k * n + j] * ip);
        out[e] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && !(p > 0.0)) state[1] = 1;
}

// M = -(sym(B^-1) - shift on each singular component's block).
__global__ __launch_bounds__(256) void k_coarse_finish(const double *__restrict__ inv, double *__restrict__ M, int n,
                                                      const int32_t *__restrict__ lab, const double *__restrict__ shift)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)n * n;
         e += (int64_t)gridDim.x * blockDim.x)
    {
        const int a = (int)(e / n), b = (int)(e % n);
        double v = 0.5 * (inv[e] + inv[(size_t)b * n + a]);
        if (lab[a] == lab[b]) v -= shift[a];
        M[e] = -v;
    }
}

}  // namespace

// The pseudo-inverse of the coarsest level (h->levels.back()) with n unknowns into h->dCoarseM / dCoarseOrder.
int coarse_pinv_device(nppe_solver *h, int n)
{
    DevLevel &C = h->levels.back();
    h->nCoarse = n;
    if (n == 0) return NPPE_OK;
    if (n > kMaxDenseCoarse) return fail(h, NPPE_ERR_INTERNAL, "coarsest level too large for the dense solve");
    HIP_TRY(h, hipMemsetAsync(h->dCoarseState, 0, 4 * sizeof(int32_t), h->stream), "coarse solve");
    const double *w = h->dCoarseW;
    const int ns = C.host.ns, nm = C.host.nm;
    k_coarse_assemble<<<1, kThreads, 0, h->stream>>>(C.box, C.types, w, w + ns, w + ns + nm, h->sigma, n,
                                                     h->dCoarseId, h->dCoarseOrder, h->dCoarseWork, h->dCoarseLab,
                                                     h->dCoarseAux, h->dCoarseShift, h->dCoarseState);
    HIP_TRY(h, hipGetLastError(), "coarse solve");
    const int64_t nn = (int64_t)n * n;
    const int G = (int)std::min<int64_t>(1024, (nn + 255) / 256);
    double *a = h->dCoarseWork, *b = h->dCoarseM;
    for (int k = 0; k < n; k++)
    {
        k_gauss_jordan<<<G, 256, 0, h->stream>>>(a, b, n, k, h->dCoarseState);
        std::swap(a, b);
    }
    HIP_TRY(h, hipGetLastError(), "coarse solve");
    // the inverse is in a; the result goes to dCoarseM
    double *inv = a;
    if (inv == h->dCoarseM)
    {
        HIP_TRY(h, hipMemcpyAsync(h->dCoarseWork, inv, (size_t)nn * sizeof(double), hipMemcpyDeviceToDevice, h->stream),
                "coarse solve");
        inv = h->dCoarseWork;
    }
    k_coarse_finish<<<G, 256, 0, h->stream>>>(inv, h->dCoarseM, n, h->dCoarseLab, h->dCoarseShift);
    HIP_TRY(h, hipGetLastError(), "coarse solve");
    int32_t st[4] = {0, 0, 0, 0};
    HIP_TRY(h, hipMemcpyAsync(st, h->dCoarseState, sizeof(st), hipMemcpyDeviceToHost, h->stream), "coarse solve");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "coarse solve");
    if (st[0] != n) return fail(h, NPPE_ERR_INTERNAL, "coarsest level: unknown count differs from the expected one");
    if (st[1]) return fail(h, NPPE_ERR_INTERNAL, "coarsest operator is not positive definite after the null-mode shift");
    return NPPE_OK;
}

}  // namespace nppe
