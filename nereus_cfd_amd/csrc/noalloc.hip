// Entry points that allocate nothing, for callers that cannot allocate device memory (code under stream capture,
// tight serving loops): their scratch (block partials and the final sum) lives in a file-scope device array of a
// fixed size instead of buffers the library allocates. Host code never passes the array's name to a kernel (in host
// code it is a stand-in, not a device address): the address comes from hipGetSymbolAddress, looked up once per
// device with that device current (at nppe_create, else at the entry point's first call on the device) and kept
// per device. Every call on one device shares that device's array, so calls that overlap on one GPU race; each
// call writes every entry it reads before reading it, so nothing is cleared between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/nppe.h"
#include "fields.h"
#include "handle.h"
#include "op64.h"
#include "rowwalk.h"

namespace nppe
{
namespace
{

// Block partials of one row-walking reduction (four interior rows per block: boxes of up to 262,144 interior rows,
// e.g. 512 x 512 rows of a 3D grid), then the slot of the final sum. 512 KiB + 64 B per device.
constexpr int kScratchPartials = 65536;
constexpr int kScratchSum = kScratchPartials;
static __device__ __align__(256) double scratch[kScratchPartials + 8];

constexpr int kMaxDevices = 64;
std::mutex gScratchMu;
double *gScratch[kMaxDevices] = {};

// ||b + A p^n||^2 over the unknowns in one pass: the row walk numbers the unknowns, A p^n is the reference's row
// (apply64) with p^n read straight from the caller's MAC array at the cell and its coupled neighbours (the
// reference's p^n vector is zero off the unknowns, and a row only reads unknowns), so no box copy of p^n is
// needed. The same rows per block, products and sums in the same order as the gather + apply64 + k_norm2_sum
// path of nppe_increment_reference_norm_device: the partials, and the norm, are bit-identical to it.
template <int DIM>
__global__ __launch_bounds__(rows::kThreads) void k_refnorm_fused(FieldGeom G, Tables64 T,
                                                                  const uint8_t *__restrict__ ty,
                                                                  const double *__restrict__ b,
                                                                  const double *__restrict__ pOld,
                                                                  double *__restrict__ partials)
{
#pragma clang fp contract(off)
    const int lb = rows::xcd_block();
    const int r = lb * rows::kRowsPerBlock + (threadIdx.x >> 6);
    const rows::Row R = rows::interior_row<DIM>(G, r);
    const int64_t sM = G.B.sM, sS = G.B.sS;
    // the MAC array's strides along the mid and slow axes (3D: j -> Nz, i -> Nx Nz; 2D: i -> Nx)
    const int64_t macM = DIM == 3 ? (int64_t)G.Nz : (int64_t)G.Nx, macS = (int64_t)G.Nx * G.Nz;
    double acc = 0.0;
    rows::walk_unknowns<DIM>(G, R, r, [&](int f, int64_t id, unsigned) {
#pragma clang fp contract(off)
        const int64_t c = DIM == 3 ? ((int64_t)R.i * G.Nx + R.j) * G.Nz + f : (int64_t)R.i * G.Nx + f;
        const int s = DIM == 3 ? R.i : 0, m = DIM == 3 ? R.j : R.i;
        auto xat = [&](int64_t d) {
            const int64_t e = d == 0 ? 0 : (d == 1 ? 1 : (d == -1 ? -1 : (d == sM ? macM : (d == -sM ? -macM
                              : (d == sS ? macS : -macS)))));
            return pOld[c + e];
        };
        const double y = apply64(G.B, T, ty, s, m, f, R.pitched + f, xat);
        const double t = b[id] + y;
        acc += t * t;
    });
    const double sum = rows::block_sum(acc);
    if (threadIdx.x == 0) partials[lb] = sum;
}

}  // namespace

// The scratch array's address on the current device (looked up once per device).
int noalloc_scratch(int device, double **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return NPPE_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lock(gScratchMu);
    if (!gScratch[device])
    {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NPPE_ERR_DEVICE;
        gScratch[device] = static_cast<double *>(p);
    }
    *out = gScratch[device];
    return NPPE_OK;
}

}  // namespace nppe

using namespace nppe;

extern "C" int nppe_increment_reference_norm_noalloc(nppe_solver *h, const double *d_b, const double *d_p_old,
                                                     double *norm)
{
    if (!h || !norm) return NPPE_ERR_ARGUMENT;
    if (!h->committed) return fail(h, NPPE_ERR_NO_GRID, "no grid committed");
    if (h->n > 0 && (!d_b || !d_p_old)) return fail(h, NPPE_ERR_ARGUMENT, "null b or p_old");
    HIP_TRY(h, hipSetDevice(h->device), "set device");
    h->err.clear();
    *norm = 0.0;
    if (h->n == 0) return NPPE_OK;
    const int nb = field_row_blocks(h->fg);
    if (nb > kScratchPartials)
        return fail(h, NPPE_ERR_SIZE, "increment reference norm without allocation: more than 262144 interior rows");
    double *ptr = nullptr;
    if (int rc = noalloc_scratch(h->device, &ptr)) return fail(h, rc, "scratch address");
    const uint8_t *ty = h->levels[0].types;
    if (h->dim == 3)
        k_refnorm_fused<3><<<nb, rows::kThreads, 0, h->stream>>>(h->fg, h->t64, ty, d_b, d_p_old, ptr);
    else
        k_refnorm_fused<2><<<nb, rows::kThreads, 0, h->stream>>>(h->fg, h->t64, ty, d_b, d_p_old, ptr);
    HIP_TRY(h, hipGetLastError(), "reference norm");
    HIP_TRY(h, launch_sum_partials(h->stream, ptr, nb, ptr + kScratchSum), "reference norm");
    HIP_TRY(h, hipMemcpyAsync(h->hScal, ptr + kScratchSum, sizeof(double), hipMemcpyDeviceToHost, h->stream),
            "reference norm");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "reference norm");
    *norm = std::sqrt(h->hScal[0]);
    return NPPE_OK;
}
