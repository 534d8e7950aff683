// The solver handle (struct nppe_solver of include/nppe.h) and the host helpers every translation unit of the
// library that works on a handle shares: error reporting, budgeted device allocation and stream-ordered uploads.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nppe.h"
#include "../../include/nppe_profile.h"
#include "fields.h"
#include "hierarchy.h"
#include "nppe_internal.h"

namespace nppe
{

struct BatchPlan;

// One multigrid level on the device.
struct DevLevel
{
    HostLevel host;
    Box box{};
    uint8_t *types = nullptr;
    uint8_t *c8 = nullptr;
    uint16_t *c16 = nullptr;
    uint32_t *rowc = nullptr;   // per-row code descriptors (launch_row_codes), ns * nm
    float *x[2] = {nullptr, nullptr};   // ping-pong pair of the smoother (level 0 uses the solver's z pair)
    float *b = nullptr;
    float *tab = nullptr;
    Tables32 t32{};
    Level32 v{};
    int tileRows = 6;        // smoother tile: tileRows rows per block (3D: 128 cells wide), planesPerChunk planes
    int planesPerChunk = 8;
    int proPlanes = 8;       // planes per chunk of the sweep that folds the coarse correction in
    int post2Planes = 8;     // planes per chunk of the post-smoothing pair in one march (launch_rbgs2)
    int pre2Planes = 8;      // planes per chunk of the pre-smoothing and restriction in one march
    int fuseRows = 12;       // tile rows and planes per chunk of the last pre-smoothing sweep with the restriction
    int fusePlanes = 8;      // folded in (launch_rbgs_restrict)
    double cellsBox = 0.0;   // ns * nm * nf (the box, unpitched)
};

}  // namespace nppe

struct nppe_solver
{
    nppe_options opt{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    bool committed = false;
    int dim = 3, Nx = 0, Ny = 0, Nz = 0;
    int64_t n = 0;
    bool singular = false;
    bool nonUniform = false;
    double dh = 0.0;

    std::vector<int32_t> orderHost;   // unpitched box index of each unknown
    int32_t *dOrder = nullptr;        // pitched box index of each unknown
    int64_t unknownCap = 0;           // entries of dOrder and the IDP staging vectors: the box's interior cells
    uint8_t *dLayoutStage = nullptr;  // nppe_update_layout: the new cell types (MAC order)
    long long *dLayoutInfo = nullptr; // its unknown counts per level (and Dirichlet-face counts), then the total
    bool chainStopsHere = false;      // coarsening stopped at the last level for a reason the layout cannot change
    bool hostStale = false;           // the host mirror of the layout (levels' types, orderHost) predates an update
    double sigma = 0.0;               // A-units of the volume-form coefficients (see build)
    std::vector<nppe::DevLevel> levels;
    double *dTab64 = nullptr;
    nppe::Tables64 t64{};
    double *dX64[2] = {nullptr, nullptr};
    double *dB64 = nullptr;
    float *dE = nullptr, *dR = nullptr, *dP = nullptr, *dP2 = nullptr, *dQ = nullptr, *dZ = nullptr, *dZ2 = nullptr;
    double *dR64 = nullptr, *dP64 = nullptr, *dQ64 = nullptr, *dP64b = nullptr;
    float *dPRing = nullptr;   // FP64 Krylov: the ring of kXRing FP32 directions, slot j at dPRing + j * ringStride
    int64_t ringStride = 0;
    bool krylov64 = false;   // MGPCG with the FP64 Krylov recursion (nppe_options::krylov)
    int nuPre = 2, nuPost = 2;   // sweeps before / after the coarse correction (nppe_options, auto resolved)
    double *dCompB = nullptr, *dCompX = nullptr, *dCompG = nullptr;
    double *dCoarseM = nullptr;
    int32_t *dCoarseOrder = nullptr;
    // building the coarsest pseudo-inverse on the device (coarse.hip): Gauss-Jordan's second matrix, the level's
    // unknown map over its box, FP64 widths, component labels, counts, shifts, status
    double *dCoarseWork = nullptr;
    int32_t *dCoarseId = nullptr;
    double *dCoarseW = nullptr;
    int32_t *dCoarseLab = nullptr;
    int32_t *dCoarseAux = nullptr;
    double *dCoarseShift = nullptr;
    int32_t *dCoarseState = nullptr;
    int nCoarse = 0;
    bool coarseDense = false;
    nppe::SolveState *dState = nullptr;
    nppe::SolveState *hState = nullptr;   // pinned
    double *hScal = nullptr;              // pinned: scalars read back by the allocation-free entry points
    double *dPartials = nullptr;
    int64_t partialsCap = 0;
    uint64_t bytesHeld = 0;
    bool budgetHit = false;
    int coarseNu = 0;   // sweeps on levels >= 2 (0: the same as the fine levels)
    bool useGraphs = true;                 // replay the inner iterations from a captured hipGraph
    bool tileTiming = true;                // time the sweep tile candidates at commit (see choose_tiles)
    hipGraphExec_t innerGraph = nullptr;   // one inner iteration, captured at the first solve
    hipGraphExec_t chunkGraph = nullptr;   // chunkGraphIters inner iterations back to back (a solve's first chunk)
    int chunkGraphIters = 0;
    double graphBytesPerIter = 0.0;
    uint64_t graphEpoch = 0;              // advanced whenever the captured iterations are dropped (drop_graphs)
    uint64_t serial = 0;                  // unique per handle of the process (a batched plan's key)
    nppe::BatchPlan *batch = nullptr;     // the batched solve's plan last built with this handle first (batch.cpp)
    double rate = 0.0;   // observed per-iteration reduction of the inner residual norm (0: none yet)
    int lastFirstPass = 0;   // iterations of the last FP64 Krylov solve's first pass (the next one's prediction)
    std::vector<void *> allocations;

    // MAC-field kernels around the solve (step.cpp): geometry, MAC width tables, row starts of the IDP numbering,
    // device scalars of their reductions, and staging for the host-array entry points (allocated on first use)
    nppe::FieldGeom fg{};
    double *dFieldTab = nullptr;
    int32_t *dRowStart = nullptr;
    double *dStepScal = nullptr;
    double *dCentre = nullptr;   // MAC::CentrePos of the x, y, z axes (LevelPressure's inflow shift), per call
    double *dStage[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};

    // profiling (nppe_profile.h)
    bool profiling = false;
    uint32_t profMask = ~0u;
    std::vector<hipEvent_t> evPool;
    struct Rec
    {
        int kind, iter;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Rec> recs;
    size_t evNext = 0;
    int curIter = -1;
    double profMs[NPPE_KERNEL_KINDS] = {};
    double profBytes[NPPE_KERNEL_KINDS] = {};
    int64_t profCount[NPPE_KERNEL_KINDS] = {};
};

namespace nppe
{

inline int fail(nppe_solver *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    return code;
}

inline int hip_fail(nppe_solver *h, hipError_t e, const char *what)
{
    if (h && h->budgetHit)
        return fail(h, NPPE_ERR_BUDGET, std::string(what) + ": the HBM budget of the handle is exhausted");
    return fail(h, NPPE_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(h, call, what)                                   \
    do                                                           \
    {                                                            \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return hip_fail((h), e_, (what));  \
    } while (0)

// Serialises device allocation and release against stream capture: a capture on one handle's stream must not see
// another thread's allocator synchronise the device.
extern std::mutex gCaptureMu;

template <class T>
hipError_t dalloc(nppe_solver *h, T **p, size_t count)
{
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    if (h->opt.hbm_budget && h->bytesHeld + bytes > h->opt.hbm_budget)
    {
        h->budgetHit = true;
        return hipErrorOutOfMemory;
    }
    hipError_t e;
    {
        std::lock_guard<std::mutex> lock(gCaptureMu);
        e = hipMalloc((void **)p, bytes);
    }
    if (e != hipSuccess) return e;
    h->allocations.push_back((void *)*p);
    h->bytesHeld += bytes;
    return hipMemsetAsync(*p, 0, bytes, h->stream);
}

// Host-to-device copy ordered on the handle's stream (never on the null stream, which would synchronise with
// other handles' streams), complete on return.
inline hipError_t upload(nppe_solver *h, void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream);
    return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
}

// A box array (pitch and plane strides multiples of 16 elements) whose element 1 sits on a 16-element boundary:
// the cell pair (2p + 1, 2p + 2) of every row then starts at an address aligned to the pair's size, so the
// whole-row sweep reads a pair with one load (rbgs_pairs_aligned checks this before choosing that kernel).
template <class T>
hipError_t dalloc_pairs(nppe_solver *h, T **p, size_t count)
{
    T *raw = nullptr;
    const hipError_t e = dalloc(h, &raw, count + 16);
    *p = raw ? raw + 1 : nullptr;
    return e;
}

// Slot of the handle's step scalars (dStepScal) that holds nppe_update_layout_device's count of changed cells.
constexpr int kStepScalDiff = 5;

// The coarsest level's dense pseudo-inverse (n unknowns) built on the device into dCoarseM / dCoarseOrder
// (coarse.hip).
int coarse_pinv_device(nppe_solver *h, int n);

// The allocation-free entry points' scratch array on `device` (current), looked up once per device (noalloc.hip).
int noalloc_scratch(int device, double **out);

// Field-kernel data of a committed grid (step.cpp); called at the end of the commit. refresh_fields recomputes the
// parts that depend on the layout (after nppe_update_layout).
int build_fields(nppe_solver *h);
int refresh_fields(nppe_solver *h);
// The solve on device vectors (solver.cpp).
int solve_device(nppe_solver *h, const double *dB, double *dX, const double *dGuess, double refNorm, nppe_result *res);
// MGPCG (FP64 Krylov) solves of several committed systems from no guess with their inner iterations batched: one
// launch per place of the iteration and kernel covers every system, each getting the bits of its solve alone
// (solver.cpp). Returns 1, having done nothing, when the systems do not all qualify (then solve them one by one);
// otherwise NPPE_OK with each system's status in rcs[q] and its result in res[q] (device_ms: the whole batch's), or a
// negative error that ended the batch.
int batch_solve_krylov64(nppe_solver *const *hs, const double *const *b, double *const *x, int count,
                         nppe_result *res, int *rcs);
void free_batch_plan(BatchPlan *p);

}  // namespace nppe
