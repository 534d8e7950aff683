// Internal types shared by the host driver (solver.cpp, hierarchy.cpp) and the kernels (kernels.hip).
//
// Box layout on the device: every level keeps the full box of its grid, ring included, in (slow, mid, fast) order
// with the fast axis padded to a pitch. In 3D slow/mid/fast are the reference's i (y), j (x), k (z)
// (MAC.h:254: index (i*Nx + j)*Nz + k); in 2D there is one slow plane and mid/fast are i (y), j (x).
#pragma once

#include <cstdint>

namespace nppe
{

constexpr int kBlockFast = 64;  // tile extent along the fast axis: one wavefront
constexpr int kBlockMid = 4;    // tile extent along the mid axis: four wavefronts per block
constexpr int kBlockThreads = kBlockFast * kBlockMid;
constexpr int kMaxDenseCoarse = 1024;  // largest coarsest level solved by the dense pseudo-inverse
constexpr int kMaxGrid = 2048;         // blocks per stencil launch: 8 per CU on 256 CUs; each loops over its tiles

// Geometry of one level's box.
struct Box
{
    int ns, nm, nf;     // extents incl. ring along slow, mid, fast (2D: ns = 1)
    int pitch;          // padded fast extent
    int threeD;         // 1 in 3D
    int tilesF, tilesM; // tiles of the interior along fast and mid
    int planes;         // slow planes of the interior (1 in 2D)
    int64_t sM, sS;     // element strides of mid and slow
    int64_t cells;      // allocated elements (ns * nm * pitch)
    __host__ __device__ int64_t at(int s, int m, int f) const { return s * sS + m * sM + f; }
    __host__ __device__ int blocks() const { return tilesF * tilesM * planes; }
};

// FP32 coefficient tables of one level, in A-units: for the face between cells n and n+1 along an axis,
// g[n] = sigma / FaceGap(n); for a Dirichlet ring neighbour the coefficient is area * h[n] with
// h[n] = 2 sigma / Width(n) (c * FaceGap / (Width/2), PressureSettings.h:307-310). Widths w[] form the areas.
// In 2D the slow tables hold a single 1.0 so that the area products reduce to one width.
struct Tables32
{
    const float *wS, *wM, *wF;
    const float *gS, *gM, *gF;
    const float *hS, *hM, *hF;
};

// One FP32 multigrid level as the kernels see it: box, tables, stencil codes (see launch_stencil_codes in
// layout.h) and, when every width of the level is equal, the constant face coefficient cu and ring-Dirichlet
// coefficient cr = 2 cu that replace the tables.
struct Level32
{
    Box B;
    Tables32 T;
    const uint8_t *c8;
    const uint16_t *c16;
    const uint32_t *rowc;   // per-row code descriptors (launch_row_codes), or null
    float cu, cr, icu;   // icu = 1 / cu
    int uniform;
};

// FP64 tables of the fine level, exactly as the reference forms them: widths and FaceGap(n) = (w[n]+w[n+1])/2.
struct Tables64
{
    const double *wS, *wM, *wF;
    const double *gS, *gM, *gF;
    double scale;   // MatrixScale3D: 1/(dh*dh) uniform, 1 stretched
    int unitCoef;   // every face coefficient is exactly 1.0: 3D uniform (PressureSolver3D.cpp:49-53), 2D square cells
    const uint8_t *c8;   // the fine level's stencil codes (a cell without a Dirichlet neighbour takes a fast path)
    const uint16_t *c16; // and their 2-bit-per-neighbour form (unit-coefficient rows beside a Dirichlet cell)
    const uint32_t *rowc; // its per-row code descriptors (launch_row_codes), or null
};

// Device-resident scalars of one solve. Written by the last block of each reduction.
// SolveState::done when a reduction inside the Krylov loop came out non-finite (NaN or infinity): the loop stops
// there, and the host reports NPPE_ERR_NONFINITE instead of running on to the iteration cap.
constexpr int kDoneNonFinite = 3;

// Directions the FP64 Krylov recursion keeps for its deferred solution update: x += alpha p is applied for
// kXRing iterations at once (every kXRing-th iteration, and once more at the end of a pass), from the directions
// in a ring of kXRing FP32 box vectors and their alphas, in iteration order -- the same operations, in the same
// order, as one update per iteration, for a sixth of the x traffic.
constexpr int kXRing = 8;

struct SolveState
{
    double rz;        // <r, z> of the current search direction
    double alpha;
    double beta;
    double rr;        // <r, r> after the last update
    double thr2;      // inner stop: rr < thr2
    double sums[4];   // FP64 reductions: [0] ||b||^2, [1] ||r||^2, [2] sum x, [3] spare
    int done;         // inner loop finished: converged (1), capped (2) or a non-finite reduction (kDoneNonFinite)
    int iters;        // inner iterations taken
    int maxIters;     // inner cap
    int first;        // next direction update is the first of the solve (beta = 0)
    unsigned ticket;  // blocks of the running reduction that have written their partial (0 between kernels)
    // FP64 Krylov solves scheduled without host round trips (solve_mgpcg): the starting <r, r>, the square of the
    // target, whether the guess is the answer, and whether the last true residual met the target
    double rr0;
    double acc2;
    int skip;
    int accepted;
    // the pass began without a guess and nothing has written x or the FP64 residual yet: x is still 0 (read as 0.0
    // by the deferred updates until the first full ring) and the residual still b (read from b by the first update)
    int zeroStart;
    double alphas[kXRing];   // alpha of the iteration whose direction sits in ring slot (iteration mod kXRing)
};

}  // namespace nppe
