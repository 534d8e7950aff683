/* nppe.h -- C ABI of the MI355X pressure Poisson backend for NEREUS's MAC-grid projection step.
 *
 * The reference selects its pressure backend through four calls in src/Methods/PPESolver/PressureSettings.h:
 *   CreatePressureBackend   (:487-514)   -> nppe_create
 *   CommitPressureMatrix    (:542-574)   -> nppe_commit_grid  (geometric: cell types + spacing tables; the CSR the
 *                                           reference assembles is optional, nppe_check_csr compares against it)
 *   SolvePressureSystem     (:578-624)   -> nppe_solve        (host FP64 vectors in IDP order, as AMGX gets them)
 *   DestroyPressureBackend  (:516-538)   -> nppe_destroy
 * The backend enum PressureBackend { CPU = 0, AMGX = 1 } (:262) gains a third value that routes to these.
 *
 * Conventions (mirroring the reference exactly):
 *   - Unknowns are the interior FLUID cells numbered lexicographically in (i, j, k), i = y slowest, k = z fastest
 *     (PressureSolver3D.cpp:133-139; 2D (i, j), PressureSolver2D.cpp:107-112). Cell (i, j, k) of the box lives at
 *     index (i*nx + j)*nz + k (MAC.h:254).
 *   - The operator is the reference's assembled matrix: negative semi-definite, scaled 1/dh^2 on a uniform grid and
 *     in volume-weighted form (face area / face gap) on a stretched one (PressureSolver3D.cpp:32-63,128-178).
 *   - Stopping: || b - A x ||_2 < tol * max(||b||_2, reference_norm) on the TRUE FP64 residual
 *     (PressureSettings.h:410, :576-577). When ||b|| <= tol * reference_norm the guess (or 0) is returned with 0
 *     iterations (:420-425).
 *   - The caller owns b, x and guess; the handle owns every device buffer.
 *   - Errors are negative return codes; nothing aborts across the ABI. One handle is used by one thread at a time;
 *     distinct handles are independent.
 */
#ifndef NPPE_H
#define NPPE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NPPE_ABI_VERSION 1

/* Return codes. NPPE_NOT_CONVERGED is a warning: the result is filled and x holds the last iterate (the reference
 * warns once and still returns its iterations, PressureSettings.h:472-479). NPPE_ERR_NONFINITE: a non-finite
 * right-hand side or guess, or a NaN/infinity that arose inside the Krylov recursion (the solve stops there; x is
 * not written). */
enum nppe_status
{
    NPPE_OK = 0,
    NPPE_NOT_CONVERGED = 1,
    NPPE_ERR_ARGUMENT = -1,
    NPPE_ERR_DEVICE = -2,
    NPPE_ERR_NO_GRID = -3,
    NPPE_ERR_SIZE = -4,
    NPPE_ERR_NONFINITE = -5,
    NPPE_ERR_BUDGET = -6,
    NPPE_ERR_MISMATCH = -7,
    NPPE_ERR_INTERNAL = -8
};

/* Solve methods. */
enum nppe_method
{
    NPPE_METHOD_MGPCG = 0, /* CG preconditioned by an FP32 geometric-multigrid V-cycle, its Krylov recursion in FP64
                              or in FP32 passes inside FP64 iterative refinement (nppe_options::krylov) */
    NPPE_METHOD_CG64 = 1   /* FP64 matrix-free Jacobi-PCG (Eigen's ConjugateGradient<..., DiagonalPreconditioner>) */
};

/* nppe_options::pre_smooth / post_smooth: chosen from the grid at commit. */
#define NPPE_SMOOTH_AUTO (-1)

/* Where MGPCG runs its Krylov recursion (the V-cycle preconditioner is FP32 either way). */
enum nppe_krylov
{
    NPPE_KRYLOV_REFINE = 0, /* FP32 CG passes inside FP64 iterative refinement (true FP64 residual between passes) */
    NPPE_KRYLOV_FP64 = 1    /* one CG with x, r, alpha, beta and A p in FP64 (directions stored in FP32) around the
                               FP32 V-cycle; the true FP64 residual is checked at its end (2D and 3D grids) */
};

/* Cell type codes (Definitions.h:191-196). */
enum nppe_cell
{
    NPPE_FLUID = 0,
    NPPE_SOLID = 1,
    NPPE_EMPTY = 2,
    NPPE_INFLOW = 3,
    NPPE_PRESSURE = 4,
    NPPE_OUTFLOW = 5
};

typedef struct nppe_options
{
    int32_t device;          /* HIP device ordinal */
    int32_t method;          /* enum nppe_method */
    double tolerance;        /* relative tolerance; eigen_tolerance's meaning (PressureSettings.h:225) */
    int32_t max_iters;       /* cap on iterations (inner iterations summed over refinement passes for MGPCG) */
    int32_t compatibility;   /* NEEDS_COMPATIBILITY_CONDITION: a singular system's solution is returned mean-free */
    int32_t pre_smooth;      /* red-black Gauss-Seidel sweeps before / after the coarse correction, or
                                NPPE_SMOOTH_AUTO (the default): 2 / 2 on uniform grids, 4 / 4 on stretched ones */
    int32_t post_smooth;
    double inner_reduction;  /* residual reduction each FP32 inner solve aims for at most (default 3e-5) */
    int32_t max_refinements; /* FP64 refinement passes (default 8); with NPPE_KRYLOV_FP64, restarts after a true
                                residual that misses the target */
    int32_t krylov;          /* enum nppe_krylov; default NPPE_KRYLOV_FP64 */
    uint64_t hbm_budget;     /* bytes of device memory the handle may hold; 0 = no limit */
} nppe_options;

/* Fills *opt with the defaults: device 0, MGPCG with the FP64 Krylov recursion, tolerance 1e-8, 2000 iterations,
 * compatibility on, automatic sweep counts, inner reduction 3e-5, 8 refinement passes, no budget. */
void nppe_default_options(nppe_options *opt);

/* The grid the reference hands CommitPressureMatrix through its MAC* (PressureSettings.h:542): box dimensions
 * including the one-cell ring, cell types, and the per-axis width tables MAC::Width(a, n) for n = 0 .. N-1
 * (MAC.h:78, ring cells included). FaceGap is recomputed from the widths with MAC::GapTable's expression
 * (MAC.cpp:165-175). */
typedef struct nppe_grid
{
    int32_t dim;                 /* 2 or 3 */
    int32_t nx, ny, nz;          /* MAC Nx, Ny, Nz (ring included); nz = 1 in 2D */
    int32_t non_uniform;         /* MAC::nonUniform */
    int32_t reserved0;
    double dh;                   /* MAC::dh */
    const uint8_t *cell_types;   /* nx*ny*nz bytes, index (i*nx + j)*nz + k */
    const double *width_x;       /* nx entries */
    const double *width_y;       /* ny entries */
    const double *width_z;       /* nz entries (ignored in 2D) */
} nppe_grid;

typedef struct nppe_result
{
    int32_t iterations;          /* MGPCG: FP32 inner iterations over all passes; CG64: CG iterations */
    int32_t refinements;         /* FP64 refinement passes (MGPCG); 0 for CG64 */
    int32_t status;              /* NPPE_OK or NPPE_NOT_CONVERGED */
    int32_t levels;              /* multigrid levels used */
    double rhs_norm;             /* ||b||_2 */
    double residual_norm;        /* true ||b - A x||_2 in FP64 */
    double relative_residual;    /* residual_norm / max(rhs_norm, reference_norm) */
    double device_ms;            /* device time of the solve */
    double algorithmic_bytes;    /* sum over the solve's launches of each kernel's compulsory HBM bytes */
} nppe_result;

typedef struct nppe_solver nppe_solver;

/* CreatePressureBackend (PressureSettings.h:487-514). opt may be NULL for the defaults. */
int nppe_create(nppe_solver **out, const nppe_options *opt);

/* DestroyPressureBackend (PressureSettings.h:516-538). NULL is accepted. */
int nppe_destroy(nppe_solver *h);

/* Message of the last error on this handle (empty when none); a NULL handle gives the calling thread's last
 * creation error. */
const char *nppe_last_error(const nppe_solver *h);

/* CommitPressureMatrix (PressureSettings.h:542-574) and the assembly it follows (PressureSolver3D.cpp:128-178):
 * numbers the unknowns, decides singularity (no Dirichlet face), uploads the cell types and tables and builds the
 * multigrid hierarchy. Returns the number of unknowns (>= 0) or an error. */
int64_t nppe_commit_grid(nppe_solver *h, const nppe_grid *g);

/* EnsureSystem / PressureLayoutChanged (PressureSolver3D.cpp:120-124; PressureSettings.h:300-303) for a layout
 * change of the committed box, its spacing unchanged (FLIP's moving liquid, FLIP3D.cpp:784-833): compares
 * cell_types (nx*ny*nz bytes, MAC order) with the committed layout and returns 0 when they are equal, doing
 * nothing. Otherwise it renumbers the unknowns and rebuilds every multigrid level's types and stencil codes and the
 * coarsest solve in the handle's existing buffers -- no device allocation, no tile search -- and returns 1; when
 * the new layout needs another hierarchy (a different level count, or a coarsest solve the buffers cannot hold) it
 * commits the grid again in full and returns 2. Solves after an update are bit-identical to those of a handle
 * committed with the new layout. Negative on error. */
int nppe_update_layout(nppe_solver *h, const uint8_t *cell_types);
/* The same with the cell types in device memory: the comparison runs on the device, and the types are read back
 * only when they differ. */
int nppe_update_layout_device(nppe_solver *h, const uint8_t *d_cell_types);

/* Unknown count and singularity of the committed system (PressureSystem::A.rows(), ::singular,
 * PressureSettings.h:284-297). */
int nppe_system_info(const nppe_solver *h, int64_t *n, int32_t *singular);

/* What the committed system's graph holds beyond PressureSystem::singular (no counterpart in the reference, which
 * assembles such systems without a word): its unknowns split into connected components (unknowns joined by a FLUID
 * face); a component with no Dirichlet face anywhere is "floating" -- A is singular on it, with the component's
 * indicator as null vector, so b must sum to zero over it -- and an unknown with no FLUID and no Dirichlet neighbour
 * at all is an empty row of A (the reference skips its zero diagonal, PressureSolver3D.cpp:168-169): b must be 0
 * there. The reference's compatibility fix removes the mean of b only when the whole system is singular
 * (PressureSolver3D.cpp:71-74, :300-307), so on a system with a Dirichlet face somewhere, an isolated unknown or a
 * floating pocket with nonzero b is an inconsistent system no solver converges on (MGPCG then runs to its cap and
 * returns NPPE_NOT_CONVERGED; CG64's recursion goes non-finite and returns NPPE_ERR_NONFINITE). Computed on the
 * host from the committed layout when asked. */
typedef struct nppe_system_report
{
    int64_t unknowns;            /* PressureSystem::A.rows() */
    int32_t singular;            /* PressureSystem::singular: no Dirichlet face anywhere */
    int32_t levels;              /* multigrid levels */
    int64_t components;          /* connected components of the unknowns */
    int64_t floating_components; /* components without a Dirichlet face (each one a null vector of A) */
    int64_t isolated_unknowns;   /* unknowns whose row of A is empty (no FLUID, no Dirichlet neighbour) */
    int64_t coarsest_unknowns;   /* unknowns of the coarsest level (its dense solve) */
    uint64_t device_bytes;       /* device memory the handle holds */
} nppe_system_report;
int nppe_describe_system(nppe_solver *h, nppe_system_report *out);

/* The IDP numbering of the committed system: order[q] = box index of unknown q (n entries). */
int nppe_unknown_order(const nppe_solver *h, int32_t *order);

/* Self-check for CommitPressureMatrix's triplets: compares the reference's assembled CSR (rowptr n+1, col/val nnz)
 * entry by entry with the matrix-free operator's coefficients. NPPE_OK when every entry is bit-identical,
 * NPPE_ERR_MISMATCH otherwise (last_error names the first difference). */
int nppe_check_csr(nppe_solver *h, int32_t n, int64_t nnz, const int32_t *rowptr, const int32_t *col,
                   const double *val);

/* SolvePressureSystem (PressureSettings.h:578-624): x (n entries, resized by the caller) = A^-1 b, from guess when
 * not NULL, to tol * max(||b||, reference_norm). res may be NULL. */
int nppe_solve(nppe_solver *h, const double *b, double *x, const double *guess, double reference_norm,
               nppe_result *res);

/* As nppe_solve with b, x and guess in device memory (FP64, IDP order, n entries), on the handle's stream. */
int nppe_solve_device(nppe_solver *h, const double *d_b, double *d_x, const double *d_guess, double reference_norm,
                      nppe_result *res);

/* y = A x with the matrix-free FP64 operator (host vectors, IDP order); the SpMV the reference's CG runs
 * (Eigen SparseDenseProduct over PressureSystem::A). */
int nppe_apply(nppe_solver *h, const double *x, double *y);

/* Device bytes the handle holds. */
int nppe_memory_bytes(const nppe_solver *h, uint64_t *bytes);

/* Changes the relative tolerance of later solves (a per-call eigen_tolerance; PressureSettings.h:369-371 sets the
 * AMGCL tolerance per call the same way). */
int nppe_set_tolerance(nppe_solver *h, double tolerance);

/* Upper estimate of the device bytes nppe_commit_grid will allocate for g, for admission against an HBM budget
 * before any allocation. 0 for an invalid grid. */
uint64_t nppe_estimate_bytes(const nppe_grid *g, int32_t method);
/* The same for a handle created with *opt (null: the defaults). */
uint64_t nppe_estimate_bytes_for(const nppe_grid *g, const nppe_options *opt);

/* ---- the pressure step around the solve, on MAC fields ---------------------------------------------------------
 * The reference's pressure and projection steps (PressureSolver3D::SolveScheme, Project; PressureSolver2D the same)
 * build the right-hand side from the intermediate velocity, solve, scatter the solution into p, apply the pressure
 * boundary conditions and project; these entry points do each part on the device, against the committed grid.
 * Fields are the reference's MAC arrays (MAC.h:235-254), FP64, unpitched, in device memory for the _device calls
 * and host memory for the others:
 *   u: ny*(nx+1)*nz values, u[(i*(nx+1) + j)*nz + k]     v: (ny+1)*nx*nz values, v[(i*nx + j)*nz + k]
 *   w: ny*nx*(nz+1) values, w[(i*nx + j)*(nz+1) + k]     p, p_old, prescribed: ny*nx*nz values (the cell index)
 * (2D: nz = 1 and w is not used). Vectors over the unknowns (b, x, f) have n entries in IDP order. p_old is the
 * previous step's pressure (gridAnt's p) under the incremental scheme and NULL under Chorin. prescribed, when not
 * NULL, holds MAC::PrescribedPressure at the PRESSURE ring cells (the caller evaluates the scene's boundary
 * pressure there; other entries are not read); when NULL a PRESSURE ring cell keeps the value p holds, which is
 * p_b(t) once the caller has run its EnforcePressureBC (SolveScheme does so first, PressureSolver3D.cpp:281).
 * Device pointers must be ready when the call is made; every call returns when its work on the handle's stream is
 * complete. */

/* The right-hand side of SolveScheme (PressureSolver3D.cpp:288-307; PressureSolver2D.cpp:262-280):
 * b = rho/dt * div(u) * V - scale * DirichletSum(p, p_old) (MAC::GetDivergencyAt, DirichletSum :248-260), then the
 * mean removed twice when the system is singular and the compatibility option is on. Bit-identical to the
 * reference's b before the mean removal. */
int nppe_rhs_device(nppe_solver *h, const double *d_u, const double *d_v, const double *d_w, const double *d_p,
                    const double *d_p_old, double rho, double dt, double *d_b);
int nppe_rhs(nppe_solver *h, const double *u, const double *v, const double *w, const double *p, const double *p_old,
             double rho, double dt, double *b);

/* ProjectPressure (PressureSolver3D.cpp:403-438; PressureSolver2D.cpp:370-396): u -= dt/rho * grad p on every face
 * that is not a wall, the gradient times DirichletRingFactor across a Dirichlet ring face; u, v, w in place. */
int nppe_project_device(nppe_solver *h, double *d_u, double *d_v, double *d_w, const double *d_p, double rho,
                        double dt);
int nppe_project(nppe_solver *h, double *u, double *v, double *w, const double *p, double rho, double dt);

/* MAC::EnforcePressureBC (MAC.cpp:866-871) over the ring list in MAC::BuildRingList's order (:1064-1096): EMPTY
 * ring cells 0, PRESSURE ring cells the prescribed value, every other ring cell its inward neighbour's value. */
int nppe_enforce_pressure_bc_device(nppe_solver *h, double *d_p, const double *d_prescribed);

/* IncrementReferenceNorm (PressureSolver3D.cpp:262-276; PressureSolver2D.cpp:237-251): ||b + A p^n||_2 with p^n
 * = p_old at the unknowns, the reference norm of an incremental solve. */
int nppe_increment_reference_norm_device(nppe_solver *h, const double *d_b, const double *d_p_old, double *norm);

/* nppe_increment_reference_norm_device without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses that forbid it). One fused
 * pass computes A p^n from the MAC array directly (no box copy of p^n) and reduces into a fixed device array of the
 * library (512 KiB per device) instead of the handle's buffers: grids of more than 262,144 interior rows (3D:
 * (ny-2)*(nx-2); 2D: ny-2) get NPPE_ERR_SIZE. The norm is bit-identical to nppe_increment_reference_norm_device's.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU race (one call at a time per GPU). */
int nppe_increment_reference_norm_noalloc(nppe_solver *h, const double *d_b, const double *d_p_old, double *norm);

/* SolveScheme's tail after the solve (PressureSolver3D.cpp:316-336; PressureSolver2D.cpp:287-304): with d_p_old
 * NULL (Chorin) x's mean is removed when the system is singular and the compatibility option is on
 * (LevelPressure; the level a scene's inflow p= expressions set, :188-211, needs the scene and is the caller's);
 * x is written into p at the unknowns, the ring BC applied, and with d_p_old every PRESSURE cell has p_old
 * subtracted (the increment's ring values). */
int nppe_scatter_pressure_device(nppe_solver *h, const double *d_x, double *d_p, const double *d_p_old,
                                 const double *d_prescribed);

/* AccumulatePressureIncrement (PressureSolver3D.cpp:374-393; PressureSolver2D.cpp:340-357): p = p_old + phi at
 * the unknowns (phi: p on entry), the mean removed as LevelPressure does, then the ring BC (without prescribed
 * values a PRESSURE ring cell gets p_old + phi there, p_b(t) to round-off). */
int nppe_accumulate_increment_device(nppe_solver *h, const double *d_p_old, double *d_p, const double *d_prescribed);

/* SolveLaplacian (PressureSolver3D.cpp:355-371; PressureSolver2D.cpp:322-337): L x = f on the committed system,
 * solved as A x = V f from a zero guess with no reference norm (FLIP's density correction reuses the hierarchy). */
int nppe_solve_laplacian_device(nppe_solver *h, const double *d_f, double *d_x, nppe_result *res);
int nppe_solve_laplacian(nppe_solver *h, const double *f, double *x, nppe_result *res);

/* The whole pressure step (SolveScheme after its first EnforcePressureBC, PressureSolver3D.cpp:279-336): b from
 * the fields, the reference norm when incremental, the solve from a zero guess, the scatter into p with the ring
 * BC and the incremental ring fix. The solve's NPPE_NOT_CONVERGED is returned after p is written, as the
 * reference warns and goes on. */
int nppe_pressure_step_device(nppe_solver *h, const double *d_u, const double *d_v, const double *d_w, double *d_p,
                              const double *d_p_old, double rho, double dt, nppe_result *res);

/* The projection step (Project, PressureSolver3D.cpp:396-401): ProjectPressure, then under the incremental scheme
 * (d_p_old not NULL) AccumulatePressureIncrement. */
int nppe_projection_step_device(nppe_solver *h, double *d_u, double *d_v, double *d_w, double *d_p,
                                const double *d_p_old, double rho, double dt, const double *d_prescribed);

/* ---- warm start and the inflow pressure level ------------------------------------------------------------------
 * The two parts of SolveScheme that need the previous step or the scene: BuildWarmStartGuess
 * (PressureSolver3D.cpp:215-244; PressureSolver2D.cpp:182-212) and LevelPressure's inflow shift (:15-30, :181-211;
 * 2D :13-29, :156-178). The scene's inlet levels are expressions the caller evaluates: it passes, per MAC cell, the
 * value SCENE.PressureLevelAt gives at each INFLOW ring cell's centre (NaN where it gives none; other cells are not
 * read) and the MAC::CentrePos tables the extrapolation uses. */
typedef struct nppe_inflow_level
{
    const double *d_level;   /* device, nx*ny*nz values (MAC cell index): the inlet's level at INFLOW ring cells */
    const double *centre_x;  /* host, nx values: MAC::CentrePos(X, n) for n = 0 .. nx-1 (ring cells included) */
    const double *centre_y;  /* host, ny values */
    const double *centre_z;  /* host, nz values (3D only) */
} nppe_inflow_level;

/* BuildWarmStartGuess once the caller's part of WarmStartEnabled holds (PRESSURE_SETTINGS.warmStart and a method
 * whose fluid region is fixed, MethodFixedFluidRegion): d_guess (n entries, IDP order) = the previous pressure
 * d_p_prev (MAC array) at the unknowns, its mean removed when the system is singular and the compatibility option is
 * on. *built = 0 (no guess; d_guess undefined) on a stretched grid or when d_p_prev is zero at every unknown. */
int nppe_warm_start_guess_device(nppe_solver *h, const double *d_p_prev, double *d_guess, int32_t *built);

/* nppe_scatter_pressure_device with LevelPressure's inflow shift after the mean removal (level may be NULL: no scene). */
int nppe_scatter_pressure_level_device(nppe_solver *h, const double *d_x, double *d_p, const double *d_p_old,
                                       const double *d_prescribed, const nppe_inflow_level *level);
/* nppe_accumulate_increment_device with LevelPressure's inflow shift (AccumulatePressureIncrement levels p^n + phi,
 * PressureSolver3D.cpp:374-393). */
int nppe_accumulate_increment_level_device(nppe_solver *h, const double *d_p_old, double *d_p,
                                           const double *d_prescribed, const nppe_inflow_level *level);

typedef struct nppe_step_options
{
    const double *d_p_prev;               /* gridAnt's p (MAC array, device): the warm start's source; NULL: none */
    int32_t warm_start;                   /* the caller's part of WarmStartEnabled (see above) */
    int32_t reserved0;
    const double *d_prescribed;           /* as nppe_scatter_pressure_device (NULL: PRESSURE ring cells keep p's) */
    const nppe_inflow_level *inflow_level;/* NULL: no scene (LevelPressure removes the mean only) */
    int32_t *warm_started;                /* out, may be NULL: 1 when the solve started from the warm-start guess */
} nppe_step_options;

/* nppe_pressure_step_device with the rest of SolveScheme (PressureSolver3D.cpp:279-336): under Chorin (d_p_old NULL)
 * the solve starts from BuildWarmStartGuess's guess when o->warm_start is set and a guess is built (:309-314; the
 * AMGX-only closed-box gate does not apply to this backend), and LevelPressure applies the scene's inflow level;
 * o may be NULL (then it is nppe_pressure_step_device). */
int nppe_pressure_step_ex_device(nppe_solver *h, const double *d_u, const double *d_v, const double *d_w, double *d_p,
                                 const double *d_p_old, double rho, double dt, const nppe_step_options *o,
                                 nppe_result *res);
/* nppe_projection_step_device whose accumulation applies the scene's inflow level (level may be NULL). */
int nppe_projection_step_ex_device(nppe_solver *h, double *d_u, double *d_v, double *d_w, double *d_p,
                                   const double *d_p_old, double rho, double dt, const double *d_prescribed,
                                   const nppe_inflow_level *level);

/* ---- many systems at once (no counterpart in the reference, which solves one system per process) ------------- */

/* Solves count independent committed systems together. b[i], x[i] are host (device == 0) or device (device == 1)
 * FP64 vectors in IDP order of hs[i]; results go to res[i] (may be NULL). With device vectors and every system an
 * MGPCG system with the FP64 Krylov recursion whose V-cycle runs the whole-row kernels (3D, unit coefficients, rows
 * of at most 256 cells), the inner iterations are batched: one launch per V-cycle level and kernel covers all
 * systems, the batch's iteration replays as one graph, all on hs[0]'s stream, and res[i].device_ms is the whole
 * batch's time. The first call on a set of handles (or after one of them changed its grid or layout) records and
 * merges the launches and times the tile shapes of the shared sweeps and restrictions (tens of milliseconds); later
 * calls reuse that plan. Otherwise each handle runs on its own stream, driven by `workers` host threads (0: one per system,
 * at most 64). Returns NPPE_OK when every solve returned NPPE_OK, the first negative code otherwise, or
 * NPPE_NOT_CONVERGED. Each result is bit-identical to the same solve run alone. */
int nppe_batch_solve(nppe_solver *const *hs, const double *const *b, double *const *x, int32_t count,
                     int32_t workers, int32_t device, nppe_result *res);

/* A solve server: requests arrive at any time, are admitted against an HBM budget into `slots` concurrent solver
 * slots as slots and memory free up, and run concurrently on the device (continuous batching). Grids are
 * registered once under a key and stay resident (committed) between requests while the budget allows; idle
 * resident grids are evicted least recently used. */
typedef struct nppe_server nppe_server;
int nppe_server_create(nppe_server **out, const nppe_options *opt, int32_t slots, uint64_t hbm_budget);
int nppe_server_destroy(nppe_server *s);
/* Registers (copies) a grid under key. */
int nppe_server_add_grid(nppe_server *s, uint64_t key, const nppe_grid *g);
/* Queues a solve of the grid registered under key with right-hand side b (copied; IDP order) to the given relative
 * tolerance; the solution is written to x (host, caller-owned, must stay valid until the request is waited on).
 * Returns a ticket (>= 0) or an error. */
int64_t nppe_server_submit(nppe_server *s, uint64_t key, const double *b, double tolerance, double *x);
/* Blocks until the request is done; its result, and its latency from submit to completion in milliseconds. */
int nppe_server_wait(nppe_server *s, int64_t ticket, nppe_result *res, double *latency_ms);
/* The error message of the request last waited on ("" when it succeeded); valid until the next wait. */
const char *nppe_server_last_error(const nppe_server *s);
/* Device bytes the server holds now, and the most it has held. */
int nppe_server_memory(const nppe_server *s, uint64_t *now, uint64_t *peak);

#ifdef __cplusplus
}
#endif
#endif
