/* CPU oracle for the NEREUS pressure Poisson path: a plain-C restatement of the reference's MAC spacing tables,
 * matrix assembly and FP64 Krylov solve, used ONLY by the tests and by bench.py's CPU-baseline leg as the checker.
 * Nothing in the product library links or calls it.
 *
 * Parity status of each function (how it is pinned):
 *   oracle_build_axis     - bit-exact against the compiled reference's MAC tables (golden fixtures), uniform,
 *                           hard and graded refinement; widths sum to the domain.
 *   oracle_gap_table      - bit-exact against the reference's FaceGap tables (golden fixtures).
 *   oracle_assemble       - bit-exact CSR (pattern and values) against the compiled reference's assembly for
 *                           2D/3D, uniform/stretched, solids, EMPTY/PRESSURE/FLUID ring cells (golden fixtures);
 *                           plus symmetry, A*1 = 0 on Neumann rows and the linear-field closed form.
 *   oracle_cg             - iteration counts and solutions against the reference's Eigen CG (golden fixtures);
 *                           exact solve on tiny systems against a dense factorisation.
 *   oracle_subtract_mean  - closed form (mean of the result is 0 to round-off).
 *   oracle_rhs            - bit-exact against the compiled reference's GetDivergencyAt / DirichletSum expression
 *                           (golden fixtures: 2D/3D, uniform/stretched, PRESSURE ring cells, incremental);
 *                           plus the closed form of a linear velocity field and of a PRESSURE ring.
 *   oracle_project        - bit-exact against the compiled reference's ProjectPressure (golden fixtures); plus the
 *                           discrete divergence of the projected field vanishing after a solve.
 *   oracle_enforce_pressure_bc - bit-exact against the compiled reference's MAC::EnforcePressureBC with a known
 *                           boundary pressure (golden fixtures: corner and edge cells of the sequential ring list).
 *   oracle_scatter_pressure - bit-exact against the compiled reference's SolveScheme tail (scatter, ring BC and
 *                           incremental ring fix) where no mean is removed; within 1e-14 of it where LevelPressure
 *                           removes the mean (Eigen's vectorised mean sums in another order).
 *   oracle_accumulate_increment - the same bars against the compiled reference's AccumulatePressureIncrement.
 *   oracle_level_inflow   - against the compiled reference's LevelPressure with a scene of inlet p= levels
 *                           (level_*.npz: 2D/3D, uniform/stretched), to 1e-14 (the mean sums in another order);
 *                           a constant level c on every inlet makes the extrapolated ring values average c.
 *   oracle_warm_start_guess - against the compiled reference's BuildWarmStartGuess (level_*.npz), bit-exact where
 *                           no mean is removed, 1e-14 where it is; a zero previous pressure gives no guess.
 */
#ifndef NPPE_ORACLE_H
#define NPPE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cell types (reference src/headers/Core/Definitions.h:191-196). */
enum { ORACLE_FLUID = 0, ORACLE_SOLID = 1, ORACLE_EMPTY = 2, ORACLE_INFLOW = 3, ORACLE_PRESSURE = 4, ORACLE_OUTFLOW = 5 };
/* Refinement modes (Definitions.h:238-244). */
enum { ORACLE_REFINE_HARD = 0, ORACLE_REFINE_RAMP = 1, ORACLE_REFINE_SYMMETRIC = 2, ORACLE_REFINE_RAMP_BACK = 3 };

typedef struct
{
    double start, end;
    int32_t factor, mode;
} oracle_refinement;

/* Widths of one axis including the two ring cells (MAC::BuildAxis + MAC::AddRing). Returns the count, or -1 when
 * `cap` is too small. */
int oracle_build_axis(double a0, double af, int nBase, double dh, const oracle_refinement *refs, int nrefs,
                      int nonUniform, double *widths, int cap);

/* FaceGap(a, m) for m = -1 .. n-1, stored at gap[m + 1] (MAC::GapTable). */
void oracle_gap_table(const double *w, int n, double *gap);

typedef struct
{
    int32_t dim;        /* 2 or 3 */
    int32_t Nx, Ny, Nz; /* cells including the ring; Nz = 1 in 2D */
    int32_t nonUniform;
    double dh;
    const uint8_t *types; /* index (i*Nx + j)*Nz + k */
    const double *wx, *wy, *wz; /* widths incl. ring: Nx, Ny, Nz entries (wz unused in 2D) */
} oracle_grid;

/* Unknowns: interior FLUID cells. */
int64_t oracle_count_unknowns(const oracle_grid *g);

/* Assembles the reference's CSR matrix. order[n] gets each unknown's box index; rowptr[n+1], col/val up to
 * 7n (5n in 2D) entries. Returns nnz; *singular set when no Dirichlet face exists. */
int64_t oracle_assemble(const oracle_grid *g, int32_t *order, int32_t *rowptr, int32_t *col, double *val,
                        int32_t *singular);

void oracle_spmv(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *x, double *y);

/* Jacobi (precond = 0) or unpreconditioned (precond = 1) CG with Eigen 3.3.7's stopping rule; x holds the guess on
 * entry. Returns iterations; *tol_error gets sqrt(|r|^2/|b|^2) of the recursive residual. */
int oracle_cg(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b, double *x,
              double tol, int maxIters, int precond, double *tol_error);

/* SolveEigenPressure's absolute-floor mapping (PressureSettings.h:411-427) around oracle_cg. */
int oracle_solve_pressure(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b,
                          double *x, const double *guess, double tol, int maxIters, double referenceNorm);

void oracle_subtract_mean(double *v, int64_t n);

/* The pressure equation's right-hand side in IDP order (order[n] from oracle_assemble): b = rho/dt * div(u) * V -
 * scale * DirichletSum, the expression of PressureSolver3D.cpp:288-298 / PressureSolver2D.cpp:262-271 with
 * MAC::GetDivergencyAt, the cell volume and DirichletSum (PRESSURE neighbours; p_b - p_b^n when pOld is given).
 * u, v, w, p, pOld are MAC arrays: u[(i*(Nx+1) + j)*Nz + k], v[(i*Nx + j)*Nz + k] (i <= Ny-1+1),
 * w[(i*Nx + j)*(Nz+1) + k], p[(i*Nx + j)*Nz + k]; in 2D w is unused and Nz = 1. */
void oracle_rhs(const oracle_grid *g, const int32_t *order, int64_t n, const double *u, const double *v,
                const double *w, const double *p, const double *pOld, double rho, double dt, double *b);

/* The projection u -= dt/rho * grad p on every face that is not a wall, with DirichletRingFactor on faces whose
 * outer cell is a Dirichlet ring cell (PressureSolver3D.cpp:399-438; PressureSolver2D.cpp:370-396), in place. */
void oracle_project(const oracle_grid *g, double *u, double *v, double *w, const double *p, double rho, double dt);

/* MAC::EnforcePressureBC (MAC.cpp:866-871) in the order of MAC::BuildRingList (MAC.cpp:1064-1096; RecordRing,
 * MAC.h:482-486): every ring cell on the y faces, then the x faces, then (3D) the z faces; an EMPTY cell gets 0, a
 * PRESSURE cell prescribed[c] (left as it is when prescribed is NULL), any other the inward neighbour's current
 * value. p is the MAC pressure array. */
void oracle_enforce_pressure_bc(const oracle_grid *g, double *p, const double *prescribed);

/* SolveScheme's tail after the solve (PressureSolver3D.cpp:316-336; PressureSolver2D.cpp:287-304) with x (IDP
 * order) as the solve's answer: when pOld is NULL (Chorin) and `level` is set (NeedsCompatibility), x's mean is
 * removed (LevelPressure without a scene); x is scattered into p, the ring BC applied, and with pOld (incremental)
 * every PRESSURE cell then has pOld subtracted. */
void oracle_scatter_pressure(const oracle_grid *g, const int32_t *order, int64_t n, const double *x, double *p,
                             const double *pOld, int level, const double *prescribed);

/* AccumulatePressureIncrement (PressureSolver3D.cpp:374-393; PressureSolver2D.cpp:340-357): p = pOld + phi at the
 * unknowns (phi: p on entry), the mean removed when `level` is set, then the ring BC. */
void oracle_accumulate_increment(const oracle_grid *g, const int32_t *order, int64_t n, const double *pOld,
                                 double *p, int level, const double *prescribed);

/* LevelPressure's inflow shift (PressureSolver3D.cpp:15-30, :181-211; PressureSolver2D.cpp:13-29, :156-178) on a
 * mean-free x (IDP order): level[c] is the scene's level at ring cell c (NaN: none), centre_* MAC::CentrePos per axis.
 * Returns the number of contributing ring cells. */
int oracle_level_inflow(const oracle_grid *g, const int32_t *order, int64_t n, double *x, const double *level,
                        const double *centre_x, const double *centre_y, const double *centre_z);

/* BuildWarmStartGuess (PressureSolver3D.cpp:222-244) once WarmStartEnabled holds: returns 1 and x0, or 0 (no guess). */
int oracle_warm_start_guess(const oracle_grid *g, const int32_t *order, int64_t n, const double *p_prev, int level,
                            double *x0);

/* |b - A x|_2 / |b|_2 in FP64. */
double oracle_relres(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b,
                     const double *x);

#ifdef __cplusplus
}
#endif
#endif
