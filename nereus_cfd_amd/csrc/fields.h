// Launchers of the MAC-field kernels around the solve (fields.hip): the pressure equation's right-hand side from
// the velocity field, the projection, the pressure boundary conditions and the increment bookkeeping of the
// incremental scheme. Every launcher enqueues on the given stream and returns the HIP status of the launch.
//
// MAC arrays are the reference's, unpitched, in device memory (MAC.h:235-254): p[(i*Nx + j)*Nz + k],
// u[(i*(Nx+1) + j)*Nz + k], v[(i*Nx + j)*Nz + k] for i <= Ny, w[(i*Nx + j)*(Nz+1) + k]; in 2D Nz = 1 and w is
// absent. Vectors over the unknowns (b, x) are in IDP order.
#pragma once

#include <hip/hip_runtime.h>

#include "nppe_internal.h"

namespace nppe
{

// The committed grid as the field kernels see it.
struct FieldGeom
{
    Box B;                     // the fine level's box: 3D (slow, mid, fast) = (i, j, k); 2D one plane, (i, j)
    int dim, Nx, Ny, Nz;       // MAC dimensions, ring included (Nz = 1 in 2D)
    int nonUniform;            // MAC::nonUniform
    double dh;                 // MAC::dh
    const double *wx, *wy, *wz;   // MAC::Width tables (wz: one entry in 2D)
    const uint8_t *types;      // cell types in the pitched box
    const uint8_t *c8;         // stencil codes in the pitched box: bit 6 = unknown, bit 7 = has a Dirichlet neighbour
    const int32_t *rowStart;   // unknowns before each interior row (3D rows (i, j), 2D rows i)
    int rows;                  // interior rows: 3D (Ny-2)(Nx-2), 2D Ny-2
    int64_t n;                 // unknowns
};

// Blocks of a row-walking launch (four rows per block), which is also the number of partial sums it writes; the
// same for a launch over every row of the box (ring rows included).
int field_row_blocks(const FieldGeom &G);
int field_box_blocks(const FieldGeom &G);

// b[id] = rho/dt * div(u) * V - scale * DirichletSum(p, pOld) for every unknown (pOld null: Chorin).
hipError_t launch_rhs(hipStream_t s, const FieldGeom &G, const double *u, const double *v, const double *w,
                      const double *p, const double *pOld, double rhoDt, double scale, double *b);
// u -= f * grad p on every face that is not a wall (f = dt / rho), in place.
hipError_t launch_project(hipStream_t s, const FieldGeom &G, double *u, double *v, double *w, const double *p,
                          double f);
// The ring list of MAC::EnforcePressureBC, pass by pass (y faces, x faces, z faces): an EMPTY ring cell gets 0, a
// PRESSURE one prescribed[c] (its own value when prescribed is null), any other the inward neighbour's value.
hipError_t launch_pressure_bc(hipStream_t s, const FieldGeom &G, double *p, const double *prescribed);
// p[c] = x[id] - (sumX ? *sumX / n : 0) at the unknowns.
hipError_t launch_scatter_p(hipStream_t s, const FieldGeom &G, const double *x, const double *sumX, double *p);
// p[c] -= pOld[c] at every PRESSURE cell of the box; with restore, p[c] = pOld[c] + p[c] at the PRESSURE cells of
// the ring only (each once, where the ring list visits edge and corner cells more than once).
hipError_t launch_ring_fix(hipStream_t s, const FieldGeom &G, const double *pOld, double *p, bool restore = false);
// partials[block] = sum over the unknowns of pOld[c] + p[c].
hipError_t launch_sum_increment(hipStream_t s, const FieldGeom &G, const double *pOld, const double *p,
                                double *partials);
// p[c] = (pOld[c] + p[c]) - (sum ? *sum / n : 0) at the unknowns.
hipError_t launch_accumulate(hipStream_t s, const FieldGeom &G, const double *pOld, const double *sum, double *p);
// box[pitched c] = pOld[c] at the unknowns (the other cells are not written).
hipError_t launch_gather_box(hipStream_t s, const FieldGeom &G, const double *pOld, double *box);
// partials[block] = sum over the unknowns of (b[id] + y[pitched c])^2.
hipError_t launch_norm2_sum(hipStream_t s, const FieldGeom &G, const double *b, const double *y, double *partials);
// b[id] = f[id] * V(cell) (V = 1.0 on a uniform grid).
hipError_t launch_scale_volume(hipStream_t s, const FieldGeom &G, const double *f, double *b);

// LevelPressure's inflow shift on p (MAC array, mean already removed at the unknowns): partials[block] = the sum
// of the INFLOW ring cells' terms, partials[blocks + block] their count, over inflow_level_blocks(G) blocks; level
// is the scene's value per MAC cell (NaN: none), centres MAC::CentrePos of the x, y, z axes back to back.
int inflow_level_blocks(const FieldGeom &G);
hipError_t launch_inflow_level(hipStream_t s, const FieldGeom &G, const double *p, const double *level,
                               const double *centres, double *partials);
// p[c] += sumCount[0] / sumCount[1] at the unknowns when sumCount[1] > 0.
hipError_t launch_add_shift(hipStream_t s, const FieldGeom &G, const double *sumCount, double *p);
// x[id] = p[c] at the unknowns; partials[block] = the block's sum of p[c]^2 (field_row_blocks(G) blocks).
hipError_t launch_gather_idp(hipStream_t s, const FieldGeom &G, const double *p, double *x, double *partials);

// partials[block] = the number of cells whose type in newTypes (MAC order, unpitched) differs from the committed
// one, partials[blocks + block] the number of cells whose code is out of range (> 5); *blocks = the blocks.
hipError_t launch_diff_types(hipStream_t s, const FieldGeom &G, const uint8_t *newTypes, double *partials, int *blocks);

// Linear vectors of n entries.
// partials[block] = sum of x over the block's share; returns the block count through *blocks.
hipError_t launch_sum_linear(hipStream_t s, const double *x, int64_t n, double *partials, int *blocks);
// x[q] -= *sum / n.
hipError_t launch_sub_mean(hipStream_t s, double *x, int64_t n, const double *sum);
// *dst = sum of partials[0 .. count) in a fixed order (one block).
hipError_t launch_sum_partials(hipStream_t s, const double *partials, int count, double *dst);

}  // namespace nppe
