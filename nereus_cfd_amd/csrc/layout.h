// Launchers of the layout kernels (layout.hip): a new cell layout of the committed box made into every multigrid
// level's types and stencil codes, the IDP numbering and the counts that decide the hierarchy's shape, on the
// device. Every launcher enqueues on the given stream and returns the HIP status of the launch.
#pragma once

#include <hip/hip_runtime.h>

#include "fields.h"
#include "nppe_internal.h"

namespace nppe
{

// Coarse types of level c from level f (cell-centred 2:1 coarsening, hierarchy.h coarsen): a coarse cell is EMPTY if
// any child is EMPTY/PRESSURE, else FLUID if any child is FLUID, else SOLID. Every cell of the coarse box, ring
// included; the pitch padding is not written.
hipError_t launch_coarsen_types(hipStream_t s, const Box &Bf, const uint8_t *tf, const Box &Bc, uint8_t *tc);

// Stencil codes of every interior cell of a level from its types (zero for a cell that is not an unknown). code8:
// bits 0-5 = neighbour q (slow+, slow-, mid+, mid-, fast+, fast-) is a coupled unknown; bit 6 = the cell is an
// unknown; bit 7 = the cell has a Dirichlet neighbour (read code16 for it). code16: two bits per neighbour q at bit
// 2q -- 0 Neumann, 1 coupled unknown, 2 Dirichlet with the face coefficient (an interior EMPTY/PRESSURE cell, or a
// FLUID ring cell), 3 Dirichlet ring cell (EMPTY/PRESSURE on the ring: coefficient times FaceGap/(Width/2)); bit 12 =
// the cell is an unknown. Built at commit and after every layout update.
hipError_t launch_stencil_codes(hipStream_t s, const Box &B, const uint8_t *t, uint8_t *c8, uint16_t *c16);

// Per row (s, m) of the box, ring rows included, at rowc[m * ns + s]: bits 0-7 the code8 of the row's first
// interior cell (f = 1), 8-15 that of its cells f = 2 .. nf - 3 when they all have one code (else the code of f = 2,
// or 0 for rows of fewer than 3 cells), 16-23 that of its last interior cell (f = nf - 2), bit 31 set when cells
// 2 .. nf - 3 do share one code. A sweep then builds such a row's codes instead of reading them. Built after the
// stencil codes, at commit and after every layout update.
hipError_t launch_row_codes(hipStream_t s, const Box &B, const uint8_t *c8, uint32_t *rowc);

// counts[0] += unknowns of the level (code8 bit 6), counts[1] += unknowns with a Dirichlet face (bit 7), with
// partial sums per block in partials (at most 2 * kMaxCountBlocks entries) finished by launch_finish_counts.
constexpr int kMaxCountBlocks = 1024;
hipError_t launch_count_codes(hipStream_t s, const Box &B, const uint8_t *c8, unsigned long long *partials,
                              int *blocks);
hipError_t launch_finish_counts(hipStream_t s, const unsigned long long *partials, int blocks, long long *out);

// The fine level's IDP numbering: per interior row the unknowns (code8 bit 6), their exclusive prefix sums into
// G.rowStart (one block), then order[id] = the pitched box index of unknown id. *total gets the unknown count.
hipError_t launch_number_rows(hipStream_t s, const FieldGeom &G, int32_t *rowStart, long long *total);
hipError_t launch_write_order(hipStream_t s, const FieldGeom &G, int32_t *order);

}  // namespace nppe
