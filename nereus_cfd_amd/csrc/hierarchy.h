// Host-side construction of the solver's grid hierarchy: unknown numbering, coarse cell types and spacing,
// FP32/FP64 coefficient tables and the dense pseudo-inverse of the coarsest operator.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "nppe_internal.h"

namespace nppe
{

// One level on the host, axes in (slow, mid, fast) order. Types are unpitched: index (s*nm + m)*nf + f.
struct HostLevel
{
    int ns = 1, nm = 0, nf = 0;
    bool threeD = true;
    std::vector<uint8_t> types;
    std::vector<double> w[3];   // widths incl. ring along slow, mid, fast (2D: w[0] = {1.0})
    int64_t unknowns = 0;
    bool uniform = false;       // every width of every axis equal
};

Box make_box(int ns, int nm, int nf, bool threeD);

// The fine level from the reference's grid (dims incl. ring, types in MAC order, MAC width tables).
HostLevel fine_level(int dim, int nx, int ny, int nz, const uint8_t *types, const double *wx, const double *wy,
                     const double *wz);

// Cell-centred 2:1 coarsening of every axis with interior extent > 1: a coarse cell is Dirichlet (EMPTY) if any
// child is EMPTY/PRESSURE, else FLUID if any child is FLUID, else SOLID; its width is the sum of its children's.
HostLevel coarsen(const HostLevel &f);

// FP32 tables (widths, sigma/gap, 2 sigma/width) of a level, packed as wS wM wF gS gM gF hS hM hF; offsets[9].
std::vector<float> tables32(const HostLevel &L, double sigma, int offsets[9]);

// FP64 tables of the fine level: wS wM wF gS gM gF (gap = FaceGap as MAC::GapTable forms it); offsets[6].
std::vector<double> tables64(const HostLevel &L, int offsets[6]);

}  // namespace nppe
