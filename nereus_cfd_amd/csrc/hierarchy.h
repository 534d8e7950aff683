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

// Per-cell stencil codes of a level (unpitched box order). code8: bits 0-5 = neighbour q (slow+, slow-, mid+,
// mid-, fast+, fast-) is a coupled unknown; bit 6 = the cell is an unknown; bit 7 = the cell has a Dirichlet
// neighbour (read code16 for it). code16: two bits per neighbour q at bit 2q -- 0 Neumann, 1 coupled unknown,
// 2 Dirichlet with the face coefficient (an interior EMPTY/PRESSURE cell, or a FLUID ring cell), 3 Dirichlet ring
// cell (EMPTY/PRESSURE on the ring: coefficient times FaceGap/(Width/2)); bit 12 = the cell is an unknown.
void stencil_codes(const HostLevel &L, std::vector<uint8_t> &code8, std::vector<uint16_t> &code16);

// FP32 tables (widths, sigma/gap, 2 sigma/width) of a level, packed as wS wM wF gS gM gF hS hM hF; offsets[9].
std::vector<float> tables32(const HostLevel &L, double sigma, int offsets[9]);

// FP64 tables of the fine level: wS wM wF gS gM gF (gap = FaceGap as MAC::GapTable forms it); offsets[6].
std::vector<double> tables64(const HostLevel &L, int offsets[6]);

// Dense pseudo-inverse (row-major n x n) of the level's operator in A-units, sigma-scaled like tables32, with
// each connected component that has no Dirichlet face treated as singular (its constant mode projected out).
// order receives the unpitched box index of each unknown, lexicographic.
bool coarse_pseudo_inverse(const HostLevel &L, double sigma, std::vector<int64_t> &order, std::vector<double> &M,
                           std::string &err);

}  // namespace nppe
