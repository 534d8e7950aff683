/* nppe_synth.h -- seeded synthetic right-hand sides for the benchmark and the tests. No counterpart in the
 * reference, whose right-hand side is the divergence of the intermediate velocity (PressureSolver3D.cpp:288-298). */
#ifndef NPPE_SYNTH_H
#define NPPE_SYNTH_H

#include <stdint.h>

#include "nppe.h"

#ifdef __cplusplus
extern "C" {
#endif

/* b over the unknowns of g (IDP order): sum over `modes` modes of a_m * prod_d cos(pi k_{m,d} x_d), x_d the cell
 * centre scaled to [0, 1] along each axis, plus noise * N(0, 1) per cell; a_m ~ U(-1, 1), k_{m,d} ~ U{1..4}. All
 * draws come from one std::mt19937_64(seed) stream: the mode parameters first (a_m, then k_{m,0..dim-1}, mode by
 * mode), then the noise in IDP order. Uniforms are (u64 >> 11) * 2^-53; normals are Box-Muller pairs. With
 * zero_mean the mean is removed twice (PressureSolver3D.cpp:300-307). Returns the number of unknowns written. */
int64_t nppe_synth_cosine_rhs(const nppe_grid *g, uint64_t seed, int32_t modes, double noise, int32_t zero_mean,
                              double *b);

#ifdef __cplusplus
}
#endif
#endif
