/* nppe_profile.h -- per-kernel timing of a solver handle (no counterpart in the reference, whose only timing is the
 * manager's wall clock around the whole pressure step, SimulationManager.h:345-353).
 *
 * When enabled, every launch of a solve is bracketed by HIP events on the handle's stream; after each solve the
 * durations of the launches that did work are summed per kernel kind, together with the algorithmic (compulsory
 * HBM) bytes each launch moves. achieved GB/s of a kind = bytes / total_ms / 1e6. */
#ifndef NPPE_PROFILE_H
#define NPPE_PROFILE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nppe_kernel_kind
{
    NPPE_K_SMOOTH = 0,       /* red-black Gauss-Seidel sweep (both colours), FP32, levels >= 1 */
    NPPE_K_RESTRICT = 1,     /* residual + restriction, FP32 */
    NPPE_K_PROLONG = 2,      /* prolongation + correction, FP32 */
    NPPE_K_COARSE = 3,       /* coarsest-level dense FP64 solve */
    NPPE_K_CG_DOT = 4,       /* <r, z>, FP32 */
    NPPE_K_CG_DIRECTION = 5, /* p = z + beta p, FP32 */
    NPPE_K_CG_SPMV = 6,      /* q = A p with <p, q>, FP32 */
    NPPE_K_CG_UPDATE = 7,    /* e += alpha p, r -= alpha q with <r, r>, FP32 */
    NPPE_K_RESIDUAL64 = 8,   /* FP64 x update + true residual r = b - A x */
    NPPE_K_SCATTER = 9,      /* compacted -> box */
    NPPE_K_GATHER = 10,      /* box -> compacted (with the mean projection) */
    NPPE_K_MEAN = 11,        /* FP64 sum for the mean projection */
    NPPE_K_CG64_INIT = 12,
    NPPE_K_CG64_SPMV = 13,   /* FP64 q = A p with <p, q> */
    NPPE_K_CG64_UPDATE = 14, /* FP64 x, r update with <r, r>, <r, M^-1 r> */
    NPPE_K_CG64_DIRECTION = 15,
    NPPE_K_SMOOTH0 = 16,     /* red-black Gauss-Seidel sweep (both colours), FP32, the finest level */
    NPPE_K_SMOOTH_RESTRICT = 17, /* the last pre-smoothing sweep with the residual restriction folded in, FP32 */
    NPPE_KERNEL_KINDS = 18
};

struct nppe_solver;

/* Turns event timing on or off and clears the totals. */
int nppe_profile_enable(struct nppe_solver *h, int on);
/* Restricts the timing to the kinds whose bit (1 << kind) is set in mask (all kinds after nppe_profile_enable);
 * launches of the other kinds run without events around them. */
int nppe_profile_select(struct nppe_solver *h, uint32_t mask);
/* Totals since the last nppe_profile_enable: launches that did work, their summed duration and algorithmic bytes. */
int nppe_profile_read(const struct nppe_solver *h, int kind, int64_t *launches, double *total_ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
