// The reference's assembled matrix row as device code: the FP64 operator with the reference's exact arithmetic
// (PressureSolver3D.cpp:47-87, :145-170; PressureSolver2D.cpp:117-143), shared by the solver kernels (kernels.hip)
// and the allocation-free entry points (noalloc.hip). Built on the fly from the cell types of the cell and its six
// neighbours and from 1D spacing tables:
//   FLUID neighbour             -> diag -= c; off-diagonal c when the neighbour is an unknown (not on the ring)
//   EMPTY / PRESSURE neighbour  -> diag -= c * rf, rf = FaceGap / (Width/2) on the domain ring, 1 inside
//   SOLID / INFLOW / OUTFLOW    -> nothing (Neumann)
#pragma once

#include <hip/hip_runtime.h>

#include "nppe_internal.h"

namespace nppe
{

constexpr uint8_t kFluid = 0, kEmpty = 2, kPressure = 4;

__device__ __forceinline__ bool is_dirichlet(uint8_t t) { return t == kEmpty || t == kPressure; }

// ---- FP64 operator with the reference's exact arithmetic ----------------------------------------------------
struct Op64
{
    double v[6];   // scale * c[q]: the off-diagonal values
    double dval;   // scale * diag (the diagonal entry; present iff diag != 0)
    double diag;
    unsigned coupled;
};

__device__ __forceinline__ Op64 op64(const Box &B, const Tables64 &T, const uint8_t *ty, int s, int m, int f,
                                     int64_t idx)
{
#pragma clang fp contract(off)
    Op64 o;
    double c[6];
    if (T.unitCoef)
    {
#pragma unroll
        for (int q = 0; q < 6; q++) c[q] = 1.0;
    }
    else if (B.threeD)
    {
        // FaceCoefs3D (PressureSolver3D.cpp:47-63): ay = Width(X,j)*Width(Z,k), ax = Width(Y,i)*Width(Z,k),
        // az = Width(X,j)*Width(Y,i); slow = Y, mid = X, fast = Z.
        const double ay = T.wM[m] * T.wF[f];
        const double ax = T.wS[s] * T.wF[f];
        const double az = T.wM[m] * T.wS[s];
        c[0] = ay / T.gS[s];
        c[1] = ay / T.gS[s - 1];
        c[2] = ax / T.gM[m];
        c[3] = ax / T.gM[m - 1];
        c[4] = az / T.gF[f];
        c[5] = az / T.gF[f - 1];
    }
    else
    {
        // PressureSolver2D.cpp:125-128; mid = Y (i), fast = X (j).
        c[0] = c[1] = 0.0;
        c[2] = T.wF[f] / T.gM[m];
        c[3] = T.wF[f] / T.gM[m - 1];
        c[4] = T.wM[m] / T.gF[f];
        c[5] = T.wM[m] / T.gF[f - 1];
    }
    const int64_t off[6] = {B.sS, -B.sS, B.sM, -B.sM, 1, -1};
    const bool onRing[6] = {s + 1 == B.ns - 1, s - 1 == 0, m + 1 == B.nm - 1, m - 1 == 0, f + 1 == B.nf - 1,
                            f - 1 == 0};
    uint8_t t[6];
#pragma unroll
    for (int q = 0; q < 6; q++) t[q] = (q < 2 && !B.threeD) ? (uint8_t)1 : ty[idx + off[q]];
    double diag = 0.0;
    unsigned cp = 0;
#pragma unroll
    for (int q = 0; q < 6; q++)
        if (t[q] == kFluid)
        {
            diag -= c[q];
            if (!onRing[q]) cp |= 1u << q;
        }
#pragma unroll
    for (int q = 0; q < 6; q++)
        if (is_dirichlet(t[q]))
        {
            double rf = 1.0;
            if (onRing[q])
            {
                // DirichletRingFactor (PressureSettings.h:307-310): FaceGap on the neighbour's side / half width.
                const double *g = q < 2 ? T.gS : (q < 4 ? T.gM : T.gF);
                const double *w = q < 2 ? T.wS : (q < 4 ? T.wM : T.wF);
                const int n = q < 2 ? s : (q < 4 ? m : f);
                rf = g[(q & 1) ? n - 1 : n] / (0.5 * w[n]);
            }
            diag -= c[q] * rf;
        }
#pragma unroll
    for (int q = 0; q < 6; q++) o.v[q] = T.scale * c[q];
    o.diag = diag;
    o.dval = T.scale * diag;
    o.coupled = cp;
    return o;
}

// Row q of A times x, summed in ascending column order from 0.0 as the CSR product does
// (Eigen SparseDenseProduct, row-major: tmp += value * x(col)). X(offset) fetches x at idx + offset.
template <class X>
__device__ __forceinline__ double row64(const Box &B, const Op64 &o, int64_t idx, X xat)
{
#pragma clang fp contract(off)
    double t = 0.0;
    if (o.coupled & 2u) t += o.v[1] * xat(-B.sS);
    if (o.coupled & 8u) t += o.v[3] * xat(-B.sM);
    if (o.coupled & 32u) t += o.v[5] * xat(-1);
    if (o.diag != 0.0) t += o.dval * xat(0);
    if (o.coupled & 16u) t += o.v[4] * xat(1);
    if (o.coupled & 4u) t += o.v[2] * xat(B.sM);
    if (o.coupled & 1u) t += o.v[0] * xat(B.sS);
    (void)idx;
    return t;
}

// (A x)_idx in FP64 with the reference's exact arithmetic. A unit-coefficient (3D uniform) cell with no Dirichlet
// neighbour has off-diagonal scale and diagonal -n * scale, summed in the same column order, so it skips the
// coefficient work; every other cell goes through op64.
template <class X>
__device__ __forceinline__ double apply64(const Box &B, const Tables64 &T, const uint8_t *ty, int s, int m, int f,
                                          int64_t idx, X xat)
{
#pragma clang fp contract(off)
    if (T.unitCoef)
    {
        const unsigned code = T.c8[idx];
        if (!(code & 128u))
        {
            const double v = T.scale * 1.0;
            const double dv = T.scale * (-(double)__popc(code & 63u));
            double t = 0.0;
            if (code & 2u) t += v * xat(-B.sS);
            if (code & 8u) t += v * xat(-B.sM);
            if (code & 32u) t += v * xat(-1);
            if (code & 63u) t += dv * xat(0);
            if (code & 16u) t += v * xat(1);
            if (code & 4u) t += v * xat(B.sM);
            if (code & 1u) t += v * xat(B.sS);
            return t;
        }
    }
    const Op64 o = op64(B, T, ty, s, m, f, idx);
    return row64(B, o, idx, xat);
}

}  // namespace nppe
