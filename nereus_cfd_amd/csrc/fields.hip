// MAC-field kernels around the pressure solve, for gfx950: the right-hand side from the intermediate velocity
// (PressureSolver3D.cpp:288-298 / PressureSolver2D.cpp:262-271), the projection (PressureSolver3D.cpp:403-438 /
// PressureSolver2D.cpp:370-396), MAC::EnforcePressureBC's ring list (MAC.cpp:866-871, :1064-1096), SolveScheme's
// scatter and ring fix (PressureSolver3D.cpp:316-336), AccumulatePressureIncrement (:374-393), the increment
// reference norm (:262-276) and SolveLaplacian's volume scaling (:355-371).
//
// Every kernel is one pass over HBM-resident fields. The cell kernels give a wavefront one row of the box along the
// fast axis (k in 3D, j in 2D, the contiguous index of every MAC array), so each of the row's loads is a coalesced
// 64-lane read; a ballot over the row's stencil codes numbers the unknowns in IDP order as the wave walks the row.
// Consecutive rows stay on one XCD, where the neighbouring rows a cell reads (j +- 1, i +- 1) are in its L2.
//
// Floating point: the same operations in the same order as the reference, with contraction off, so every value is
// the reference's bit for bit; only the sums (means, norms) reduce in a different order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fields.h"
#include "rowwalk.h"

namespace nppe
{
namespace
{

using namespace rows;

constexpr uint8_t kSolid = 1, kEmpty = 2, kPressure = 4;

__device__ __forceinline__ bool dirichlet(uint8_t t) { return t == kEmpty || t == kPressure; }

// MAC::FaceGap(a, m) = (Width(m) + Width(m+1)) / 2 for an interior m (MAC.cpp:165-175).
__device__ __forceinline__ double face_gap(const double *w, int m) { return 0.5 * (w[m] + w[m + 1]); }

// DirichletRingFactor (PressureSettings.h:307-310): FaceGap / (Width / 2) toward the ring neighbour on side s.
__device__ __forceinline__ double ring_factor(const double *w, int n, int s)
{
    return face_gap(w, s < 0 ? n - 1 : n) / (0.5 * w[n]);
}

__device__ __forceinline__ double known_p(const double *p, const double *pOld, int64_t c)
{
    return pOld ? p[c] - pOld[c] : p[c];
}

// ---- right-hand side ---------------------------------------------------------------------------------------
// NU: the stretched (volume-weighted) form, G.nonUniform.
template <int DIM, bool NU>
__global__ __launch_bounds__(kThreads) void k_rhs(FieldGeom G, const double *__restrict__ u,
                                                 const double *__restrict__ v, const double *__restrict__ w,
                                                 const double *__restrict__ p, const double *__restrict__ pOld,
                                                 double rhoDt, double scale, double *__restrict__ b)
{
#pragma clang fp contract(off)
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    const int Nx = G.Nx, Ny = G.Ny, Nz = G.Nz;
    if (!R.active) return;   // no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    int64_t base = G.rowStart[r];
    for (int f0 = 1; f0 <= R.nfast; f0 += 64)
    {
        // the face values are loaded with the stencil code, before the ballot that finds the unknowns (every
        // interior cell's faces exist), so both loads are in flight together
        const int f = f0 + lane;
        unsigned code = 0u;
        double u0 = 0.0, u1 = 0.0, v0 = 0.0, v1 = 0.0, w0 = 0.0, w1 = 0.0;
        if (f <= R.nfast)
        {
            code = G.c8[R.pitched + f];
            const int i = R.i;
            if (DIM == 3)
            {
                const int j = R.j, k = f;
                u0 = u[((int64_t)i * (Nx + 1) + j) * Nz + k];
                u1 = u[((int64_t)i * (Nx + 1) + j + 1) * Nz + k];
                v0 = v[((int64_t)i * Nx + j) * Nz + k];
                v1 = v[((int64_t)(i + 1) * Nx + j) * Nz + k];
                const int64_t wi = ((int64_t)i * Nx + j) * (Nz + 1) + k;
                w0 = w[wi];
                w1 = w[wi + 1];
            }
            else
            {
                const int j = f;
                u0 = u[(int64_t)i * (Nx + 1) + j];
                u1 = u[(int64_t)i * (Nx + 1) + j + 1];
                v0 = v[(int64_t)i * Nx + j];
                v1 = v[(int64_t)(i + 1) * Nx + j];
            }
        }
        const bool unk = (code & 64u) != 0u;
        const uint64_t m = __ballot(unk);
        const int64_t id = base + __popcll(m & below);
        base += __popcll(m);
        if (!unk) continue;
        double div, vol, ds = 0.0;
        const int64_t pc = R.pitched + f;
        if (DIM == 3)
        {
            const int i = R.i, j = R.j, k = f;
            if (NU)
            {
                // MAC::GetDivergencyAt, stretched: the net face flux over the cell volume (MAC.cpp:353-360)
                const double dx = G.wx[j], dy = G.wy[i], dz = G.wz[k];
                const double fu = (u1 - u0) * dy * dz, fv = (v1 - v0) * dx * dz, fw = (w1 - w0) * dx * dy;
                div = (fu + fv + fw) / (dx * dy * dz);
                vol = G.wx[j] * G.wy[i] * G.wz[k];
            }
            else
            {
                const double du = (u1 - u0) / G.dh, dv = (v1 - v0) / G.dh, dw = (w1 - w0) / G.dh;
                div = du + dv + dw;
                vol = 1.0;
            }
            if (code & 128u)
            {
                // DirichletSum (PressureSolver3D.cpp:248-260): PRESSURE neighbours (i+1, i-1, j+1, j-1, k+1, k-1),
                // FaceCoefs3D times RingFactor3D times the known pressure
                const int64_t off[6] = {G.B.sS, -G.B.sS, G.B.sM, -G.B.sM, 1, -1};
                const int64_t moff[6] = {(int64_t)Nx * Nz, -(int64_t)Nx * Nz, Nz, -Nz, 1, -1};
                const int64_t c = ((int64_t)i * Nx + j) * Nz + k;
#pragma unroll 1
                for (int a = 0; a < 6; a++)
                {
                    if (G.types[pc + off[a]] != kPressure) continue;
                    double cq = 1.0;
                    if (NU)
                    {
                        const double ax = G.wy[i] * G.wz[k], ay = G.wx[j] * G.wz[k], az = G.wx[j] * G.wy[i];
                        cq = a == 0   ? ay / face_gap(G.wy, i)
                             : a == 1 ? ay / face_gap(G.wy, i - 1)
                             : a == 2 ? ax / face_gap(G.wx, j)
                             : a == 3 ? ax / face_gap(G.wx, j - 1)
                             : a == 4 ? az / face_gap(G.wz, k)
                                      : az / face_gap(G.wz, k - 1);
                    }
                    const int sgn = (a & 1) ? -1 : 1;
                    const int na = (a < 2 ? i : (a < 4 ? j : k)) + sgn, Na = a < 2 ? Ny : (a < 4 ? Nx : Nz);
                    double rf = 1.0;
                    if (na == 0 || na == Na - 1)
                        rf = a < 2 ? ring_factor(G.wy, i, sgn) : (a < 4 ? ring_factor(G.wx, j, sgn) : ring_factor(G.wz, k, sgn));
                    ds += cq * rf * known_p(p, pOld, c + moff[a]);
                }
            }
        }
        else
        {
            const int i = R.i, j = f;
            if (NU)
            {
                const double dx = G.wx[j], dy = G.wy[i];
                const double fu = (u1 - u0) * dy, fv = (v1 - v0) * dx;
                div = (fu + fv) / (dx * dy);
                vol = G.wx[j] * G.wy[i];
            }
            else
            {
                const double du = (u1 - u0) / G.dh, dv = (v1 - v0) / G.dh;
                div = du + dv;
                vol = 1.0;
            }
            if (code & 128u)
            {
                // DirichletSum 2D (PressureSolver2D.cpp:223-235): N, S, E, W
                const int64_t off[4] = {G.B.sM, -G.B.sM, 1, -1};
                const int64_t moff[4] = {Nx, -Nx, 1, -1};
                const int64_t c = (int64_t)i * Nx + j;
#pragma unroll 1
                for (int a = 0; a < 4; a++)
                {
                    if (G.types[pc + off[a]] != kPressure) continue;
                    const double cf = a < 2 ? G.wx[j] / face_gap(G.wy, a == 0 ? i : i - 1)
                                            : G.wy[i] / face_gap(G.wx, a == 2 ? j : j - 1);
                    const int sgn = (a & 1) ? -1 : 1;
                    const int na = (a < 2 ? i : j) + sgn, Na = a < 2 ? Ny : Nx;
                    double rf = 1.0;
                    if (na == 0 || na == Na - 1) rf = a < 2 ? ring_factor(G.wy, i, sgn) : ring_factor(G.wx, j, sgn);
                    ds += cf * rf * known_p(p, pOld, c + moff[a]);
                }
            }
        }
        b[id] = rhoDt * div * vol - scale * ds;
    }
}

// ---- projection --------------------------------------------------------------------------------------------
// ProjectionGradient (PressureSolver3D.cpp:405-414): the face gradient over FaceGap, times DirichletRingFactor when
// the face's outer cell is a Dirichlet ring cell. n: the face's index along the axis, N: the axis length.
// The ring cells' types (t[ring0], t[ringN]) are read only for the first and last face of the axis.
__device__ __forceinline__ double proj_grad(double pc, double pm, const double *w, int N, int n,
                                            const uint8_t *__restrict__ t, int64_t ring0, int64_t ringN)
{
#pragma clang fp contract(off)
    const double grad = (pc - pm) / face_gap(w, n - 1);
    if (n == 1 && dirichlet(t[ring0])) return grad * ring_factor(w, 1, -1);
    if (n == N - 1 && dirichlet(t[ringN])) return grad * ring_factor(w, N - 2, +1);
    return grad;
}

// One thread per cell (i, j, k) of [1, Ny-1] x [1, Nx-1] x [1, Nz-1]: the u, v and w faces on its low sides
// (those inside the reference's loop ranges). Each face has one owner, and only p is read across cells.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_project(FieldGeom G, double *__restrict__ u, double *__restrict__ v,
                                                     double *__restrict__ w, const double *__restrict__ p, double fac,
                                                     int rows)
{
#pragma clang fp contract(off)
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= rows) return;   // no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const int Nx = G.Nx, Ny = G.Ny, Nz = G.Nz;
    const uint8_t *t = G.types;
    if (DIM == 3)
    {
        const int i = 1 + r / (Nx - 1), j = 1 + r % (Nx - 1);
        // the faces k = 1 .. Nz-2 one per lane; the last w face (k = Nz-1) goes with the lane of k = Nz-2, so a
        // row of Nz - 2 = 64 m cells takes m steps, not m + 1
        for (int k = 1 + lane; k <= Nz - 2; k += 64)
        {
            const int64_t c = ((int64_t)i * Nx + j) * Nz + k, cp = G.B.at(i, j, k);
            const uint8_t tc = t[cp];
            const double pc = p[c];
            if (k == Nz - 2 && i <= Ny - 2 && j <= Nx - 2 && t[cp + 1] != kSolid && tc != kSolid)
            {
                const double gr = proj_grad(p[c + 1], pc, G.wz, Nz, Nz - 1, t, G.B.at(i, j, 0), cp + 1);
                double *wf = &w[((int64_t)i * Nx + j) * (Nz + 1) + k + 1];
                *wf = *wf - fac * gr;
            }
            if (i <= Ny - 2 && tc != kSolid && t[cp - G.B.sM] != kSolid)
            {
                const double gr = proj_grad(pc, p[c - Nz], G.wx, Nx, j, t, G.B.at(i, 0, k), G.B.at(i, Nx - 1, k));
                double *uf = &u[((int64_t)i * (Nx + 1) + j) * Nz + k];
                *uf = *uf - fac * gr;
            }
            if (j <= Nx - 2 && tc != kSolid && t[cp - G.B.sS] != kSolid)
            {
                const double gr =
                    proj_grad(pc, p[c - (int64_t)Nx * Nz], G.wy, Ny, i, t, G.B.at(0, j, k), G.B.at(Ny - 1, j, k));
                double *vf = &v[((int64_t)i * Nx + j) * Nz + k];
                *vf = *vf - fac * gr;
            }
            if (i <= Ny - 2 && j <= Nx - 2 && tc != kSolid && t[cp - 1] != kSolid)
            {
                const double gr = proj_grad(pc, p[c - 1], G.wz, Nz, k, t, G.B.at(i, j, 0), G.B.at(i, j, Nz - 1));
                double *wf = &w[((int64_t)i * Nx + j) * (Nz + 1) + k];
                *wf = *wf - fac * gr;
            }
        }
    }
    else
    {
        const int i = 1 + r;
        // the faces j = 1 .. Nx-2 one per lane; the last u face (j = Nx-1) goes with the lane of j = Nx-2
        for (int j = 1 + lane; j <= Nx - 2; j += 64)
        {
            const int64_t c = (int64_t)i * Nx + j, cp = G.B.at(0, i, j);
            const uint8_t tc = t[cp];
            const double pc = p[c];
            if (j == Nx - 2 && i <= Ny - 2 && t[cp + 1] != kSolid && tc != kSolid)
            {
                const double gr = proj_grad(p[c + 1], pc, G.wx, Nx, Nx - 1, t, G.B.at(0, i, 0), cp + 1);
                u[(int64_t)i * (Nx + 1) + j + 1] = u[(int64_t)i * (Nx + 1) + j + 1] - fac * gr;
            }
            if (i <= Ny - 2 && tc != kSolid && t[cp - 1] != kSolid)
            {
                const double gr = proj_grad(pc, p[c - 1], G.wx, Nx, j, t, G.B.at(0, i, 0), G.B.at(0, i, Nx - 1));
                u[(int64_t)i * (Nx + 1) + j] = u[(int64_t)i * (Nx + 1) + j] - fac * gr;
            }
            if (tc != kSolid && t[cp - G.B.sM] != kSolid)
            {
                const double gr = proj_grad(pc, p[c - Nx], G.wy, Ny, i, t, G.B.at(0, 0, j), G.B.at(0, Ny - 1, j));
                v[(int64_t)i * Nx + j] = v[(int64_t)i * Nx + j] - fac * gr;
            }
        }
    }
}

// ---- pressure boundary conditions ------------------------------------------------------------------------------
// One pass of the ring list (RecordRing, MAC.h:482-486): the ring cells on both sides of one axis. Within a pass no
// destination is another entry's source (every axis has at least three cells), so its entries are independent;
// the passes run in the list's order, and a later pass reads what an earlier one wrote (edges and corners).
// Thread q covers cell (a, b) of the face plane, on side q & 1.
__global__ __launch_bounds__(256) void k_ring_pass(FieldGeom G, int axis, double *__restrict__ p,
                                                   const double *__restrict__ prescribed, int64_t count)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int side = (int)(q & 1);
    const int64_t e = q >> 1;
    const int Nx = G.Nx, Ny = G.Ny, Nz = G.Nz;
    int i, j, k, si, sj, sk;
    if (G.dim == 3)
    {
        if (axis == 0)        // y faces: (0 | Ny-1, j, k), j slow, k fast
        {
            j = (int)(e / Nz); k = (int)(e % Nz);
            i = side ? Ny - 1 : 0; si = side ? Ny - 2 : 1; sj = j; sk = k;
        }
        else if (axis == 1)   // x faces: (i, 0 | Nx-1, k)
        {
            i = (int)(e / Nz); k = (int)(e % Nz);
            j = side ? Nx - 1 : 0; sj = side ? Nx - 2 : 1; si = i; sk = k;
        }
        else                  // z faces: (i, j, 0 | Nz-1)
        {
            i = (int)(e / Nx); j = (int)(e % Nx);
            k = side ? Nz - 1 : 0; sk = side ? Nz - 2 : 1; si = i; sj = j;
        }
    }
    else
    {
        k = sk = 0;
        if (axis == 0) { j = (int)e; i = side ? Ny - 1 : 0; si = side ? Ny - 2 : 1; sj = j; }
        else { i = (int)e; j = side ? Nx - 1 : 0; sj = side ? Nx - 2 : 1; si = i; }
    }
    const int64_t dst = ((int64_t)i * Nx + j) * Nz + k, src = ((int64_t)si * Nx + sj) * Nz + sk;
    const uint8_t t = G.dim == 3 ? G.types[G.B.at(i, j, k)] : G.types[G.B.at(0, i, j)];
    if (t == kEmpty) p[dst] = 0.0;
    else if (t == kPressure) { if (prescribed) p[dst] = prescribed[dst]; }
    else p[dst] = p[src];
}

// ---- scatter, increment, gathers ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_scatter_p(FieldGeom G, const double *__restrict__ x,
                                                       const double *__restrict__ sumX, double *__restrict__ p)
{
#pragma clang fp contract(off)
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    const double mean = sumX ? *sumX / (double)G.n : 0.0;
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t id, unsigned) {
#pragma clang fp contract(off)
        const int64_t c = DIM == 3 ? ((int64_t)R.i * G.Nx + R.j) * G.Nz + f : (int64_t)R.i * G.Nx + f;
        p[c] = sumX ? x[id] - mean : x[id];
    });
}

// Every cell of the box, one wave per (i, j) row (2D: per i row), lanes along the fast axis. RESTORE = false:
// p -= pOld at every PRESSURE cell; RESTORE = true: p = pOld + p at the PRESSURE cells of the ring only.
template <int DIM, bool RESTORE>
__global__ __launch_bounds__(kThreads) void k_ring_fix(FieldGeom G, const double *__restrict__ pOld,
                                                      double *__restrict__ p, int rows)
{
#pragma clang fp contract(off)
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int lane = threadIdx.x & 63;
    const int nf = DIM == 3 ? G.Nz : G.Nx;
    const int i = DIM == 3 ? r / G.Nx : r, j = DIM == 3 ? r % G.Nx : 0;
    const int64_t pitched = DIM == 3 ? G.B.at(i, j, 0) : G.B.at(0, i, 0);
    const bool ringRow = DIM == 3 ? (i == 0 || i == G.Ny - 1 || j == 0 || j == G.Nx - 1) : (i == 0 || i == G.Ny - 1);
    for (int f = lane; f < nf; f += 64)
        if (G.types[pitched + f] == kPressure)
        {
            const int64_t c = (int64_t)r * nf + f;
            if (!RESTORE) p[c] = p[c] - pOld[c];
            else if (ringRow || f == 0 || f == nf - 1) p[c] = pOld[c] + p[c];
        }
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_sum_increment(FieldGeom G, const double *__restrict__ pOld,
                                                           const double *__restrict__ p, double *__restrict__ partials)
{
#pragma clang fp contract(off)
    const int lb = xcd_block();
    const int r = lb * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    double acc = 0.0;
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t, unsigned) {
#pragma clang fp contract(off)
        const int64_t c = DIM == 3 ? ((int64_t)R.i * G.Nx + R.j) * G.Nz + f : (int64_t)R.i * G.Nx + f;
        acc += pOld[c] + p[c];
    });
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partials[lb] = s;
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_accumulate(FieldGeom G, const double *__restrict__ pOld,
                                                        const double *__restrict__ sum, double *__restrict__ p)
{
#pragma clang fp contract(off)
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    const double mean = sum ? *sum / (double)G.n : 0.0;
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t, unsigned) {
#pragma clang fp contract(off)
        const int64_t c = DIM == 3 ? ((int64_t)R.i * G.Nx + R.j) * G.Nz + f : (int64_t)R.i * G.Nx + f;
        const double v = pOld[c] + p[c];
        p[c] = sum ? v - mean : v;
    });
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_gather_box(FieldGeom G, const double *__restrict__ pOld,
                                                        double *__restrict__ box)
{
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t, unsigned) {
#pragma clang fp contract(off)
        const int64_t c = DIM == 3 ? ((int64_t)R.i * G.Nx + R.j) * G.Nz + f : (int64_t)R.i * G.Nx + f;
        box[R.pitched + f] = pOld[c];
    });
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_norm2_sum(FieldGeom G, const double *__restrict__ b,
                                                       const double *__restrict__ y, double *__restrict__ partials)
{
#pragma clang fp contract(off)
    const int lb = xcd_block();
    const int r = lb * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    double acc = 0.0;
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t id, unsigned) {
#pragma clang fp contract(off)
        const double t = b[id] + y[R.pitched + f];
        acc += t * t;
    });
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partials[lb] = s;
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_scale_volume(FieldGeom G, const double *__restrict__ f_,
                                                          double *__restrict__ b)
{
#pragma clang fp contract(off)
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t id, unsigned) {
#pragma clang fp contract(off)
        // CellVolume3D / CellVolume2D (PressureSolver3D.cpp:39-43): exactly 1.0 on a uniform grid
        double vol = 1.0;
        if (G.nonUniform) vol = DIM == 3 ? G.wx[R.j] * G.wy[R.i] * G.wz[f] : G.wx[f] * G.wy[R.i];
        b[id] = f_[id] * vol;
    });
}

// LevelPressure's inflow shift (PressureSolver3D.cpp:15-30, :181-211; PressureSolver2D.cpp:13-29, :156-178) on
// p at the unknowns, after its mean was removed there: every ring cell of the x, y (3D: and z) faces whose cell is
// INFLOW, whose two inward neighbours are FLUID unknowns and where the scene gives a level (level[c] not NaN) adds
// level - (p1 + (p1 - p2) (sR - s1) / (s1 - s2)), the reference's expression with its operation order (contraction
// off); centres: MAC::CentrePos of the x, y, z axes back to back. Block partials: the sums, then the counts.
constexpr uint8_t kInflow = 3;
constexpr int kLevelBlocks = 256;

template <int DIM>
__global__ __launch_bounds__(kThreads) void k_inflow_level(FieldGeom G, const double *__restrict__ p,
                                                          const double *__restrict__ level,
                                                          const double *__restrict__ centres,
                                                          double *__restrict__ partials)
{
#pragma clang fp contract(off)
    const int Nx = G.Nx, Ny = G.Ny, Nz = G.Nz;
    const double *cx = centres, *cy = centres + Nx, *cz = centres + Nx + Ny;
    // ring-face items: x faces (i, k), y faces (j, k), z faces (i, j), two sides each (2D: x faces i, y faces j)
    const int64_t segX = DIM == 3 ? (int64_t)(Ny - 2) * (Nz - 2) : (int64_t)(Ny - 2);
    const int64_t segY = DIM == 3 ? (int64_t)(Nx - 2) * (Nz - 2) : (int64_t)(Nx - 2);
    const int64_t segZ = DIM == 3 ? (int64_t)(Ny - 2) * (Nx - 2) : 0;
    const int64_t total = 2 * (segX + segY + segZ);
    double sum = 0.0, cnt = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < total; t += (int64_t)gridDim.x * kThreads)
    {
        int64_t u = t >> 1;
        const bool hi = (t & 1) != 0;
        int ri, rj, rk, di = 0, dj = 0, dk = 0, N;
        const double *cen;
        if (u < segX)
        {
            ri = 1 + (int)(DIM == 3 ? u / (Nz - 2) : u);
            rk = DIM == 3 ? 1 + (int)(u % (Nz - 2)) : 0;
            rj = hi ? Nx - 1 : 0;
            dj = hi ? -1 : 1;
            N = Nx;
            cen = cx;
        }
        else if ((u -= segX) < segY)
        {
            rj = 1 + (int)(DIM == 3 ? u / (Nz - 2) : u);
            rk = DIM == 3 ? 1 + (int)(u % (Nz - 2)) : 0;
            ri = hi ? Ny - 1 : 0;
            di = hi ? -1 : 1;
            N = Ny;
            cen = cy;
        }
        else
        {
            u -= segY;
            ri = 1 + (int)(u / (Nx - 2));
            rj = 1 + (int)(u % (Nx - 2));
            rk = hi ? Nz - 1 : 0;
            dk = hi ? -1 : 1;
            N = Nz;
            cen = cz;
        }
        if (N < 4) continue;   // the second inward cell is the far ring cell: no unknown
        const int n0 = di ? ri : (dj ? rj : rk);
        auto typ = [&](int i, int j, int k) { return G.types[G.B.at(DIM == 3 ? i : 0, DIM == 3 ? j : i, DIM == 3 ? k : j)]; };
        auto mac = [&](int i, int j, int k) { return ((int64_t)i * Nx + j) * Nz + k; };
        if (typ(ri, rj, rk) != kInflow || typ(ri + di, rj + dj, rk + dk) != 0 ||
            typ(ri + 2 * di, rj + 2 * dj, rk + 2 * dk) != 0)
            continue;
        const double lv = level[mac(ri, rj, rk)];
        if (lv != lv) continue;   // NaN: the scene sets no level here
        const double p1 = p[mac(ri + di, rj + dj, rk + dk)], p2 = p[mac(ri + 2 * di, rj + 2 * dj, rk + 2 * dk)];
        const int step = di + dj + dk;
        const double sR = cen[n0], s1 = cen[n0 + step], s2 = cen[n0 + 2 * step];
        sum += lv - (p1 + (p1 - p2) * (sR - s1) / (s1 - s2));
        cnt += 1.0;
    }
    const double bs = block_sum(sum);
    __syncthreads();   // block_sum's LDS is used again
    const double bc = block_sum(cnt);
    if (threadIdx.x == 0)
    {
        partials[blockIdx.x] = bs;
        partials[gridDim.x + blockIdx.x] = bc;
    }
}

// p[c] += sum / count at the unknowns when count > 0 (LevelPressure's p.array() += sum / n).
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_add_shift(FieldGeom G, const double *__restrict__ sumCount,
                                                       double *__restrict__ p)
{
#pragma clang fp contract(off)
    const double cnt = sumCount[1];
    if (!(cnt > 0.0)) return;
    const double shift = sumCount[0] / cnt;
    const int r = xcd_block() * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t, unsigned) {
        const int64_t c = DIM == 3 ? ((int64_t)R.i * G.Nx + R.j) * G.Nz + f : (int64_t)R.i * G.Nx + f;
        p[c] += shift;
    });
}

// x[id] = p[c] at the unknowns (IDP order), with the block partials of sum p[c]^2 (BuildWarmStartGuess's gather
// and its zero test).
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_gather_idp(FieldGeom G, const double *__restrict__ p,
                                                        double *__restrict__ x, double *__restrict__ partials)
{
    const int lb = xcd_block();
    const int r = lb * kRowsPerBlock + (threadIdx.x >> 6);
    const Row R = interior_row<DIM>(G, r);
    double acc = 0.0;
    walk_unknowns<DIM>(G, R, r, [&](int f, int64_t id, unsigned) {
        const int64_t c = DIM == 3 ? ((int64_t)R.i * G.Nx + R.j) * G.Nz + f : (int64_t)R.i * G.Nx + f;
        const double v = p[c];
        x[id] = v;
        acc += v * v;
    });
    const double sum = block_sum(acc);
    if (threadIdx.x == 0) partials[lb] = sum;
}

// EnsureSystem's layout comparison (PressureSettings.h:300-303): every cell of the box, one wave per (i, j) row.
template <int DIM>
__global__ __launch_bounds__(kThreads) void k_diff_types(FieldGeom G, const uint8_t *__restrict__ t, int rows,
                                                        double *__restrict__ partials)
{
    const int lb = xcd_block();
    const int r = lb * kRowsPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    double acc = 0.0, bad = 0.0;
    if (r < rows)
    {
        const int nf = DIM == 3 ? G.Nz : G.Nx;
        const int64_t pitched = DIM == 3 ? G.B.at(r / G.Nx, r % G.Nx, 0) : G.B.at(0, r, 0);
        for (int f = lane; f < nf; f += 64)
        {
            const uint8_t v = t[(int64_t)r * nf + f];
            acc += v != G.types[pitched + f] ? 1.0 : 0.0;
            bad += v > 5 ? 1.0 : 0.0;
        }
    }
    const double s = block_sum(acc);
    __syncthreads();   // block_sum's LDS is used again
    const double sb = block_sum(bad);
    if (threadIdx.x == 0)
    {
        partials[lb] = s;
        partials[gridDim.x + lb] = sb;
    }
}

// ---- linear reductions -------------------------------------------------------------------------------------
constexpr int kLinearBlocks = 1024;

__global__ __launch_bounds__(kThreads) void k_sum_linear(const double *__restrict__ x, int64_t n,
                                                         double *__restrict__ partials)
{
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kThreads)
        acc += x[q];
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_sub_mean(double *__restrict__ x, int64_t n,
                                                       const double *__restrict__ sum)
{
    const double mean = *sum / (double)n;
    for (int64_t q = (int64_t)blockIdx.x * kThreads + threadIdx.x; q < n; q += (int64_t)gridDim.x * kThreads)
        x[q] -= mean;
}

__global__ __launch_bounds__(1024) void k_sum_partials(const double *__restrict__ partials, int count,
                                                       double *__restrict__ dst)
{
    __shared__ double red[16];
    double acc = 0.0;
    for (int q = threadIdx.x; q < count; q += 1024) acc += partials[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double s = 0.0;
        for (int q = 0; q < 16; q++) s += red[q];
        *dst = s;
    }
}

int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(kLinearBlocks, (n + kThreads - 1) / kThreads)); }

}  // namespace

int field_row_blocks(const FieldGeom &G) { return std::max(1, (G.rows + kRowsPerBlock - 1) / kRowsPerBlock); }

int field_box_blocks(const FieldGeom &G)
{
    const int rows = G.dim == 3 ? G.Ny * G.Nx : G.Ny;
    return std::max(1, (rows + kRowsPerBlock - 1) / kRowsPerBlock);
}

hipError_t launch_diff_types(hipStream_t s, const FieldGeom &G, const uint8_t *newTypes, double *partials, int *blocks)
{
    const int rows = G.dim == 3 ? G.Ny * G.Nx : G.Ny;
    *blocks = field_box_blocks(G);
    if (G.dim == 3) k_diff_types<3><<<*blocks, kThreads, 0, s>>>(G, newTypes, rows, partials);
    else k_diff_types<2><<<*blocks, kThreads, 0, s>>>(G, newTypes, rows, partials);
    return hipGetLastError();
}

hipError_t launch_rhs(hipStream_t s, const FieldGeom &G, const double *u, const double *v, const double *w,
                      const double *p, const double *pOld, double rhoDt, double scale, double *b)
{
    if (G.n == 0) return hipSuccess;
    const int nb = field_row_blocks(G);
    if (G.dim == 3)
    {
        if (G.nonUniform) k_rhs<3, true><<<nb, kThreads, 0, s>>>(G, u, v, w, p, pOld, rhoDt, scale, b);
        else k_rhs<3, false><<<nb, kThreads, 0, s>>>(G, u, v, w, p, pOld, rhoDt, scale, b);
    }
    else
    {
        if (G.nonUniform) k_rhs<2, true><<<nb, kThreads, 0, s>>>(G, u, v, w, p, pOld, rhoDt, scale, b);
        else k_rhs<2, false><<<nb, kThreads, 0, s>>>(G, u, v, w, p, pOld, rhoDt, scale, b);
    }
    return hipGetLastError();
}

hipError_t launch_project(hipStream_t s, const FieldGeom &G, double *u, double *v, double *w, const double *p,
                          double f)
{
    const int rows = G.dim == 3 ? (G.Ny - 1) * (G.Nx - 1) : G.Ny - 1;
    const int blocks = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
    if (G.dim == 3) k_project<3><<<blocks, kThreads, 0, s>>>(G, u, v, w, p, f, rows);
    else k_project<2><<<blocks, kThreads, 0, s>>>(G, u, v, w, p, f, rows);
    return hipGetLastError();
}

hipError_t launch_pressure_bc(hipStream_t s, const FieldGeom &G, double *p, const double *prescribed)
{
    const int passes = G.dim == 3 ? 3 : 2;
    for (int a = 0; a < passes; a++)
    {
        int64_t plane;
        if (G.dim == 3) plane = a == 0 ? (int64_t)G.Nx * G.Nz : (a == 1 ? (int64_t)G.Ny * G.Nz : (int64_t)G.Ny * G.Nx);
        else plane = a == 0 ? G.Nx : G.Ny;
        const int64_t count = 2 * plane;
        k_ring_pass<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(G, a, p, prescribed, count);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_scatter_p(hipStream_t s, const FieldGeom &G, const double *x, const double *sumX, double *p)
{
    if (G.n == 0) return hipSuccess;
    if (G.dim == 3) k_scatter_p<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, x, sumX, p);
    else k_scatter_p<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, x, sumX, p);
    return hipGetLastError();
}

hipError_t launch_ring_fix(hipStream_t s, const FieldGeom &G, const double *pOld, double *p, bool restore)
{
    const int rows = G.dim == 3 ? G.Ny * G.Nx : G.Ny;
    const int blocks = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
    if (G.dim == 3)
    {
        if (restore) k_ring_fix<3, true><<<blocks, kThreads, 0, s>>>(G, pOld, p, rows);
        else k_ring_fix<3, false><<<blocks, kThreads, 0, s>>>(G, pOld, p, rows);
    }
    else
    {
        if (restore) k_ring_fix<2, true><<<blocks, kThreads, 0, s>>>(G, pOld, p, rows);
        else k_ring_fix<2, false><<<blocks, kThreads, 0, s>>>(G, pOld, p, rows);
    }
    return hipGetLastError();
}

hipError_t launch_sum_increment(hipStream_t s, const FieldGeom &G, const double *pOld, const double *p,
                                double *partials)
{
    if (G.dim == 3) k_sum_increment<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, pOld, p, partials);
    else k_sum_increment<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, pOld, p, partials);
    return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t s, const FieldGeom &G, const double *pOld, const double *sum, double *p)
{
    if (G.n == 0) return hipSuccess;
    if (G.dim == 3) k_accumulate<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, pOld, sum, p);
    else k_accumulate<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, pOld, sum, p);
    return hipGetLastError();
}

hipError_t launch_gather_box(hipStream_t s, const FieldGeom &G, const double *pOld, double *box)
{
    if (G.n == 0) return hipSuccess;
    if (G.dim == 3) k_gather_box<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, pOld, box);
    else k_gather_box<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, pOld, box);
    return hipGetLastError();
}

hipError_t launch_norm2_sum(hipStream_t s, const FieldGeom &G, const double *b, const double *y, double *partials)
{
    if (G.dim == 3) k_norm2_sum<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, b, y, partials);
    else k_norm2_sum<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, b, y, partials);
    return hipGetLastError();
}

hipError_t launch_scale_volume(hipStream_t s, const FieldGeom &G, const double *f, double *b)
{
    if (G.n == 0) return hipSuccess;
    if (G.dim == 3) k_scale_volume<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, f, b);
    else k_scale_volume<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, f, b);
    return hipGetLastError();
}

hipError_t launch_sum_linear(hipStream_t s, const double *x, int64_t n, double *partials, int *blocks)
{
    const int nb = blocks_for(n);
    *blocks = nb;
    k_sum_linear<<<nb, kThreads, 0, s>>>(x, n, partials);
    return hipGetLastError();
}

hipError_t launch_sub_mean(hipStream_t s, double *x, int64_t n, const double *sum)
{
    if (n == 0) return hipSuccess;
    k_sub_mean<<<blocks_for(n), kThreads, 0, s>>>(x, n, sum);
    return hipGetLastError();
}

hipError_t launch_sum_partials(hipStream_t s, const double *partials, int count, double *dst)
{
    k_sum_partials<<<1, 1024, 0, s>>>(partials, count, dst);
    return hipGetLastError();
}


int inflow_level_blocks(const FieldGeom &G)
{
    const int64_t segX = G.dim == 3 ? (int64_t)(G.Ny - 2) * (G.Nz - 2) : (int64_t)(G.Ny - 2);
    const int64_t segY = G.dim == 3 ? (int64_t)(G.Nx - 2) * (G.Nz - 2) : (int64_t)(G.Nx - 2);
    const int64_t segZ = G.dim == 3 ? (int64_t)(G.Ny - 2) * (G.Nx - 2) : 0;
    const int64_t total = 2 * (segX + segY + segZ);
    return (int)std::max<int64_t>(1, std::min<int64_t>(kLevelBlocks, (total + kThreads - 1) / kThreads));
}

hipError_t launch_inflow_level(hipStream_t s, const FieldGeom &G, const double *p, const double *level,
                               const double *centres, double *partials)
{
    const int nb = inflow_level_blocks(G);
    if (G.dim == 3) k_inflow_level<3><<<nb, kThreads, 0, s>>>(G, p, level, centres, partials);
    else k_inflow_level<2><<<nb, kThreads, 0, s>>>(G, p, level, centres, partials);
    return hipGetLastError();
}

hipError_t launch_add_shift(hipStream_t s, const FieldGeom &G, const double *sumCount, double *p)
{
    if (G.dim == 3) k_add_shift<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, sumCount, p);
    else k_add_shift<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, sumCount, p);
    return hipGetLastError();
}

hipError_t launch_gather_idp(hipStream_t s, const FieldGeom &G, const double *p, double *x, double *partials)
{
    if (G.dim == 3) k_gather_idp<3><<<field_row_blocks(G), kThreads, 0, s>>>(G, p, x, partials);
    else k_gather_idp<2><<<field_row_blocks(G), kThreads, 0, s>>>(G, p, x, partials);
    return hipGetLastError();
}

}  // namespace nppe
