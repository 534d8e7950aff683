// Host-side hierarchy construction; see hierarchy.h.
#include "hierarchy.h"

#include <algorithm>
#include <cmath>

namespace nppe
{

namespace
{
constexpr uint8_t kFluid = 0, kSolid = 1, kEmpty = 2, kPressure = 4;
inline bool is_dirichlet(uint8_t t) { return t == kEmpty || t == kPressure; }
inline int rank_of(uint8_t t) { return is_dirichlet(t) ? 2 : (t == kFluid ? 1 : 0); }
}  // namespace

Box make_box(int ns, int nm, int nf, bool threeD)
{
    Box b{};
    b.ns = ns;
    b.nm = nm;
    b.nf = nf;
    b.threeD = threeD ? 1 : 0;
    b.pitch = (nf + 15) / 16 * 16;
    b.sM = b.pitch;
    b.sS = (int64_t)nm * b.pitch;
    b.cells = (int64_t)ns * nm * b.pitch;
    b.tilesF = std::max(1, (nf - 2 + kBlockFast - 1) / kBlockFast);
    b.tilesM = std::max(1, (nm - 2 + kBlockMid - 1) / kBlockMid);
    b.planes = threeD ? std::max(1, ns - 2) : 1;
    return b;
}

HostLevel fine_level(int dim, int nx, int ny, int nz, const uint8_t *types, const double *wx, const double *wy,
                     const double *wz)
{
    HostLevel L;
    L.threeD = dim == 3;
    if (L.threeD)
    {
        // slow = Y (i), mid = X (j), fast = Z (k): the MAC index (i*Nx + j)*Nz + k is already (s*nm + m)*nf + f.
        L.ns = ny;
        L.nm = nx;
        L.nf = nz;
        L.w[0].assign(wy, wy + ny);
        L.w[1].assign(wx, wx + nx);
        L.w[2].assign(wz, wz + nz);
    }
    else
    {
        // 2D: one slow plane, mid = Y (i), fast = X (j): index i*Nx + j.
        L.ns = 1;
        L.nm = ny;
        L.nf = nx;
        L.w[0].assign(1, 1.0);
        L.w[1].assign(wy, wy + ny);
        L.w[2].assign(wx, wx + nx);
    }
    const size_t cells = (size_t)L.ns * L.nm * L.nf;
    L.types.assign(types, types + cells);
    const int s0 = L.threeD ? 1 : 0, s1 = L.threeD ? L.ns - 2 : 0;
    for (int s = s0; s <= s1; s++)
        for (int m = 1; m <= L.nm - 2; m++)
            for (int f = 1; f <= L.nf - 2; f++)
                if (L.types[((size_t)s * L.nm + m) * L.nf + f] == kFluid) L.unknowns++;
    double w0 = L.w[1].empty() ? 0.0 : L.w[1][0];
    bool uni = true;
    for (int a = L.threeD ? 0 : 1; a < 3; a++)
        for (double v : L.w[a]) uni = uni && v == w0;
    L.uniform = uni;
    return L;
}

namespace
{
// Children of coarse index C along an axis with fine interior extent n (fine ring at 0 and n+1).
inline void children(int C, int mc, int n, int &lo, int &hi)
{
    if (C == 0) { lo = hi = 0; return; }
    if (C == mc + 1) { lo = hi = n + 1; return; }
    lo = 2 * C - 1;
    hi = std::min(2 * C, n);
}

std::vector<double> coarse_widths(const std::vector<double> &w)
{
    const int n = (int)w.size() - 2, mc = (n + 1) / 2;
    std::vector<double> c(mc + 2);
    for (int C = 1; C <= mc; C++)
    {
        const int lo = 2 * C - 1, hi = std::min(2 * C, n);
        double s = 0.0;
        for (int q = lo; q <= hi; q++) s += w[q];
        c[C] = s;
    }
    c[0] = c[1];
    c[mc + 1] = c[mc];
    return c;
}
}  // namespace

HostLevel coarsen(const HostLevel &f)
{
    HostLevel c;
    c.threeD = f.threeD;
    const int nsi = f.ns - 2, nmi = f.nm - 2, nfi = f.nf - 2;
    const int msc = f.threeD ? (nsi + 1) / 2 : 0, mmc = (nmi + 1) / 2, mfc = (nfi + 1) / 2;
    c.ns = f.threeD ? msc + 2 : 1;
    c.nm = mmc + 2;
    c.nf = mfc + 2;
    c.w[0] = f.threeD ? coarse_widths(f.w[0]) : f.w[0];
    c.w[1] = coarse_widths(f.w[1]);
    c.w[2] = coarse_widths(f.w[2]);
    c.types.assign((size_t)c.ns * c.nm * c.nf, kSolid);
    for (int S = 0; S < c.ns; S++)
    {
        int slo = 0, shi = 0;
        if (f.threeD) children(S, msc, nsi, slo, shi);
        for (int M = 0; M < c.nm; M++)
        {
            int mlo, mhi;
            children(M, mmc, nmi, mlo, mhi);
            for (int F = 0; F < c.nf; F++)
            {
                int flo, fhi;
                children(F, mfc, nfi, flo, fhi);
                int best = -1;
                uint8_t bt = kSolid;
                for (int s = slo; s <= shi; s++)
                    for (int m = mlo; m <= mhi; m++)
                        for (int q = flo; q <= fhi; q++)
                        {
                            const uint8_t t = f.types[((size_t)s * f.nm + m) * f.nf + q];
                            const int r = rank_of(t);
                            if (r > best) { best = r; bt = t; }
                        }
                c.types[((size_t)S * c.nm + M) * c.nf + F] = is_dirichlet(bt) ? kEmpty : (bt == kFluid ? kFluid : kSolid);
            }
        }
    }
    const int s0 = c.threeD ? 1 : 0, s1 = c.threeD ? c.ns - 2 : 0;
    for (int s = s0; s <= s1; s++)
        for (int m = 1; m <= c.nm - 2; m++)
            for (int q = 1; q <= c.nf - 2; q++)
                if (c.types[((size_t)s * c.nm + m) * c.nf + q] == kFluid) c.unknowns++;
    double w0 = c.w[1][0];
    bool uni = true;
    for (int a = c.threeD ? 0 : 1; a < 3; a++)
        for (double v : c.w[a]) uni = uni && v == w0;
    c.uniform = uni;
    return c;
}

std::vector<float> tables32(const HostLevel &L, double sigma, int off[9])
{
    std::vector<float> out;
    for (int a = 0; a < 3; a++)
    {
        off[a] = (int)out.size();
        for (double v : L.w[a]) out.push_back((float)v);
    }
    for (int a = 0; a < 3; a++)
    {
        off[3 + a] = (int)out.size();
        const std::vector<double> &w = L.w[a];
        const int N = (int)w.size();
        for (int n = 0; n < std::max(1, N - 1); n++)
            out.push_back(N >= 2 ? (float)(sigma / (0.5 * (w[n] + w[n + 1]))) : 0.f);
    }
    for (int a = 0; a < 3; a++)
    {
        off[6 + a] = (int)out.size();
        for (double v : L.w[a]) out.push_back((float)(2.0 * sigma / v));
    }
    return out;
}

std::vector<double> tables64(const HostLevel &L, int off[6])
{
    std::vector<double> out;
    for (int a = 0; a < 3; a++)
    {
        off[a] = (int)out.size();
        for (double v : L.w[a]) out.push_back(v);
    }
    for (int a = 0; a < 3; a++)
    {
        off[3 + a] = (int)out.size();
        const std::vector<double> &w = L.w[a];
        const int N = (int)w.size();
        // MAC::GapTable (MAC.cpp:165-175) for the faces m = 0 .. N-2 (no clamping needed there).
        for (int m = 0; m < std::max(1, N - 1); m++) out.push_back(N >= 2 ? 0.5 * (w[m] + w[m + 1]) : 0.0);
    }
    return out;
}

}  // namespace nppe
