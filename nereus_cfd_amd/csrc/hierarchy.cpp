// Host-side hierarchy construction; see hierarchy.h.
#include "hierarchy.h"

#include <algorithm>
#include <cmath>
#include <queue>

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

void stencil_codes(const HostLevel &L, std::vector<uint8_t> &code8, std::vector<uint16_t> &code16)
{
    const int ns = L.ns, nm = L.nm, nf = L.nf;
    const size_t cells = (size_t)ns * nm * nf;
    code8.assign(cells, 0);
    code16.assign(cells, 0);
    auto at = [&](int s, int m, int f) { return ((size_t)s * nm + m) * nf + f; };
    const int s0 = L.threeD ? 1 : 0, s1 = L.threeD ? ns - 2 : 0;
    const int ds[6] = {1, -1, 0, 0, 0, 0}, dm[6] = {0, 0, 1, -1, 0, 0}, df[6] = {0, 0, 0, 0, 1, -1};
    for (int s = s0; s <= s1; s++)
        for (int m = 1; m <= nm - 2; m++)
            for (int f = 1; f <= nf - 2; f++)
            {
                const size_t c = at(s, m, f);
                if (L.types[c] != kFluid) continue;
                uint8_t c8 = 1u << 6;
                uint16_t c16 = 1u << 12;
                for (int q = L.threeD ? 0 : 2; q < 6; q++)
                {
                    const int S = s + ds[q], M = m + dm[q], F = f + df[q];
                    const bool ring = (L.threeD && (S == 0 || S == ns - 1)) || M == 0 || M == nm - 1 || F == 0 ||
                                      F == nf - 1;
                    const uint8_t t = L.types[at(S, M, F)];
                    unsigned st = 0;
                    if (t == kFluid) st = ring ? 2u : 1u;
                    else if (is_dirichlet(t)) st = ring ? 3u : 2u;
                    if (st == 1u) c8 |= (uint8_t)(1u << q);
                    if (st >= 2u) c8 |= (uint8_t)(1u << 7);
                    c16 |= (uint16_t)(st << (2 * q));
                }
                code8[c] = c8;
                code16[c] = c16;
            }
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

bool coarse_pseudo_inverse(const HostLevel &L, double sigma, std::vector<int64_t> &order, std::vector<double> &M,
                           std::string &err)
{
    const int ns = L.ns, nm = L.nm, nf = L.nf;
    auto at = [&](int s, int m, int f) { return ((int64_t)s * nm + m) * nf + f; };
    std::vector<int> id((size_t)ns * nm * nf, -1);
    order.clear();
    const int s0 = L.threeD ? 1 : 0, s1 = L.threeD ? ns - 2 : 0;
    for (int s = s0; s <= s1; s++)
        for (int m = 1; m <= nm - 2; m++)
            for (int f = 1; f <= nf - 2; f++)
                if (L.types[at(s, m, f)] == kFluid)
                {
                    id[at(s, m, f)] = (int)order.size();
                    order.push_back(at(s, m, f));
                }
    const int n = (int)order.size();
    if (n > kMaxDenseCoarse)
    {
        err = "coarsest level too large for the dense solve";
        return false;
    }
    // B = -A (positive semi-definite) in the sigma-scaled volume form of the FP32 levels.
    std::vector<double> B((size_t)n * n, 0.0);
    std::vector<char> touchesDirichlet(n, 0);
    const std::vector<double> &wS = L.w[0], &wM = L.w[1], &wF = L.w[2];
    for (int r = 0; r < n; r++)
    {
        const int64_t c = order[r];
        const int f = (int)(c % nf), m = (int)((c / nf) % nm), s = (int)(c / ((int64_t)nf * nm));
        const double aS = wM[m] * wF[f], aM = wS[s] * wF[f], aF = wS[s] * wM[m];
        auto gap = [](const std::vector<double> &w, int q) { return 0.5 * (w[q] + w[q + 1]); };
        double cq[6] = {0, 0, 0, 0, 0, 0}, ring[6] = {0, 0, 0, 0, 0, 0};
        if (L.threeD)
        {
            cq[0] = aS * sigma / gap(wS, s);
            cq[1] = aS * sigma / gap(wS, s - 1);
            ring[0] = ring[1] = aS * 2.0 * sigma / wS[s];
        }
        cq[2] = aM * sigma / gap(wM, m);
        cq[3] = aM * sigma / gap(wM, m - 1);
        cq[4] = aF * sigma / gap(wF, f);
        cq[5] = aF * sigma / gap(wF, f - 1);
        ring[2] = ring[3] = aM * 2.0 * sigma / wM[m];
        ring[4] = ring[5] = aF * 2.0 * sigma / wF[f];
        const int ds[6] = {1, -1, 0, 0, 0, 0}, dm[6] = {0, 0, 1, -1, 0, 0}, df[6] = {0, 0, 0, 0, 1, -1};
        double diag = 0.0;
        for (int q = 0; q < 6; q++)
        {
            if (q < 2 && !L.threeD) continue;
            const int S = s + ds[q], Mm = m + dm[q], F = f + df[q];
            const bool onRing = (L.threeD && (S == 0 || S == ns - 1)) || Mm == 0 || Mm == nm - 1 || F == 0 || F == nf - 1;
            const uint8_t t = L.types[at(S, Mm, F)];
            if (t == kFluid)
            {
                diag += cq[q];
                const int col = id[at(S, Mm, F)];
                if (!onRing && col >= 0) B[(size_t)r * n + col] -= cq[q];
                else touchesDirichlet[r] = 1;
            }
            else if (is_dirichlet(t))
            {
                diag += onRing ? ring[q] : cq[q];
                touchesDirichlet[r] = 1;
            }
        }
        B[(size_t)r * n + r] += diag;
    }
    // Components of the coupling graph; a component with no Dirichlet face has the constant as its null mode.
    std::vector<int> comp(n, -1);
    std::vector<std::vector<int>> members;
    std::vector<char> singular;
    for (int r = 0; r < n; r++)
    {
        if (comp[r] >= 0) continue;
        const int cid = (int)members.size();
        members.emplace_back();
        singular.push_back(1);
        std::queue<int> q;
        q.push(r);
        comp[r] = cid;
        while (!q.empty())
        {
            const int u = q.front();
            q.pop();
            members[cid].push_back(u);
            if (touchesDirichlet[u]) singular[cid] = 0;
            for (int v = 0; v < n; v++)
                if (v != u && B[(size_t)u * n + v] != 0.0 && comp[v] < 0)
                {
                    comp[v] = cid;
                    q.push(v);
                }
        }
    }
    double dmean = 0.0;
    for (int r = 0; r < n; r++) dmean += B[(size_t)r * n + r];
    dmean = n > 0 ? dmean / n : 1.0;
    if (!(dmean > 0.0)) dmean = 1.0;
    for (size_t cid = 0; cid < members.size(); cid++)
    {
        if (!singular[cid]) continue;
        const double v = dmean / (double)members[cid].size();
        for (int a : members[cid])
            for (int b : members[cid]) B[(size_t)a * n + b] += v;
    }
    // Cholesky B = L L^T, then B^-1 column by column.
    std::vector<double> Lc(B);
    for (int j = 0; j < n; j++)
    {
        double d = Lc[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= Lc[(size_t)j * n + k] * Lc[(size_t)j * n + k];
        if (!(d > 0.0))
        {
            err = "coarsest operator is not positive definite after the null-mode shift";
            return false;
        }
        d = std::sqrt(d);
        Lc[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; i++)
        {
            double v = Lc[(size_t)i * n + j];
            for (int k = 0; k < j; k++) v -= Lc[(size_t)i * n + k] * Lc[(size_t)j * n + k];
            Lc[(size_t)i * n + j] = v / d;
        }
    }
    M.assign((size_t)n * n, 0.0);
    std::vector<double> y(n);
    for (int col = 0; col < n; col++)
    {
        for (int i = 0; i < n; i++)
        {
            double v = i == col ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) v -= Lc[(size_t)i * n + k] * y[k];
            y[i] = v / Lc[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--)
        {
            double v = y[i];
            for (int k = i + 1; k < n; k++) v -= Lc[(size_t)k * n + i] * y[k];
            y[i] = v / Lc[(size_t)i * n + i];
        }
        for (int i = 0; i < n; i++) M[(size_t)i * n + col] = y[i];
    }
    // A^+ = -(B^-1 - sum_c 1_c 1_c^T / (dmean |c|)): remove the shift on each singular component, flip the sign.
    for (size_t cid = 0; cid < members.size(); cid++)
    {
        if (!singular[cid]) continue;
        const double shift = 1.0 / dmean / (double)members[cid].size();
        for (int a : members[cid])
            for (int b : members[cid]) M[(size_t)a * n + b] -= shift;
    }
    for (double &v : M) v = -v;
    return true;
}

}  // namespace nppe
