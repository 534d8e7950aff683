// Seeded synthetic right-hand sides (include/nppe_synth.h).
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/nppe_synth.h"

namespace
{
double uniform01(std::mt19937_64 &g) { return (double)(g() >> 11) * 0x1.0p-53; }

// Centre of every cell along one axis scaled to [0, 1] over the interior (ring cells get values outside).
std::vector<double> centres(const double *w, int N)
{
    std::vector<double> c(N, 0.0);
    double L = 0.0;
    for (int q = 1; q <= N - 2; q++) L += w[q];
    double acc = 0.0;
    for (int q = 1; q <= N - 2; q++)
    {
        c[q] = (acc + 0.5 * w[q]) / L;
        acc += w[q];
    }
    return c;
}
}  // namespace

extern "C" int64_t nppe_synth_cosine_rhs(const nppe_grid *g, uint64_t seed, int32_t modes, double noise,
                                         int32_t zero_mean, double *b)
{
    if (!g || !b || !g->cell_types || (g->dim != 2 && g->dim != 3) || modes < 0) return NPPE_ERR_ARGUMENT;
    const int dim = g->dim, Nx = g->nx, Ny = g->ny, Nz = dim == 2 ? 1 : g->nz;
    std::mt19937_64 gen(seed);
    std::vector<double> a(modes);
    std::vector<int> k((size_t)modes * dim);
    for (int m = 0; m < modes; m++)
    {
        a[m] = 2.0 * uniform01(gen) - 1.0;
        for (int d = 0; d < dim; d++) k[(size_t)m * dim + d] = 1 + (int)(4.0 * uniform01(gen));
    }
    const std::vector<double> cx = centres(g->width_x, Nx), cy = centres(g->width_y, Ny);
    const std::vector<double> cz = dim == 3 ? centres(g->width_z, Nz) : std::vector<double>(1, 0.0);
    const double pi = 3.14159265358979323846;
    int64_t n = 0;
    bool haveSpare = false;
    double spare = 0.0;
    auto normal = [&]() {
        if (haveSpare)
        {
            haveSpare = false;
            return spare;
        }
        double u1 = uniform01(gen), u2 = uniform01(gen);
        if (u1 <= 0.0) u1 = 0x1.0p-53;
        const double r = std::sqrt(-2.0 * std::log(u1));
        spare = r * std::sin(2.0 * pi * u2);
        haveSpare = true;
        return r * std::cos(2.0 * pi * u2);
    };
    const int k0 = dim == 3 ? 1 : 0, k1 = dim == 3 ? Nz - 2 : 0;
    for (int i = 1; i <= Ny - 2; i++)
        for (int j = 1; j <= Nx - 2; j++)
            for (int kk = k0; kk <= k1; kk++)
            {
                if (g->cell_types[((int64_t)i * Nx + j) * Nz + kk] != NPPE_FLUID) continue;
                const double xs[3] = {cx[j], cy[i], cz[kk]};
                double v = 0.0;
                for (int m = 0; m < modes; m++)
                {
                    double p = a[m];
                    for (int d = 0; d < dim; d++) p *= std::cos(pi * k[(size_t)m * dim + d] * xs[d]);
                    v += p;
                }
                b[n++] = v + noise * normal();
            }
    if (zero_mean && n > 0)
        for (int pass = 0; pass < 2; pass++)
        {
            double s = 0.0;
            for (int64_t q = 0; q < n; q++) s += b[q];
            s /= (double)n;
            for (int64_t q = 0; q < n; q++) b[q] -= s;
        }
    return n;
}
