// The reference's pressure_dirichlet_check (tests/suite_pressure.cpp:79-232) restated against the HIP backend
// through the C++ adapter: a box whose ring is PRESSURE cells holding p = 0.3 + 1.1x - 0.7y on the boundary faces,
// zero velocity, so b is the Dirichlet term alone; the interior solution is the linear field to machine precision
// and a repeated solve is bit-identical. Also a pure-Neumann box (singular, compatibility on): the result is
// mean-free and meets the tolerance. Prints PASS/FAIL lines; exit code = number of failures.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../nereus_cfd_amd/host/PressureBackendHIP.h"

static int failures = 0;
static void check(bool ok, const std::string &name, const std::string &detail = "")
{
    std::printf("  [%s] %s%s%s\n", ok ? "PASS" : "FAIL", name.c_str(), detail.empty() ? "" : " -- ", detail.c_str());
    if (!ok) failures++;
}

static double Exact(double x, double y) { return 0.3 + 1.1 * x - 0.7 * y; }

static void Dirichlet(int dim, int n)
{
    const double dh = 1.0 / n;
    const int Nx = n + 2, Ny = n + 2, Nz = dim == 3 ? n + 2 : 1;
    std::vector<unsigned char> t((size_t)Nx * Ny * Nz, NPPE_PRESSURE);
    for (int i = 1; i < Ny - 1; i++)
        for (int j = 1; j < Nx - 1; j++)
            for (int k = dim == 3 ? 1 : 0; k < (dim == 3 ? Nz - 1 : 1); k++) t[((size_t)i * Nx + j) * Nz + k] = NPPE_FLUID;
    std::vector<double> w(n + 2, dh);
    auto centre = [&](int q) { return (q - 0.5) * dh; };
    auto face = [&](int q, int N) { return q == 0 ? 0.0 : (q == N - 1 ? 1.0 : centre(q)); };
    // b = -scale * sum over PRESSURE neighbours of c * ringfactor * p_b; uniform: c = 1, ring factor 2.
    std::vector<double> b, exact;
    for (int i = 1; i < Ny - 1; i++)
        for (int j = 1; j < Nx - 1; j++)
            for (int k = dim == 3 ? 1 : 0; k < (dim == 3 ? Nz - 1 : 1); k++)
            {
                double s = 0.0;
                const int nb[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
                for (int q = 0; q < (dim == 3 ? 6 : 4); q++)
                {
                    const int a = nb[q][0], bb = nb[q][1], c = nb[q][2];
                    if (t[((size_t)a * Nx + bb) * Nz + c] == NPPE_PRESSURE) s += 2.0 * Exact(face(bb, Nx), face(a, Ny));
                }
                b.push_back(-(1.0 / (dh * dh)) * s);
                exact.push_back(Exact(centre(j), centre(i)));
            }
    nereus_hip::HipSettings st = nereus_hip::ReadHipSettings("hip_solver = MGPCG, hip_tolerance = 1e-12\n");
    nereus_hip::HipPressureSolver *p = nereus_hip::CreatePressureBackendHIP(st, false);
    const int64_t unknowns = nereus_hip::CommitPressureMatrixHIP(p, dim, Nx, Ny, Nz, false, dh, t, w, w, w);
    check(unknowns == (int64_t)b.size(), "unknowns match the interior fluid cells");
    check(!p->singular, "a prescribed-pressure ring makes the system regular");
    std::vector<double> x1, x2;
    const int it = nereus_hip::SolvePressureSystemHIP(p, b, x1, nullptr, 0.0);
    nereus_hip::SolvePressureSystemHIP(p, b, x2, nullptr, 0.0);
    double worst = 0.0, maxAbs = 0.0, drift = 0.0;
    for (size_t q = 0; q < x1.size(); q++)
    {
        worst = std::max(worst, std::fabs(x1[q] - exact[q]));
        maxAbs = std::max(maxAbs, std::fabs(x1[q]));
        drift = std::max(drift, std::fabs(x1[q] - x2[q]));
    }
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%dD n=%d: max|p - exact| = %.3e (max|p| = %.3f), %d iterations", dim, n, worst,
                  maxAbs, it);
    check(worst < 1e-8, "interior matches the prescribed field", buf);
    check(maxAbs > 0.1, "and the field is genuinely non-zero", buf);
    std::snprintf(buf, sizeof(buf), "drift = %.3e", drift);
    check(drift == 0.0, "repeating the solve is deterministic", buf);
    // SolveLaplacian: on a uniform grid V = 1, so L x = b is the same system (PressureSolver3D.cpp:355-371)
    std::vector<double> xl;
    check(nereus_hip::SolveLaplacianHIP(p, b, xl), "SolveLaplacian solves on the committed system");
    double dl = 0.0;
    for (size_t q = 0; q < xl.size() && q < x1.size(); q++) dl = std::max(dl, std::fabs(xl[q] - x1[q]));
    std::snprintf(buf, sizeof(buf), "max|x_laplacian - x_solve| = %.3e", dl);
    check(xl.size() == x1.size() && dl == 0.0, "and gives the pressure solve's answer bit for bit", buf);
    std::vector<double> shortF(b.size() > 0 ? b.size() - 1 : 0);
    check(!nereus_hip::SolveLaplacianHIP(p, shortF, xl), "SolveLaplacian refuses an f of the wrong size");
    nereus_hip::DestroyPressureBackendHIP(p);
    check(p == nullptr, "destroy clears the handle");
}

static void Neumann(int n)
{
    const double dh = 1.0 / n;
    const int N = n + 2;
    std::vector<unsigned char> t((size_t)N * N * N, NPPE_SOLID);
    std::vector<double> b;
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++)
            for (int k = 1; k < N - 1; k++)
            {
                t[((size_t)i * N + j) * N + k] = NPPE_FLUID;
                b.push_back(std::cos(3.14159265358979 * (j - 0.5) * dh) * std::cos(2 * 3.14159265358979 * (i - 0.5) * dh));
            }
    double mean = 0.0;
    for (double v : b) mean += v;
    mean /= (double)b.size();
    for (double &v : b) v -= mean;
    std::vector<double> w(N, dh);
    nereus_hip::HipPressureSolver *p =
        nereus_hip::CreatePressureBackendHIP(nereus_hip::ReadHipSettings("hip_tolerance=1e-10"), true);
    nereus_hip::CommitPressureMatrixHIP(p, 3, N, N, N, false, dh, t, w, w, w);
    check(p->singular, "a closed box is singular");
    std::vector<double> x;
    nereus_hip::SolvePressureSystemHIP(p, b, x, nullptr, 0.0);
    double sum = 0.0, amax = 0.0;
    for (double v : x)
    {
        sum += v;
        amax = std::max(amax, std::fabs(v));
    }
    char buf[160];
    std::snprintf(buf, sizeof(buf), "relres %.3e, mean %.3e", p->last.relative_residual, sum / (double)x.size());
    check(p->last.relative_residual < 1e-10, "pure Neumann solve meets 1e-10", buf);
    check(std::fabs(sum / (double)x.size()) < 1e-12 * amax, "the solution is mean-free", buf);
    nereus_hip::DestroyPressureBackendHIP(p);
}

int main()
{
    std::printf("pressure_dirichlet_check (HIP backend)\n");
    Dirichlet(3, 20);
    Dirichlet(2, 40);
    Neumann(32);
    nereus_hip::HipSettings bad = nereus_hip::ReadHipSettings("hip_tolerance = -3, hip_solver = CG, eigen_solver = AMG");
    check(bad.tolerance == 1e-8 && bad.method == NPPE_METHOD_CG64, "settings: bad values fall back, other keys ignored");
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures);
    return failures;
}
