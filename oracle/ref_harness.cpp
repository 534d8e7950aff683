// ref_harness: drives the reference's own pressure path, compiled from its sources (oracle/Makefile, target `ref`),
// to write golden vectors and to time the reference's CPU solvers. Test infrastructure only; never linked into the
// product library.
//
//   ref_harness <case.bin> <out.bin>
//
// The case names a grid the way a NEREUS config does (dimension, GRID_SIZE, domain, NON_UNIFORM_GRID and refinement
// intervals), the cell types, and a list of solves. The harness builds the MAC grid with the reference's
// MAC::InitializeGrid (MAC.cpp:275-328), sets the cell types, assembles with
// PressureSolver{2D,3D}::InitializePressureSolver (PressureSolver3D.cpp:89-102, :128-178) and solves with the
// reference's SolvePressureSystem (PressureSettings.h:578-624) for each requested eigen_solver / preconditioner /
// tolerance. It writes the spacing tables, the unknown numbering, the CSR matrix, and each solution with its
// iteration count and wall times. Given velocity and pressure fields it also writes the pressure equation's
// right-hand side built from them with the reference's MAC::GetDivergencyAt and DirichletSum
// (PressureSolver3D.cpp:288-298 / PressureSolver2D.cpp:262-271), IncrementReferenceNorm when a previous
// pressure is given, and the fields after the reference's ProjectPressure. It then runs MAC::EnforcePressureBC
// with a known boundary pressure, the tail of SolveScheme after the solve (PressureSolver3D.cpp:316-336,
// PressureSolver2D.cpp:287-304) on a given vector standing for the solution, and, with a previous pressure,
// AccumulatePressureIncrement (PressureSolver3D.cpp:374-393 / PressureSolver2D.cpp:340-357), writing p after each.

#include <omp.h>
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "Methods/PPESolver/PressureSettings.h"
#include "headers/Core/MAC.h"
#include "headers/Core/Definitions.h"
#include "headers/Core/Scene.h"
#include "headers/Core/SimulationManager.h"
// The solvers keep their system in private statics; the harness reads them.
#define private public
#include "Methods/PPESolver/PressureSolver2D.h"
#include "Methods/PPESolver/PressureSolver3D.h"
#undef private
#include "TestMethods.h"

namespace
{
struct Reader
{
    FILE *f;
    template <class T> T get()
    {
        T v;
        if (std::fread(&v, sizeof(T), 1, f) != 1) { std::fprintf(stderr, "ref_harness: short case file\n"); std::exit(2); }
        return v;
    }
    template <class T> void get(T *p, size_t n)
    {
        if (n && std::fread(p, sizeof(T), n, f) != n) { std::fprintf(stderr, "ref_harness: short case file\n"); std::exit(2); }
    }
};

struct Writer
{
    FILE *f;
    template <class T> void put(T v) { std::fwrite(&v, sizeof(T), 1, f); }
    template <class T> void put(const T *p, size_t n) { if (n) std::fwrite(p, sizeof(T), n, f); }
};

struct SolveSpec
{
    int solver, precond, maxIters, repeats, hasGuess;
    double tol, refNorm;
    Eigen::VectorXd guess;
};

MAC GRID;

// A boundary pressure with a value on every ring cell, for MAC::PrescribedPressure (MAC.cpp:603-611).
double BoundaryP3(double x, double y, double z, double t) { return 0.25 + x - 2.0 * y + 0.5 * z + 3.0 * t; }
double BoundaryP2(double x, double y, double t) { return 0.25 + x - 2.0 * y + 3.0 * t; }

void WriteAxis(Writer &w, const MAC &g, MAC::Axis a, int N)
{
    for (int n = 0; n < N; n++) w.put<double>(g.Width(a, n));
    for (int n = -1; n < N; n++) w.put<double>(g.FaceGap(a, n));
}
} // namespace

int main(int argc, char **argv)
{
    if (argc != 3) { std::fprintf(stderr, "usage: ref_harness <case.bin> <out.bin>\n"); return 2; }
    FILE *fi = std::fopen(argv[1], "rb");
    if (!fi) { std::fprintf(stderr, "ref_harness: cannot open %s\n", argv[1]); return 2; }
    Reader r{fi};
    char magic[8];
    r.get(magic, 8);
    if (std::memcmp(magic, "NPPECASE", 8) != 0) { std::fprintf(stderr, "ref_harness: bad magic\n"); return 2; }

    const int dim = r.get<int32_t>();
    const int gridSize = r.get<int32_t>();
    double dom[6];
    r.get(dom, 6);
    const int nonUniform = r.get<int32_t>();
    int nrefs[3];
    r.get(nrefs, 3);
    std::vector<GridRefinement> refs[3];
    for (int a = 0; a < 3; a++)
        for (int q = 0; q < nrefs[a]; q++)
        {
            GridRefinement g;
            g.active = true;
            g.start = r.get<double>();
            g.end = r.get<double>();
            g.factor = r.get<int32_t>();
            g.mode = r.get<int32_t>();
            refs[a].push_back(g);
        }

    DIMENSION = dim;
    SelectMethod("ADI");
    GPU_ACCELERATION = 0;
    SIMULATION.domain = Domain{dom[0], dom[1], dom[2], dom[3], dom[4], dom[5]};
    SIMULATION.GRID_SIZE = gridSize;
    SIMULATION.RHO = 1.0;
    SIMULATION.dt = 1e-3;
    SIMULATION.time = 0.0;
    SIMULATION.grid.nonUniform = nonUniform != 0;
    SIMULATION.grid.refineX = refs[0];
    SIMULATION.grid.refineY = refs[1];
    SIMULATION.grid.refineZ = refs[2];
    PRESSURE_SETTINGS.warmStart = false;

    GRID = MAC();
    GRID.InitializeGrid(SIMULATION.domain, dim == 2);
    const int Nx = GRID.Nx, Ny = GRID.Ny, Nz = dim == 2 ? 1 : GRID.Nz;
    const size_t cells = (size_t)Nx * Ny * Nz;

    FILE *fo = std::fopen(argv[2], "wb");
    if (!fo) { std::fprintf(stderr, "ref_harness: cannot write %s\n", argv[2]); return 2; }
    Writer w{fo};
    w.put("NPPEOUT1", 8);
    w.put<int32_t>(Nx);
    w.put<int32_t>(Ny);
    w.put<int32_t>(Nz);
    w.put<int32_t>(GRID.nonUniform ? 1 : 0);
    w.put<double>(GRID.dh);
    WriteAxis(w, GRID, MAC::X, Nx);
    WriteAxis(w, GRID, MAC::Y, Ny);
    if (dim == 3) WriteAxis(w, GRID, MAC::Z, Nz);

    const int hasTypes = r.get<int32_t>();
    if (!hasTypes) { std::fclose(fo); return 0; }
    const int32_t tNx = r.get<int32_t>(), tNy = r.get<int32_t>(), tNz = r.get<int32_t>();
    if (tNx != Nx || tNy != Ny || tNz != Nz)
    {
        std::fprintf(stderr, "ref_harness: types are %dx%dx%d, the grid is %dx%dx%d\n", tNx, tNy, tNz, Nx, Ny, Nz);
        return 3;
    }
    std::vector<uint8_t> types(cells);
    r.get(types.data(), cells);
    for (int i = 0; i < Ny; i++)
        for (int j = 0; j < Nx; j++)
            for (int k = 0; k < Nz; k++)
            {
                const int t = types[((size_t)i * Nx + j) * Nz + k];
                if (dim == 2) GRID.SetSolid(i, j, t);
                else GRID.SetSolid(i, j, k, t);
            }
    PRESSURE_SETTINGS.needsCompatibilityCondition = r.get<int32_t>() != 0;

    int n = 0, nnz = 0, singular = 0;
    const RowMatrix *A = nullptr;
    PressureSystem *sys = nullptr;
    if (dim == 2)
    {
        PressureSolver2D::InitializePressureSolver(&GRID, PRESSURE_CHORIN);
        sys = &PressureSolver2D::SYSTEM;
    }
    else
    {
        PressureSolver3D::InitializePressureSolver(&GRID, PRESSURE_CHORIN);
        sys = &PressureSolver3D::SYSTEM;
    }
    A = &sys->A;
    n = (int)A->rows();
    nnz = (int)A->nonZeros();
    singular = sys->singular ? 1 : 0;
    w.put<int32_t>(n);
    w.put<int32_t>(nnz);
    w.put<int32_t>(singular);
    // The unknowns in matrix order, as box indices (i*Nx + j)*Nz + k.
    std::vector<int32_t> order(n, -1);
    for (int i = 0; i < Ny; i++)
        for (int j = 0; j < Nx; j++)
            for (int k = 0; k < Nz; k++)
            {
                const int id = dim == 2 ? PressureSolver2D::GetIDP(i, j) : PressureSolver3D::GetIDP(i, j, k);
                if (id >= 0) order[id] = (int32_t)(((size_t)i * Nx + j) * Nz + k);
            }

    const int nsolves = r.get<int32_t>();
    std::vector<SolveSpec> specs(nsolves);
    for (SolveSpec &s : specs)
    {
        s.solver = r.get<int32_t>();
        s.precond = r.get<int32_t>();
        s.tol = r.get<double>();
        s.maxIters = r.get<int32_t>();
        s.refNorm = r.get<double>();
        s.repeats = std::max(1, (int)r.get<int32_t>());
        s.hasGuess = r.get<int32_t>();
    }
    Eigen::VectorXd b(n);
    if (nsolves > 0) r.get(b.data(), (size_t)n);
    // the guesses of the solves that have one, in order
    for (SolveSpec &s : specs)
        if (s.hasGuess)
        {
            s.guess.resize(n);
            r.get(s.guess.data(), (size_t)n);
        }
    // optional fields section
    const int hasFields = r.get<int32_t>();
    double rho = 1.0, dt = 1e-3;
    int hasOld = 0;
    const size_t nu = (size_t)Ny * (Nx + 1) * Nz, nv = (size_t)(Ny + 1) * Nx * Nz,
                 nw = dim == 3 ? (size_t)Ny * Nx * (Nz + 1) : 0;
    std::vector<double> fu, fv, fw, fp, fpOld, xTail;
    double tBC = 0.0;
    if (hasFields)
    {
        rho = r.get<double>();
        dt = r.get<double>();
        hasOld = r.get<int32_t>();
        fu.resize(nu);
        fv.resize(nv);
        fw.resize(nw);
        fp.resize(cells);
        r.get(fu.data(), nu);
        r.get(fv.data(), nv);
        if (dim == 3) r.get(fw.data(), nw);
        r.get(fp.data(), cells);
        if (hasOld)
        {
            fpOld.resize(cells);
            r.get(fpOld.data(), cells);
        }
        tBC = r.get<double>();
        xTail.resize((size_t)n);
        r.get(xTail.data(), (size_t)n);
    }
    // optional extras: a scene (inlet faces with p= levels) for LevelPressure on a given vector, and a previous
    // pressure for BuildWarmStartGuess
    int hasExtras = 0;
    std::string sceneText;
    double tLevel = 0.0;
    std::vector<double> xLevel, pPrev;
    int hasWarm = 0;
    if (std::fread(&hasExtras, sizeof(int32_t), 1, fi) != 1) hasExtras = 0;
    if (hasExtras)
    {
        const int len = r.get<int32_t>();
        sceneText.resize((size_t)len);
        if (len) r.get(&sceneText[0], (size_t)len);
        tLevel = r.get<double>();
        xLevel.resize((size_t)n);
        r.get(xLevel.data(), (size_t)n);
        hasWarm = r.get<int32_t>();
        if (hasWarm)
        {
            pPrev.resize(cells);
            r.get(pPrev.data(), cells);
        }
    }
    // output flags (timing runs of large systems): bit 0 leaves the unknown order and the matrix out of the output,
    // bit 1 the solutions (iterations and times are always written)
    int outFlags = 0;
    if (std::fread(&outFlags, sizeof(int32_t), 1, fi) != 1) outFlags = 0;
    std::fclose(fi);
    if (!(outFlags & 1))
    {
        w.put(order.data(), order.size());
        w.put(A->outerIndexPtr(), (size_t)n + 1);
        w.put(A->innerIndexPtr(), (size_t)nnz);
        w.put(A->valuePtr(), (size_t)nnz);
    }

    w.put<int32_t>(nsolves);
    for (const SolveSpec &s : specs)
    {
        sys->eigen.solver = s.solver;
        sys->eigen.preconditioner = s.precond;
        sys->eigen.tolerance = s.tol;
        sys->eigen.maxIters = s.maxIters;
        Eigen::VectorXd x;
        int iters = 0;
        std::vector<double> times;
        for (int rep = 0; rep < s.repeats; rep++)
        {
            const double t0 = omp_get_wtime();
            iters = SolvePressureSystem(*sys, b, x, s.hasGuess ? &s.guess : nullptr, s.refNorm);
            times.push_back(omp_get_wtime() - t0);
        }
        w.put<int32_t>(iters);
        w.put<int32_t>((int32_t)times.size());
        w.put(times.data(), times.size());
        if (!(outFlags & 2)) w.put(x.data(), (size_t)n);
    }
    w.put<int32_t>(omp_get_max_threads());

    if (hasFields)
    {
        SIMULATION.RHO = rho;
        SIMULATION.dt = dt;
        // load the MAC fields through the public setters (MAC.h:235-246, :369-377)
        MAC OLD = GRID;
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j <= Nx; j++)
                for (int k = 0; k < Nz; k++)
                    if (dim == 2) GRID.SetU(i, j, fu[(size_t)i * (Nx + 1) + j]);
                    else GRID.SetU(i, j, k, fu[((size_t)i * (Nx + 1) + j) * Nz + k]);
        for (int i = 0; i <= Ny; i++)
            for (int j = 0; j < Nx; j++)
                for (int k = 0; k < Nz; k++)
                    if (dim == 2) GRID.SetV(i, j, fv[(size_t)i * Nx + j]);
                    else GRID.SetV(i, j, k, fv[((size_t)i * Nx + j) * Nz + k]);
        if (dim == 3)
            for (int i = 0; i < Ny; i++)
                for (int j = 0; j < Nx; j++)
                    for (int k = 0; k <= Nz; k++) GRID.SetW(i, j, k, fw[((size_t)i * Nx + j) * (Nz + 1) + k]);
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j < Nx; j++)
                for (int k = 0; k < Nz; k++)
                {
                    const size_t c = ((size_t)i * Nx + j) * Nz + k;
                    if (dim == 2)
                    {
                        GRID.SetP(i, j, fp[c]);
                        if (hasOld) OLD.SetP(i, j, fpOld[c]);
                    }
                    else
                    {
                        GRID.SetP(i, j, k, fp[c]);
                        if (hasOld) OLD.SetP(i, j, k, fpOld[c]);
                    }
                }
        // the right-hand side: the expression of SolveScheme with the reference's own GetDivergencyAt and
        // DirichletSum; MatrixScale and the cell volume are the solvers' one-line helpers
        const double scale = GRID.nonUniform ? 1.0 : 1.0 / (GRID.dh * GRID.dh);
        const double rhoDt = SIMULATION.RHO / SIMULATION.dt;
        Eigen::VectorXd rb(n);
        for (int q = 0; q < n; q++)
        {
            const int c = order[q];
            const int k = c % Nz, j = (c / Nz) % Nx, i = c / (Nz * Nx);
            if (dim == 2)
                rb(q) = rhoDt * GRID.GetDivergencyAt(i, j) * (GRID.nonUniform ? GRID.CellVolume(i, j) : 1.0) -
                        scale * PressureSolver2D::DirichletSum(&GRID, hasOld ? &OLD : nullptr, i, j);
            else
                rb(q) = rhoDt * GRID.GetDivergencyAt(i, j, k) * (GRID.nonUniform ? GRID.CellVolume(i, j, k) : 1.0) -
                        scale * PressureSolver3D::DirichletSum(&GRID, hasOld ? &OLD : nullptr, i, j, k);
        }
        w.put(rb.data(), (size_t)n);
        double ref = 0.0;
        if (hasOld)
            ref = dim == 2 ? PressureSolver2D::IncrementReferenceNorm(&OLD, rb)
                           : PressureSolver3D::IncrementReferenceNorm(&OLD, rb);
        w.put<double>(ref);
        if (dim == 2) PressureSolver2D::ProjectPressure(&GRID);
        else PressureSolver3D::ProjectPressure(&GRID);
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j <= Nx; j++)
                for (int k = 0; k < Nz; k++) w.put<double>(dim == 2 ? GRID.GetU(i, j) : GRID.GetU(i, j, k));
        for (int i = 0; i <= Ny; i++)
            for (int j = 0; j < Nx; j++)
                for (int k = 0; k < Nz; k++) w.put<double>(dim == 2 ? GRID.GetV(i, j) : GRID.GetV(i, j, k));
        if (dim == 3)
            for (int i = 0; i < Ny; i++)
                for (int j = 0; j < Nx; j++)
                    for (int k = 0; k <= Nz; k++) w.put<double>(GRID.GetW(i, j, k));

        auto putP = [&](MAC &g) {
            for (int i = 0; i < Ny; i++)
                for (int j = 0; j < Nx; j++)
                    for (int k = 0; k < Nz; k++) w.put<double>(dim == 2 ? g.GetP(i, j) : g.GetP(i, j, k));
        };
        SIMULATION.time = tBC;
        SIMULATION.PressureBoundaryFunction = BoundaryP3;
        SIMULATION.PressureBoundaryFunction2D = BoundaryP2;
        GRID.EnforcePressureBC(tBC);
        putP(GRID);
        // SolveScheme's tail with xTail as the solve's answer: level (Chorin), scatter, ring BC, ring fix
        // (incremental)
        Eigen::VectorXd x = Eigen::Map<Eigen::VectorXd>(xTail.data(), n);
        if (!hasOld)
        {
            if (dim == 2) PressureSolver2D::LevelPressure(&GRID, x);
            else PressureSolver3D::LevelPressure(&GRID, x);
        }
        for (int q = 0; q < n; q++)
        {
            const int c = order[q];
            const int k = c % Nz, j = (c / Nz) % Nx, i = c / (Nz * Nx);
            if (dim == 2) GRID.SetP(i, j, x(q));
            else GRID.SetP(i, j, k, x(q));
        }
        GRID.EnforcePressureBC(tBC);
        if (hasOld)
            for (int i = 0; i < Ny; i++)
                for (int j = 0; j < Nx; j++)
                    for (int k = 0; k < Nz; k++)
                        if (dim == 2 ? GRID.GetSolid(i, j) == PRESSURE_CELL : GRID.GetSolid(i, j, k) == PRESSURE_CELL)
                        {
                            if (dim == 2) GRID.SetP(i, j, GRID.GetP(i, j) - OLD.GetP(i, j));
                            else GRID.SetP(i, j, k, GRID.GetP(i, j, k) - OLD.GetP(i, j, k));
                        }
        putP(GRID);
        if (hasOld)
        {
            if (dim == 2) PressureSolver2D::AccumulatePressureIncrement(&OLD, &GRID);
            else PressureSolver3D::AccumulatePressureIncrement(&OLD, &GRID);
            putP(GRID);
        }
    }
    if (hasExtras)
    {
        // LevelPressure (PressureSolver{2D,3D}.cpp) with the scene's inlet levels: the per-cell level the scene gives
        // at each ring cell's centre (NaN where it gives none), the centre tables, and the levelled vector
        const bool loaded = !sceneText.empty() && SCENE.LoadText(sceneText, "<harness>");
        if (!sceneText.empty() && !loaded) std::fprintf(stderr, "ref_harness: scene: %s\n", SCENE.error.c_str());
        w.put<int32_t>(loaded ? 1 : 0);
        SIMULATION.time = tLevel;
        for (int i = 0; i < Ny; i++)
            for (int j = 0; j < Nx; j++)
                for (int k = 0; k < Nz; k++)
                {
                    double lv = std::numeric_limits<double>::quiet_NaN();
                    const bool ring = i == 0 || i == Ny - 1 || j == 0 || j == Nx - 1 ||
                                      (dim == 3 && (k == 0 || k == Nz - 1));
                    double v = 0.0;
                    if (loaded && ring &&
                        SCENE.PressureLevelAt(GRID.CentrePos(MAC::X, j), GRID.CentrePos(MAC::Y, i),
                                              dim == 3 ? GRID.CentrePos(MAC::Z, k) : 0.0, tLevel, v))
                        lv = v;
                    w.put<double>(lv);
                }
        for (int j = 0; j < Nx; j++) w.put<double>(GRID.CentrePos(MAC::X, j));
        for (int i = 0; i < Ny; i++) w.put<double>(GRID.CentrePos(MAC::Y, i));
        if (dim == 3)
            for (int k = 0; k < Nz; k++) w.put<double>(GRID.CentrePos(MAC::Z, k));
        Eigen::VectorXd xl = Eigen::Map<Eigen::VectorXd>(xLevel.data(), n);
        if (dim == 2) PressureSolver2D::LevelPressure(&GRID, xl);
        else PressureSolver3D::LevelPressure(&GRID, xl);
        w.put(xl.data(), (size_t)n);
        // BuildWarmStartGuess from a previous pressure, with warm_start on under a fixed-region method (ADI)
        w.put<int32_t>(hasWarm);
        if (hasWarm)
        {
            MAC ANT = GRID;
            for (int i = 0; i < Ny; i++)
                for (int j = 0; j < Nx; j++)
                    for (int k = 0; k < Nz; k++)
                    {
                        const double v = pPrev[((size_t)i * Nx + j) * Nz + k];
                        if (dim == 2) ANT.SetP(i, j, v);
                        else ANT.SetP(i, j, k, v);
                    }
            PRESSURE_SETTINGS.warmStart = true;
            Eigen::VectorXd x0 = Eigen::VectorXd::Zero(n);
            const bool built = dim == 2 ? PressureSolver2D::BuildWarmStartGuess(&ANT, x0)
                                        : PressureSolver3D::BuildWarmStartGuess(&ANT, x0);
            PRESSURE_SETTINGS.warmStart = false;
            if (!built) x0.setZero();   // no guess: the vector is not meaningful
            w.put<int32_t>(built ? 1 : 0);
            w.put(x0.data(), (size_t)n);
        }
    }
    std::fclose(fo);
    return 0;
}
