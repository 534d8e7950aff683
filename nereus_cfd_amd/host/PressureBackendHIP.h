// PressureBackendHIP.h -- C++17 host side of the MI355X pressure backend, shaped like the reference's backend layer
// (src/Methods/PPESolver/PressureSettings.h:483-624) so that NEREUS can route PressureBackend = HIP through it:
//
//   CreatePressureBackend   (PressureSettings.h:487-514) -> CreatePressureBackendHIP
//   DestroyPressureBackend  (PressureSettings.h:516-538) -> DestroyPressureBackendHIP
//   CommitPressureMatrix    (PressureSettings.h:542-574) -> CommitPressureMatrixHIP   (reads the MAC grid, not triplets)
//   SolvePressureSystem     (PressureSettings.h:578-624) -> SolvePressureSystemHIP    (returns iterations)
//   ReadEigenSettings       (PressureSettings.h:212-255) -> ReadHipSettings           (the hip_* keys)
//
// Behaviour mirrored from the reference: the solve returns its iteration count and a non-converged solve warns once
// on stderr and still returns (PressureSettings.h:472-479); vectors are FP64 in IDP order; a zero-size system
// returns 0 iterations. Errors of the C ABI (negative codes) throw std::runtime_error, the C++ analogue of
// AMGX_SAFE_CALL's abort. Header-only; link with libnppe.so.
#ifndef NEREUS_PRESSURE_BACKEND_HIP_H
#define NEREUS_PRESSURE_BACKEND_HIP_H

#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdlib>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nppe.h"

namespace nereus_hip
{

// The backend's PRESSURE-block keys. They carry a hip_ prefix: an AMGX build passes every key that is not eigen_*
// verbatim to AMGX (PressureSettings.h:158-173), so AmgxConfigString must skip hip_* keys as it skips eigen_*.
struct HipSettings
{
    int method = NPPE_METHOD_MGPCG;   // hip_solver = MGPCG | CG
    double tolerance = 1e-8;          // hip_tolerance
    int maxIters = 2000;              // hip_max_iters
    int preSmooth = NPPE_SMOOTH_AUTO;   // hip_pre_smooth (AUTO or a count)
    int postSmooth = NPPE_SMOOTH_AUTO;  // hip_post_smooth
    double innerReduction = 3e-5;     // hip_inner_reduction
    int krylov = NPPE_KRYLOV_FP64;    // hip_krylov = FP64 | REFINE
    int device = 0;                   // hip_device
};

inline std::string Trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// The PRESSURE block as (key, VALUE) entries: entries end at a newline or a comma, '#' starts a comment, an entry
// without '=' is skipped; keys are trimmed, values trimmed and upper-cased (values are case-insensitive).
inline std::vector<std::pair<std::string, std::string>> PressureEntries(const std::string &text)
{
    std::vector<std::pair<std::string, std::string>> entries;
    size_t begin = 0;
    while (begin <= text.size())
    {
        size_t end = text.find_first_of(",\n", begin);
        if (end == std::string::npos) end = text.size();
        std::string entry = text.substr(begin, end - begin);
        entry.erase(std::min(entry.size(), entry.find('#')));
        const size_t eq = entry.find('=');
        if (eq != std::string::npos)
        {
            std::string value = Trim(entry.substr(eq + 1));
            for (char &c : value) c = (char)std::toupper((unsigned char)c);
            entries.emplace_back(Trim(entry.substr(0, eq)), value);
        }
        begin = end + 1;
    }
    return entries;
}

// The hip_* keys of the PRESSURE block (the other backends' keys are theirs). A key this backend does not know, a
// word it does not know or a number that does not parse (or is out of range) keeps the default and says so on
// stderr; reading never throws.
inline HipSettings ReadHipSettings(const std::string &text)
{
    HipSettings s;
    auto ignored = [](const std::string &key, const std::string &value, const char *why) {
        std::cerr << "[nereus_hip] PRESSURE " << key << " = '" << value << "' ignored: " << why << std::endl;
    };
    auto real = [](const std::string &v, double &out) {
        char *end = nullptr;
        const double d = std::strtod(v.c_str(), &end);
        if (v.empty() || *end != '\0') return false;
        out = d;
        return true;
    };
    auto whole = [](const std::string &v, int &out) {
        char *end = nullptr;
        const long d = std::strtol(v.c_str(), &end, 10);
        if (v.empty() || *end != '\0' || d < INT32_MIN || d > INT32_MAX) return false;
        out = (int)d;
        return true;
    };
    // pre/post smoothing: a sweep count or AUTO
    auto sweeps = [&](const std::string &v, int &out) {
        if (v == "AUTO")
        {
            out = NPPE_SMOOTH_AUTO;
            return true;
        }
        return whole(v, out);
    };
    for (const auto &kv : PressureEntries(text))
    {
        const std::string &key = kv.first, &v = kv.second;
        if (key.compare(0, 4, "hip_") != 0) continue;
        bool ok = true;
        if (key == "hip_solver")
        {
            if (v == "MGPCG" || v == "CG") s.method = v == "CG" ? NPPE_METHOD_CG64 : NPPE_METHOD_MGPCG;
            else ignored(key, v, "expected MGPCG or CG; MGPCG kept");
        }
        else if (key == "hip_krylov")
        {
            if (v == "FP64" || v == "REFINE") s.krylov = v == "REFINE" ? NPPE_KRYLOV_REFINE : NPPE_KRYLOV_FP64;
            else ignored(key, v, "expected FP64 or REFINE; FP64 kept");
        }
        else if (key == "hip_tolerance") ok = real(v, s.tolerance);
        else if (key == "hip_inner_reduction") ok = real(v, s.innerReduction);
        else if (key == "hip_max_iters") ok = whole(v, s.maxIters);
        else if (key == "hip_device") ok = whole(v, s.device);
        else if (key == "hip_pre_smooth") ok = sweeps(v, s.preSmooth);
        else if (key == "hip_post_smooth") ok = sweeps(v, s.postSmooth);
        else ignored(key, v, "not a key of this backend");
        if (!ok) ignored(key, v, "not a number; default kept");
    }
    // values outside their ranges fall back to the defaults
    if (!(s.tolerance > 0.0)) s.tolerance = 1e-8;
    if (s.maxIters <= 0) s.maxIters = 2000;
    if (s.preSmooth < 1 && s.preSmooth != NPPE_SMOOTH_AUTO) s.preSmooth = NPPE_SMOOTH_AUTO;
    if (s.postSmooth < 0 && s.postSmooth != NPPE_SMOOTH_AUTO) s.postSmooth = NPPE_SMOOTH_AUTO;
    if (!(s.innerReduction > 0.0 && s.innerReduction < 1.0)) s.innerReduction = 3e-5;
    return s;
}

// The handle PressureSystem would hold next to its AMGXSolver* (PressureSettings.h:295): raw new/delete, as the
// reference stores its AMGX handle.
struct HipPressureSolver
{
    nppe_solver *h = nullptr;
    int64_t n = 0;
    bool singular = false;
    bool warned = false;
    nppe_result last{};
};

inline void Check(int rc, const nppe_solver *h, const char *what)
{
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + nppe_last_error(h));
}

inline HipPressureSolver *CreatePressureBackendHIP(const HipSettings &s, bool needsCompatibility)
{
    nppe_options o;
    nppe_default_options(&o);
    o.device = s.device;
    o.method = s.method;
    o.tolerance = s.tolerance;
    o.max_iters = s.maxIters;
    o.pre_smooth = s.preSmooth;
    o.post_smooth = s.postSmooth;
    o.inner_reduction = s.innerReduction;
    o.krylov = s.krylov;
    o.compatibility = needsCompatibility ? 1 : 0;
    HipPressureSolver *p = new HipPressureSolver();
    const int rc = nppe_create(&p->h, &o);
    if (rc != NPPE_OK)
    {
        delete p;
        Check(rc, nullptr, "CreatePressureBackendHIP");
    }
    return p;
}

inline void DestroyPressureBackendHIP(HipPressureSolver *&p)
{
    if (!p) return;
    nppe_destroy(p->h);
    delete p;
    p = nullptr;
}

// The MAC grid as CommitPressureMatrix sees it through its MAC* argument: MAC::Nx/Ny/Nz (ring included),
// MAC::CellTypes() (MAC.h:167), MAC::Width(a, n) for every n (MAC.h:78), nonUniform and dh. Returns the number of
// unknowns, which must equal the row count the reference assembled (PressureSolver3D.cpp:133-139).
inline int64_t CommitPressureMatrixHIP(HipPressureSolver *p, int dim, int Nx, int Ny, int Nz, bool nonUniform,
                                       double dh, const std::vector<unsigned char> &cellTypes,
                                       const std::vector<double> &widthX, const std::vector<double> &widthY,
                                       const std::vector<double> &widthZ)
{
    nppe_grid g{};
    g.dim = dim;
    g.nx = Nx;
    g.ny = Ny;
    g.nz = dim == 2 ? 1 : Nz;
    g.non_uniform = nonUniform ? 1 : 0;
    g.dh = dh;
    g.cell_types = cellTypes.data();
    g.width_x = widthX.data();
    g.width_y = widthY.data();
    g.width_z = dim == 3 ? widthZ.data() : nullptr;
    const int64_t n = nppe_commit_grid(p->h, &g);
    Check((int)(n < 0 ? n : 0), p->h, "CommitPressureMatrixHIP");
    int32_t sing = 0;
    nppe_system_info(p->h, &p->n, &sing);
    p->singular = sing != 0;
    return n;
}

// x = A^-1 b to tol * max(||b||, referenceNorm), from guess when not null; x is resized to n. Returns iterations.
inline int SolvePressureSystemHIP(HipPressureSolver *p, const std::vector<double> &b, std::vector<double> &x,
                                  const std::vector<double> *guess, double referenceNorm)
{
    x.resize((size_t)p->n);
    if (p->n == 0) return 0;
    if ((int64_t)b.size() != p->n) throw std::runtime_error("SolvePressureSystemHIP: b has the wrong size");
    const int rc = nppe_solve(p->h, b.data(), x.data(), guess ? guess->data() : nullptr, referenceNorm, &p->last);
    Check(rc, p->h, "SolvePressureSystemHIP");
    if (rc == NPPE_NOT_CONVERGED && !p->warned)
    {
        p->warned = true;
        std::cerr << "Warning: the HIP pressure solve did not converge (relative residual "
                  << p->last.relative_residual << " after " << p->last.iterations
                  << " iterations). Reported once." << std::endl;
    }
    return p->last.iterations;
}

// SolveLaplacian (PressureSolver3D.cpp:355-371): L x = f on the committed system, solved as A x = V f from zero.
// False without solving when the system has no unknowns or f has the wrong size, as the reference returns.
inline bool SolveLaplacianHIP(HipPressureSolver *p, const std::vector<double> &f, std::vector<double> &x)
{
    if (p->n == 0 || (int64_t)f.size() != p->n) return false;
    x.resize((size_t)p->n);
    const int rc = nppe_solve_laplacian(p->h, f.data(), x.data(), &p->last);
    Check(rc, p->h, "SolveLaplacianHIP");
    if (rc == NPPE_NOT_CONVERGED && !p->warned)
    {
        p->warned = true;
        std::cerr << "Warning: the HIP Laplacian solve did not converge (relative residual "
                  << p->last.relative_residual << "). Reported once." << std::endl;
    }
    return true;
}

// The pressure step of SolveScheme after its first EnforcePressureBC (PressureSolver3D.cpp:279-336) on MAC fields in
// device memory: b from the velocity, the reference norm when d_pOld is given (incremental), the solve, the scatter
// into d_p with the ring BC and the incremental ring fix. Returns the iterations (the step's counter).
inline int PressureStepHIP(HipPressureSolver *p, const double *d_u, const double *d_v, const double *d_w,
                           double *d_p, const double *d_pOld, double rho, double dt)
{
    const int rc = nppe_pressure_step_device(p->h, d_u, d_v, d_w, d_p, d_pOld, rho, dt, &p->last);
    Check(rc, p->h, "PressureStepHIP");
    if (rc == NPPE_NOT_CONVERGED && !p->warned)
    {
        p->warned = true;
        std::cerr << "Warning: the HIP pressure solve did not converge (relative residual "
                  << p->last.relative_residual << " after " << p->last.iterations
                  << " iterations). Reported once." << std::endl;
    }
    return p->last.iterations;
}

// The projection step (Project, PressureSolver3D.cpp:396-401) on device-resident MAC fields: ProjectPressure, then
// with d_pOld AccumulatePressureIncrement; d_prescribed (nullable) holds MAC::PrescribedPressure at the PRESSURE
// ring cells.
inline void ProjectionStepHIP(HipPressureSolver *p, double *d_u, double *d_v, double *d_w, double *d_p,
                              const double *d_pOld, double rho, double dt, const double *d_prescribed)
{
    Check(nppe_projection_step_device(p->h, d_u, d_v, d_w, d_p, d_pOld, rho, dt, d_prescribed), p->h,
          "ProjectionStepHIP");
}

}  // namespace nereus_hip

#endif
