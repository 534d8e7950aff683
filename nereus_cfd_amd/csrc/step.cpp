// The pressure step around the solve on MAC fields (include/nppe.h, "the pressure step"): the right-hand side, the
// projection, the pressure boundary conditions, the incremental scheme's reference norm, scatter and accumulation,
// SolveLaplacian, and the two composed steps -- the host side of fields.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "handle.h"
#include "kernels.h"
#include "layout.h"

using namespace nppe;

namespace
{

constexpr int kScalBC = 0;        // dStepScal slots: the two compatibility sums of b
constexpr int kScalMeanX = 2;     // sum of x (LevelPressure of the Chorin scatter)
constexpr int kScalMeanP = 3;     // sum of p_old + phi (LevelPressure of the accumulation)
constexpr int kScalNorm = 4;      // ||b + A p^n||^2
static_assert(kStepScalDiff == 5, "slots 5, 6: the layout comparison's counts (nppe_update_layout_device)");
constexpr int kScalLevel = 7;     // LevelPressure's inflow sum and count (slots 7, 8)
constexpr int kScalWarm = 9;      // the warm start's ||p_prev||^2 at the unknowns
constexpr int kScalWarmMean = 10; // the sum of the warm-start guess (its mean removal)
constexpr int kScalCount = 12;

bool needs_compatibility(const nppe_solver *h) { return h->singular && h->opt.compatibility && h->n > 0; }

int ready(nppe_solver *h)
{
    if (!h) return NPPE_ERR_ARGUMENT;
    if (!h->committed) return fail(h, NPPE_ERR_NO_GRID, "no grid committed");
    HIP_TRY(h, hipSetDevice(h->device), "set device");
    h->err.clear();
    return NPPE_OK;
}

int fields_args(nppe_solver *h, const void *u, const void *v, const void *w, const void *p)
{
    if (!u || !v || !p || (h->dim == 3 && !w)) return fail(h, NPPE_ERR_ARGUMENT, "null field");
    return NPPE_OK;
}

int scalars_args(nppe_solver *h, double rho, double dt)
{
    if (!std::isfinite(rho) || !std::isfinite(dt) || rho == 0.0 || dt == 0.0)
        return fail(h, NPPE_ERR_ARGUMENT, "rho and dt must be finite and non-zero");
    return NPPE_OK;
}

// Element counts of the MAC arrays u, v, w, p.
void field_sizes(const nppe_solver *h, size_t sz[4])
{
    const size_t Nx = h->Nx, Ny = h->Ny, Nz = h->Nz;
    sz[0] = Ny * (Nx + 1) * Nz;
    sz[1] = (Ny + 1) * Nx * Nz;
    sz[2] = h->dim == 3 ? Ny * Nx * (Nz + 1) : 0;
    sz[3] = Ny * Nx * Nz;
}

// b -= mean(b), twice (PressureSolver3D.cpp:300-307, the second pass takes the round-off left by the first).
hipError_t compatibility(nppe_solver *h, double *b)
{
    for (int pass = 0; pass < 2; pass++)
    {
        int nb = 0;
        hipError_t e = launch_sum_linear(h->stream, b, h->n, h->dPartials, &nb);
        if (e == hipSuccess) e = launch_sum_partials(h->stream, h->dPartials, nb, h->dStepScal + kScalBC + pass);
        if (e == hipSuccess) e = launch_sub_mean(h->stream, b, h->n, h->dStepScal + kScalBC + pass);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t rhs(nppe_solver *h, const double *u, const double *v, const double *w, const double *p,
               const double *pOld, double rho, double dt, double *b)
{
    const double rhoDt = rho / dt;
    const double scale = h->nonUniform ? 1.0 : 1.0 / (h->dh * h->dh);
    hipError_t e = launch_rhs(h->stream, h->fg, u, v, w, p, pOld, rhoDt, scale, b);
    if (e == hipSuccess && needs_compatibility(h)) e = compatibility(h, b);
    return e;
}

// LevelPressure's inflow shift (PressureSolver3D.cpp:15-30, :196-211) on p at the unknowns, whose mean is already
// removed: the ring terms reduced in a fixed order, then p += sum / count when any ring cell contributed.
hipError_t level_inflow(nppe_solver *h, double *p, const nppe_inflow_level *lv)
{
    const size_t Nx = h->Nx, Ny = h->Ny, Nz = h->Nz;
    hipError_t e = hipMemcpyAsync(h->dCentre, lv->centre_x, Nx * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->dCentre + Nx, lv->centre_y, Ny * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess && h->dim == 3)
        e = hipMemcpyAsync(h->dCentre + Nx + Ny, lv->centre_z, Nz * sizeof(double), hipMemcpyHostToDevice, h->stream);
    const int nb = inflow_level_blocks(h->fg);
    if (e == hipSuccess) e = launch_inflow_level(h->stream, h->fg, p, lv->d_level, h->dCentre, h->dPartials);
    if (e == hipSuccess) e = launch_sum_partials(h->stream, h->dPartials, nb, h->dStepScal + kScalLevel);
    if (e == hipSuccess) e = launch_sum_partials(h->stream, h->dPartials + nb, nb, h->dStepScal + kScalLevel + 1);
    if (e == hipSuccess) e = launch_add_shift(h->stream, h->fg, h->dStepScal + kScalLevel, p);
    return e;
}

int level_args(nppe_solver *h, const nppe_inflow_level *lv)
{
    if (lv && (!lv->d_level || !lv->centre_x || !lv->centre_y || (h->dim == 3 && !lv->centre_z)))
        return fail(h, NPPE_ERR_ARGUMENT, "inflow level: null level array or centre table");
    return NPPE_OK;
}

hipError_t scatter(nppe_solver *h, const double *x, double *p, const double *pOld, const double *prescribed,
                   const nppe_inflow_level *lv = nullptr)
{
    const double *sumX = nullptr;
    hipError_t e = hipSuccess;
    const bool level = !pOld && needs_compatibility(h);
    if (level)
    {
        int nb = 0;
        e = launch_sum_linear(h->stream, x, h->n, h->dPartials, &nb);
        if (e == hipSuccess) e = launch_sum_partials(h->stream, h->dPartials, nb, h->dStepScal + kScalMeanX);
        sumX = h->dStepScal + kScalMeanX;
    }
    if (e == hipSuccess) e = launch_scatter_p(h->stream, h->fg, x, sumX, p);
    if (e == hipSuccess && level && lv) e = level_inflow(h, p, lv);
    if (e == hipSuccess) e = launch_pressure_bc(h->stream, h->fg, p, prescribed);
    if (e == hipSuccess && pOld) e = launch_ring_fix(h->stream, h->fg, pOld, p);
    return e;
}

hipError_t accumulate(nppe_solver *h, const double *pOld, double *p, const double *prescribed,
                      const nppe_inflow_level *lv = nullptr)
{
    const double *sum = nullptr;
    hipError_t e = hipSuccess;
    if (needs_compatibility(h))
    {
        e = launch_sum_increment(h->stream, h->fg, pOld, p, h->dPartials);
        if (e == hipSuccess)
            e = launch_sum_partials(h->stream, h->dPartials, field_row_blocks(h->fg), h->dStepScal + kScalMeanP);
        sum = h->dStepScal + kScalMeanP;
    }
    if (e == hipSuccess) e = launch_accumulate(h->stream, h->fg, pOld, sum, p);
    if (e == hipSuccess && sum && lv) e = level_inflow(h, p, lv);
    // without prescribed values the PRESSURE ring cells get p_old + phi, p_b(t) to round-off
    if (e == hipSuccess && !prescribed) e = launch_ring_fix(h->stream, h->fg, pOld, p, true);
    if (e == hipSuccess) e = launch_pressure_bc(h->stream, h->fg, p, prescribed);
    return e;
}

// ||b + A p^n||: p^n into the pitched FP64 box at the unknowns, A applied with the reference's arithmetic.
int reference_norm(nppe_solver *h, const double *b, const double *pOld, double *norm)
{
    *norm = 0.0;
    if (h->n == 0) return NPPE_OK;
    const DevLevel &L0 = h->levels[0];
    HIP_TRY(h, launch_gather_box(h->stream, h->fg, pOld, h->dX64[1]), "gather p_old");
    HIP_TRY(h, launch_apply64(h->stream, L0.box, h->t64, L0.types, h->dX64[1], h->dB64), "apply A");
    HIP_TRY(h, launch_norm2_sum(h->stream, h->fg, b, h->dB64, h->dPartials), "reference norm");
    HIP_TRY(h, launch_sum_partials(h->stream, h->dPartials, field_row_blocks(h->fg), h->dStepScal + kScalNorm),
            "reference norm");
    double s = 0.0;
    HIP_TRY(h, hipMemcpyAsync(&s, h->dStepScal + kScalNorm, sizeof(double), hipMemcpyDeviceToHost, h->stream),
            "reference norm");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "reference norm");
    *norm = std::sqrt(s);
    return NPPE_OK;
}

// BuildWarmStartGuess (PressureSolver3D.cpp:222-244; PressureSolver2D.cpp:191-212) once the caller's part of
// WarmStartEnabled holds (warm_start on, a method whose fluid region is fixed): the previous pressure at the
// unknowns, its mean removed when the compatibility condition applies; no guess on a stretched grid
// (WarmStartEnabled, :215-220) or when that pressure is zero at every unknown.
int warm_guess(nppe_solver *h, const double *pPrev, double *guess, bool *built)
{
    *built = false;
    if (h->nonUniform || h->n == 0) return NPPE_OK;
    HIP_TRY(h, launch_gather_idp(h->stream, h->fg, pPrev, guess, h->dPartials), "warm start");
    HIP_TRY(h, launch_sum_partials(h->stream, h->dPartials, field_row_blocks(h->fg), h->dStepScal + kScalWarm),
            "warm start");
    HIP_TRY(h, hipMemcpyAsync(h->hScal + 1, h->dStepScal + kScalWarm, sizeof(double), hipMemcpyDeviceToHost,
                              h->stream),
            "warm start");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "warm start");
    if (h->hScal[1] == 0.0) return NPPE_OK;   // the first step: solve cold
    if (needs_compatibility(h))
    {
        int nb = 0;
        HIP_TRY(h, launch_sum_linear(h->stream, guess, h->n, h->dPartials, &nb), "warm start");
        HIP_TRY(h, launch_sum_partials(h->stream, h->dPartials, nb, h->dStepScal + kScalWarmMean), "warm start");
        HIP_TRY(h, launch_sub_mean(h->stream, guess, h->n, h->dStepScal + kScalWarmMean), "warm start");
    }
    *built = true;
    return NPPE_OK;
}

// Staging for the host-array entry points: u, v, w, p, p_old, allocated on first use and kept with the grid.
int stage(nppe_solver *h)
{
    if (h->dStage[0]) return NPPE_OK;
    size_t sz[4];
    field_sizes(h, sz);
    const size_t n[5] = {sz[0], sz[1], sz[2], sz[3], sz[3]};
    for (int q = 0; q < 5; q++) HIP_TRY(h, dalloc(h, &h->dStage[q], n[q]), "allocate field staging");
    return NPPE_OK;
}

hipError_t put(nppe_solver *h, double *dst, const double *src, size_t count)
{
    return count ? hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, h->stream) : hipSuccess;
}

hipError_t get(nppe_solver *h, double *dst, const double *src, size_t count)
{
    return count ? hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, h->stream) : hipSuccess;
}

}  // namespace

int nppe::build_fields(nppe_solver *h)
{
    const DevLevel &F = h->levels[0];
    const int Nx = h->Nx, Ny = h->Ny, Nz = h->Nz, dim = h->dim;
    // MAC width tables in MAC axis order (the levels keep them in slow/mid/fast order)
    const HostLevel &H = F.host;
    const std::vector<double> &wx = dim == 3 ? H.w[1] : H.w[2];
    const std::vector<double> &wy = dim == 3 ? H.w[0] : H.w[1];
    std::vector<double> tab;
    tab.insert(tab.end(), wx.begin(), wx.end());
    tab.insert(tab.end(), wy.begin(), wy.end());
    if (dim == 3) tab.insert(tab.end(), H.w[2].begin(), H.w[2].end());
    else tab.push_back(h->dh);
    HIP_TRY(h, dalloc(h, &h->dFieldTab, tab.size()), "allocate field tables");
    HIP_TRY(h, upload(h, h->dFieldTab, tab.data(), tab.size() * sizeof(double)), "upload field tables");
    const int rows = dim == 3 ? (Ny - 2) * (Nx - 2) : Ny - 2;
    HIP_TRY(h, dalloc(h, &h->dRowStart, (size_t)std::max(rows, 1)), "allocate row starts");
    HIP_TRY(h, dalloc(h, &h->dStepScal, kScalCount), "allocate step scalars");
    HIP_TRY(h, dalloc(h, &h->dCentre, (size_t)Nx + Ny + Nz), "allocate centre tables");
    FieldGeom &G = h->fg;
    G.B = F.box;
    G.dim = dim;
    G.Nx = Nx;
    G.Ny = Ny;
    G.Nz = Nz;
    G.nonUniform = h->nonUniform ? 1 : 0;
    G.dh = h->dh;
    G.wx = h->dFieldTab;
    G.wy = h->dFieldTab + Nx;
    G.wz = h->dFieldTab + Nx + Ny;
    G.types = F.types;
    G.c8 = F.c8;
    G.rowStart = h->dRowStart;
    G.rows = rows;
    // the row-walking reductions write one partial per block
    if ((int64_t)field_row_blocks(G) > h->partialsCap || 2 * (int64_t)field_box_blocks(G) > h->partialsCap ||
        2 * (int64_t)inflow_level_blocks(G) > h->partialsCap ||
        h->partialsCap < 1024)
        return fail(h, NPPE_ERR_INTERNAL, "partials buffer too small for the field reductions");
    return refresh_fields(h);
}

int nppe::refresh_fields(nppe_solver *h)
{
    // unknowns before each interior row (rows in IDP order: i, then j in 3D), counted and scanned on the device
    long long *total = h->dLayoutInfo;
    HIP_TRY(h, launch_number_rows(h->stream, h->fg, h->dRowStart, total), "row starts");
    long long count = 0;
    HIP_TRY(h, hipMemcpyAsync(&count, total, sizeof(count), hipMemcpyDeviceToHost, h->stream), "row starts");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "row starts");
    if (count != h->n) return fail(h, NPPE_ERR_INTERNAL, "row starts disagree with the unknown count");
    h->fg.n = h->n;
    return NPPE_OK;
}

extern "C" {

int nppe_rhs_device(nppe_solver *h, const double *d_u, const double *d_v, const double *d_w, const double *d_p,
                    const double *d_p_old, double rho, double dt, double *d_b)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, d_u, d_v, d_w, d_p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    if (h->n > 0 && !d_b) return fail(h, NPPE_ERR_ARGUMENT, "null b");
    HIP_TRY(h, rhs(h, d_u, d_v, d_w, d_p, d_p_old, rho, dt, d_b), "right-hand side");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "right-hand side");
    return NPPE_OK;
}

int nppe_rhs(nppe_solver *h, const double *u, const double *v, const double *w, const double *p, const double *p_old,
             double rho, double dt, double *b)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, u, v, w, p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    if (h->n > 0 && !b) return fail(h, NPPE_ERR_ARGUMENT, "null b");
    if (int rc = stage(h)) return rc;
    size_t sz[4];
    field_sizes(h, sz);
    HIP_TRY(h, put(h, h->dStage[0], u, sz[0]), "upload u");
    HIP_TRY(h, put(h, h->dStage[1], v, sz[1]), "upload v");
    if (h->dim == 3) HIP_TRY(h, put(h, h->dStage[2], w, sz[2]), "upload w");
    HIP_TRY(h, put(h, h->dStage[3], p, sz[3]), "upload p");
    if (p_old) HIP_TRY(h, put(h, h->dStage[4], p_old, sz[3]), "upload p_old");
    HIP_TRY(h, rhs(h, h->dStage[0], h->dStage[1], h->dStage[2], h->dStage[3], p_old ? h->dStage[4] : nullptr, rho, dt,
                   h->dCompB),
            "right-hand side");
    HIP_TRY(h, get(h, b, h->dCompB, (size_t)h->n), "download b");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "right-hand side");
    return NPPE_OK;
}

int nppe_project_device(nppe_solver *h, double *d_u, double *d_v, double *d_w, const double *d_p, double rho,
                        double dt)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, d_u, d_v, d_w, d_p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    HIP_TRY(h, launch_project(h->stream, h->fg, d_u, d_v, d_w, d_p, dt / rho), "projection");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "projection");
    return NPPE_OK;
}

int nppe_project(nppe_solver *h, double *u, double *v, double *w, const double *p, double rho, double dt)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, u, v, w, p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    if (int rc = stage(h)) return rc;
    size_t sz[4];
    field_sizes(h, sz);
    HIP_TRY(h, put(h, h->dStage[0], u, sz[0]), "upload u");
    HIP_TRY(h, put(h, h->dStage[1], v, sz[1]), "upload v");
    if (h->dim == 3) HIP_TRY(h, put(h, h->dStage[2], w, sz[2]), "upload w");
    HIP_TRY(h, put(h, h->dStage[3], p, sz[3]), "upload p");
    HIP_TRY(h, launch_project(h->stream, h->fg, h->dStage[0], h->dStage[1], h->dStage[2], h->dStage[3], dt / rho),
            "projection");
    HIP_TRY(h, get(h, u, h->dStage[0], sz[0]), "download u");
    HIP_TRY(h, get(h, v, h->dStage[1], sz[1]), "download v");
    if (h->dim == 3) HIP_TRY(h, get(h, w, h->dStage[2], sz[2]), "download w");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "projection");
    return NPPE_OK;
}

int nppe_enforce_pressure_bc_device(nppe_solver *h, double *d_p, const double *d_prescribed)
{
    if (int rc = ready(h)) return rc;
    if (!d_p) return fail(h, NPPE_ERR_ARGUMENT, "null p");
    HIP_TRY(h, launch_pressure_bc(h->stream, h->fg, d_p, d_prescribed), "pressure boundary conditions");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "pressure boundary conditions");
    return NPPE_OK;
}

int nppe_increment_reference_norm_device(nppe_solver *h, const double *d_b, const double *d_p_old, double *norm)
{
    if (int rc = ready(h)) return rc;
    if (!norm || (h->n > 0 && (!d_b || !d_p_old))) return fail(h, NPPE_ERR_ARGUMENT, "null b, p_old or norm");
    return reference_norm(h, d_b, d_p_old, norm);
}

int nppe_scatter_pressure_device(nppe_solver *h, const double *d_x, double *d_p, const double *d_p_old,
                                 const double *d_prescribed)
{
    if (int rc = ready(h)) return rc;
    if (!d_p || (h->n > 0 && !d_x)) return fail(h, NPPE_ERR_ARGUMENT, "null x or p");
    HIP_TRY(h, scatter(h, d_x, d_p, d_p_old, d_prescribed), "scatter pressure");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "scatter pressure");
    return NPPE_OK;
}

int nppe_accumulate_increment_device(nppe_solver *h, const double *d_p_old, double *d_p, const double *d_prescribed)
{
    if (int rc = ready(h)) return rc;
    if (!d_p || !d_p_old) return fail(h, NPPE_ERR_ARGUMENT, "null p or p_old");
    HIP_TRY(h, accumulate(h, d_p_old, d_p, d_prescribed), "accumulate increment");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "accumulate increment");
    return NPPE_OK;
}

int nppe_solve_laplacian_device(nppe_solver *h, const double *d_f, double *d_x, nppe_result *res)
{
    if (int rc = ready(h)) return rc;
    // SolveLaplacian returns false without solving when the system has no rows (PressureSolver3D.cpp:358-359)
    if (h->n == 0) return fail(h, NPPE_ERR_SIZE, "the committed system has no unknowns");
    if (!d_f || !d_x) return fail(h, NPPE_ERR_ARGUMENT, "null f or x");
    HIP_TRY(h, launch_scale_volume(h->stream, h->fg, d_f, h->dCompB), "volume scaling");
    return solve_device(h, h->dCompB, d_x, nullptr, 0.0, res);
}

int nppe_solve_laplacian(nppe_solver *h, const double *f, double *x, nppe_result *res)
{
    if (int rc = ready(h)) return rc;
    if (h->n == 0) return fail(h, NPPE_ERR_SIZE, "the committed system has no unknowns");
    if (!f || !x) return fail(h, NPPE_ERR_ARGUMENT, "null f or x");
    HIP_TRY(h, put(h, h->dCompX, f, (size_t)h->n), "upload f");
    HIP_TRY(h, launch_scale_volume(h->stream, h->fg, h->dCompX, h->dCompB), "volume scaling");
    const int rc = solve_device(h, h->dCompB, h->dCompX, nullptr, 0.0, res);
    if (rc < 0) return rc;
    HIP_TRY(h, get(h, x, h->dCompX, (size_t)h->n), "download x");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "download x");
    return rc;
}

int nppe_pressure_step_device(nppe_solver *h, const double *d_u, const double *d_v, const double *d_w, double *d_p,
                              const double *d_p_old, double rho, double dt, nppe_result *res)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, d_u, d_v, d_w, d_p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    HIP_TRY(h, rhs(h, d_u, d_v, d_w, d_p, d_p_old, rho, dt, h->dCompB), "right-hand side");
    double ref = 0.0;
    if (d_p_old)
        if (int rc = reference_norm(h, h->dCompB, d_p_old, &ref)) return rc;
    const int rc = solve_device(h, h->dCompB, h->dCompX, nullptr, ref, res);
    if (rc < 0) return rc;
    HIP_TRY(h, scatter(h, h->dCompX, d_p, d_p_old, nullptr), "scatter pressure");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "pressure step");
    return rc;
}

int nppe_warm_start_guess_device(nppe_solver *h, const double *d_p_prev, double *d_guess, int32_t *built)
{
    if (int rc = ready(h)) return rc;
    if (!built || !d_p_prev || (h->n > 0 && !d_guess)) return fail(h, NPPE_ERR_ARGUMENT, "null p_prev, guess or flag");
    bool b = false;
    if (int rc = warm_guess(h, d_p_prev, d_guess, &b)) return rc;
    HIP_TRY(h, hipStreamSynchronize(h->stream), "warm start");
    *built = b ? 1 : 0;
    return NPPE_OK;
}

int nppe_scatter_pressure_level_device(nppe_solver *h, const double *d_x, double *d_p, const double *d_p_old,
                                       const double *d_prescribed, const nppe_inflow_level *level)
{
    if (int rc = ready(h)) return rc;
    if (!d_p || (h->n > 0 && !d_x)) return fail(h, NPPE_ERR_ARGUMENT, "null x or p");
    if (int rc = level_args(h, level)) return rc;
    HIP_TRY(h, scatter(h, d_x, d_p, d_p_old, d_prescribed, level), "scatter pressure");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "scatter pressure");
    return NPPE_OK;
}

int nppe_accumulate_increment_level_device(nppe_solver *h, const double *d_p_old, double *d_p,
                                           const double *d_prescribed, const nppe_inflow_level *level)
{
    if (int rc = ready(h)) return rc;
    if (!d_p || !d_p_old) return fail(h, NPPE_ERR_ARGUMENT, "null p or p_old");
    if (int rc = level_args(h, level)) return rc;
    HIP_TRY(h, accumulate(h, d_p_old, d_p, d_prescribed, level), "accumulate increment");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "accumulate increment");
    return NPPE_OK;
}

int nppe_pressure_step_ex_device(nppe_solver *h, const double *d_u, const double *d_v, const double *d_w, double *d_p,
                                 const double *d_p_old, double rho, double dt, const nppe_step_options *o,
                                 nppe_result *res)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, d_u, d_v, d_w, d_p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    const nppe_inflow_level *lv = o ? o->inflow_level : nullptr;
    if (int rc = level_args(h, lv)) return rc;
    if (o && o->warm_started) *o->warm_started = 0;
    HIP_TRY(h, rhs(h, d_u, d_v, d_w, d_p, d_p_old, rho, dt, h->dCompB), "right-hand side");
    double ref = 0.0;
    if (d_p_old)
        if (int rc = reference_norm(h, h->dCompB, d_p_old, &ref)) return rc;
    // SolveScheme (PressureSolver3D.cpp:309-314): only a Chorin step warm starts
    bool warm = false;
    if (o && o->warm_start && o->d_p_prev && !d_p_old)
        if (int rc = warm_guess(h, o->d_p_prev, h->dCompG, &warm)) return rc;
    const int rc = solve_device(h, h->dCompB, h->dCompX, warm ? h->dCompG : nullptr, ref, res);
    if (rc < 0) return rc;
    HIP_TRY(h, scatter(h, h->dCompX, d_p, d_p_old, o ? o->d_prescribed : nullptr, lv), "scatter pressure");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "pressure step");
    if (o && o->warm_started) *o->warm_started = warm ? 1 : 0;
    return rc;
}

int nppe_projection_step_ex_device(nppe_solver *h, double *d_u, double *d_v, double *d_w, double *d_p,
                                   const double *d_p_old, double rho, double dt, const double *d_prescribed,
                                   const nppe_inflow_level *level)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, d_u, d_v, d_w, d_p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    if (int rc = level_args(h, level)) return rc;
    HIP_TRY(h, launch_project(h->stream, h->fg, d_u, d_v, d_w, d_p, dt / rho), "projection");
    if (d_p_old) HIP_TRY(h, accumulate(h, d_p_old, d_p, d_prescribed, level), "accumulate increment");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "projection step");
    return NPPE_OK;
}

int nppe_projection_step_device(nppe_solver *h, double *d_u, double *d_v, double *d_w, double *d_p,
                                const double *d_p_old, double rho, double dt, const double *d_prescribed)
{
    if (int rc = ready(h)) return rc;
    if (int rc = fields_args(h, d_u, d_v, d_w, d_p)) return rc;
    if (int rc = scalars_args(h, rho, dt)) return rc;
    HIP_TRY(h, launch_project(h->stream, h->fg, d_u, d_v, d_w, d_p, dt / rho), "projection");
    if (d_p_old) HIP_TRY(h, accumulate(h, d_p_old, d_p, d_prescribed), "accumulate increment");
    HIP_TRY(h, hipStreamSynchronize(h->stream), "projection step");
    return NPPE_OK;
}

}  // extern "C"
