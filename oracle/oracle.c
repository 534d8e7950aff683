/* CPU oracle: plain-C restatement of the reference's pressure path. TEST INFRASTRUCTURE ONLY -- only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this, as the checker; the product library never
 * does. Every function names the reference lines it follows; see oracle.h for how each one is pinned. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- spacing tables: MAC.cpp:19-175 ---------------------------------------------------------------------------- */

/* MAC::FactorAt (MAC.cpp:19-41): the largest local refinement factor over the intervals containing x. */
static double factor_at(const oracle_refinement *refs, int nrefs, double x)
{
    double f = 1.0;
    for (int q = 0; q < nrefs; q++)
    {
        const oracle_refinement *r = &refs[q];
        if (r->factor <= 1) continue;
        if (x < r->start || x > r->end) continue;
        double local = (double)r->factor;
        const double span = r->end - r->start;
        if (span > 0.0 && r->mode != ORACLE_REFINE_HARD)
        {
            const double t = (x - r->start) / span;
            const double sN = (r->mode == ORACLE_REFINE_RAMP)        ? t
                              : (r->mode == ORACLE_REFINE_RAMP_BACK) ? (1.0 - t)
                                                                     : (1.0 - fabs(2.0 * t - 1.0));
            local = 1.0 + (r->factor - 1.0) * sN;
        }
        if (local > f) f = local;
    }
    return f;
}

typedef struct
{
    double *v;
    int n, cap, overflow;
} dvec;

static void dpush(dvec *d, double x)
{
    if (d->n >= d->cap) { d->overflow = 1; return; }
    d->v[d->n++] = x;
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* MAC::BuildAxis (MAC.cpp:43-146) followed by MAC::AddRing (:150-161); only the widths are produced (face and
 * centre positions are not read by the pressure path). */
int oracle_build_axis(double a0, double af, int nBase, double dh, const oracle_refinement *refs, int nrefs,
                      int nonUniform, double *widths, int cap)
{
    (void)af;
    int anyActive = 0, anyGraded = 0;
    for (int q = 0; q < nrefs; q++)
        if (refs[q].factor > 1)
        {
            anyActive = 1;
            if (refs[q].mode != ORACLE_REFINE_HARD) anyGraded = 1;
        }
    if (cap < 3) return -1;
    dvec size = {widths + 1, 0, cap - 2, 0};

    if (!nonUniform || !anyActive || nBase <= 0)
    {
        for (int j = 0; j < nBase; j++) dpush(&size, dh);
    }
    else
    {
        const double aEnd = a0 + nBase * dh;
        if (!anyGraded)
        {
            for (int j = 0; j < nBase; j++)
            {
                const double c = a0 + j * dh + dh / 2.0;
                int f = (int)(factor_at(refs, nrefs, c) + 0.5);
                if (f < 1) f = 1;
                for (int t = 0; t < f; t++) dpush(&size, dh / (double)f);
            }
        }
        else
        {
            double *anchors = (double *)malloc(sizeof(double) * (size_t)(2 + 2 * nrefs));
            int na = 0;
            anchors[na++] = a0;
            anchors[na++] = aEnd;
            for (int q = 0; q < nrefs; q++)
                if (refs[q].factor > 1 && refs[q].mode == ORACLE_REFINE_HARD)
                {
                    const double e2[2] = {refs[q].start, refs[q].end};
                    for (int s = 0; s < 2; s++)
                        if (e2[s] > a0 + 1e-12 && e2[s] < aEnd - 1e-12) anchors[na++] = e2[s];
                }
            qsort(anchors, (size_t)na, sizeof(double), cmp_double);
            int ne = 0;
            for (int s = 0; s < na; s++)
                if (ne == 0 || anchors[s] - anchors[ne - 1] > 1e-12) anchors[ne++] = anchors[s];
            const int guard = 4000000;
            for (int s = 0; s + 1 < ne; s++)
            {
                const double p = anchors[s], q = anchors[s + 1], L = q - p;
                int graded = 0;
                for (int m = 0; m < nrefs; m++)
                    if (refs[m].factor > 1 && refs[m].mode != ORACLE_REFINE_HARD && refs[m].start < q - 1e-12 &&
                        refs[m].end > p + 1e-12)
                        graded = 1;
                const double fm = factor_at(refs, nrefs, 0.5 * (p + q));
                if (!graded)
                {
                    int n = (int)lround(L * fm / dh);
                    if (n < 1) n = 1;
                    for (int t = 0; t < n; t++) dpush(&size, L / n);
                    continue;
                }
                const int first = size.n;
                double x = p;
                while (x < q - 1e-12 && size.n < guard && !size.overflow)
                {
                    double fa = factor_at(refs, nrefs, x);
                    if (fa < 1.0) fa = 1.0;
                    const double h = dh / fa;
                    if (h <= 0.0 || !isfinite(h)) break;
                    dpush(&size, h);
                    x += h;
                }
                if (size.n == first) dpush(&size, L);
                double sum = 0.0;
                for (int t = first; t < size.n; t++) sum += size.v[t];
                for (int t = first; t < size.n; t++) size.v[t] *= L / sum;
            }
            free(anchors);
        }
    }
    if (size.overflow || size.n == 0) return size.n == 0 && !size.overflow ? 0 : -1;
    /* AddRing: one cell outside each end mirroring the adjacent width. */
    widths[0] = size.v[0];
    widths[size.n + 1] = size.v[size.n - 1];
    return size.n + 2;
}

/* MAC::GapTable (MAC.cpp:165-175). */
void oracle_gap_table(const double *w, int n, double *gap)
{
    for (int m = -1; m <= n - 1; m++)
    {
        const double a = w[m < 0 ? 0 : (m >= n ? n - 1 : m)];
        const double b = w[m + 1 >= n ? n - 1 : m + 1];
        gap[m + 1] = 0.5 * (a + b);
    }
}

/* ---- assembly: PressureSolver3D.cpp:32-87,128-178; PressureSolver2D.cpp:31-60,100-152 ------------------------- */

static int is_dirichlet(int t) { return t == ORACLE_EMPTY || t == ORACLE_PRESSURE; }

static double face_gap(const double *w, int N, int m)
{
    /* FaceGap(a, m) = gap[m + 1] of GapTable over the N widths. */
    const double a = w[m < 0 ? 0 : (m >= N ? N - 1 : m)];
    const double b = w[m + 1 >= N ? N - 1 : m + 1];
    return 0.5 * (a + b);
}

/* DirichletRingFactor (PressureSettings.h:307-310). */
static double ring_factor_axis(const double *w, int N, int n, int s)
{
    return face_gap(w, N, s < 0 ? n - 1 : n) / (0.5 * w[n]);
}

int64_t oracle_count_unknowns(const oracle_grid *g)
{
    int64_t n = 0;
    const int Nz = g->dim == 2 ? 1 : g->Nz;
    const int k0 = g->dim == 2 ? 0 : 1, k1 = g->dim == 2 ? 1 : Nz - 1;
    for (int i = 1; i < g->Ny - 1; i++)
        for (int j = 1; j < g->Nx - 1; j++)
            for (int k = k0; k < k1; k++)
                if (g->types[((int64_t)i * g->Nx + j) * Nz + k] == ORACLE_FLUID) n++;
    return n;
}

typedef struct
{
    int32_t col;
    double val;
} entry;

static void sort_row(entry *e, int m)
{
    for (int a = 1; a < m; a++)
    {
        entry t = e[a];
        int b = a - 1;
        while (b >= 0 && e[b].col > t.col) { e[b + 1] = e[b]; b--; }
        e[b + 1] = t;
    }
}

int64_t oracle_assemble(const oracle_grid *g, int32_t *order, int32_t *rowptr, int32_t *col, double *val,
                        int32_t *singular)
{
    const int dim = g->dim, Nx = g->Nx, Ny = g->Ny, Nz = dim == 2 ? 1 : g->Nz;
    const int64_t cells = (int64_t)Nx * Ny * Nz;
    int32_t *idp = (int32_t *)malloc(sizeof(int32_t) * (size_t)cells);
    for (int64_t c = 0; c < cells; c++) idp[c] = -1;
    const int k0 = dim == 2 ? 0 : 1, k1 = dim == 2 ? 1 : Nz - 1;
    int32_t n = 0;
    /* IDP numbering, (i, j, k) lexicographic (PressureSolver3D.cpp:133-139; 2D :107-112). */
    for (int i = 1; i < Ny - 1; i++)
        for (int j = 1; j < Nx - 1; j++)
            for (int k = k0; k < k1; k++)
            {
                const int64_t c = ((int64_t)i * Nx + j) * Nz + k;
                if (g->types[c] == ORACLE_FLUID) { order[n] = (int32_t)c; idp[c] = n++; }
            }
    const double scale = g->nonUniform ? 1.0 : 1.0 / (g->dh * g->dh);
    const int nq = dim == 2 ? 4 : 6;
    int dirichlet = 0;
    int64_t nnz = 0;
    rowptr[0] = 0;
    for (int32_t row = 0; row < n; row++)
    {
        const int64_t c = order[row];
        const int k = (int)(c % Nz), j = (int)((c / Nz) % Nx), i = (int)(c / ((int64_t)Nz * Nx));
        int ni[6], nj[6], nk[6];
        double cq[6];
        if (dim == 3)
        {
            /* neighbour order (i+1, i-1, j+1, j-1, k+1, k-1): PressureSolver3D.cpp:66-68 */
            const int DI[6] = {1, -1, 0, 0, 0, 0}, DJ[6] = {0, 0, 1, -1, 0, 0}, DK[6] = {0, 0, 0, 0, 1, -1};
            for (int q = 0; q < 6; q++) { ni[q] = i + DI[q]; nj[q] = j + DJ[q]; nk[q] = k + DK[q]; }
            if (!g->nonUniform)
                for (int q = 0; q < 6; q++) cq[q] = 1.0;
            else
            {
                /* FaceCoefs3D (:47-63) with MAC::FaceArea (MAC.h:91-93). */
                const double ax = g->wy[i] * g->wz[k];
                const double ay = g->wx[j] * g->wz[k];
                const double az = g->wx[j] * g->wy[i];
                cq[0] = ay / face_gap(g->wy, Ny, i);
                cq[1] = ay / face_gap(g->wy, Ny, i - 1);
                cq[2] = ax / face_gap(g->wx, Nx, j);
                cq[3] = ax / face_gap(g->wx, Nx, j - 1);
                cq[4] = az / face_gap(g->wz, Nz, k);
                cq[5] = az / face_gap(g->wz, Nz, k - 1);
            }
        }
        else
        {
            /* N, S, E, W (PressureSolver2D.cpp:122-128) */
            const int I2[4] = {i + 1, i - 1, i, i}, J2[4] = {j, j, j + 1, j - 1};
            for (int q = 0; q < 4; q++) { ni[q] = I2[q]; nj[q] = J2[q]; nk[q] = 0; }
            cq[0] = g->wx[j] / face_gap(g->wy, Ny, i);
            cq[1] = g->wx[j] / face_gap(g->wy, Ny, i - 1);
            cq[2] = g->wy[i] / face_gap(g->wx, Nx, j);
            cq[3] = g->wy[i] / face_gap(g->wx, Nx, j - 1);
        }
        entry e[7];
        int m = 0;
        double diag = 0.0;
        for (int q = 0; q < nq; q++)
        {
            const int64_t nc = ((int64_t)ni[q] * Nx + nj[q]) * Nz + nk[q];
            if (g->types[nc] == ORACLE_FLUID)
            {
                diag -= cq[q];
                if (idp[nc] >= 0) { e[m].col = idp[nc]; e[m].val = scale * cq[q]; m++; }
                else dirichlet = 1;
            }
        }
        for (int q = 0; q < nq; q++)
        {
            const int64_t nc = ((int64_t)ni[q] * Nx + nj[q]) * Nz + nk[q];
            if (!is_dirichlet(g->types[nc])) continue;
            /* RingFactor3D (:80-87) / RingFactor2D (:54-60) */
            double rf = 1.0;
            const int onRing = ni[q] == 0 || ni[q] == Ny - 1 || nj[q] == 0 || nj[q] == Nx - 1 ||
                               (dim == 3 && (nk[q] == 0 || nk[q] == Nz - 1));
            if (onRing)
            {
                const int s = (q % 2 == 0) ? 1 : -1;
                rf = q < 2 ? ring_factor_axis(g->wy, Ny, i, s)
                           : (q < 4 ? ring_factor_axis(g->wx, Nx, j, s) : ring_factor_axis(g->wz, Nz, k, s));
            }
            diag -= cq[q] * rf;
            dirichlet = 1;
        }
        if (diag != 0.0) { e[m].col = row; e[m].val = scale * diag; m++; }
        /* setFromTriplets + makeCompressed: columns ascending within the row (PressureSettings.h:557-559). */
        sort_row(e, m);
        for (int a = 0; a < m; a++) { col[nnz] = e[a].col; val[nnz] = e[a].val; nnz++; }
        rowptr[row + 1] = (int32_t)nnz;
    }
    *singular = dirichlet ? 0 : 1;
    free(idp);
    return nnz;
}

/* ---- Krylov: Eigen 3.3.7 ConjugateGradient.h:28-91, BasicPreconditioners.h:36-91 ------------------------------- */

void oracle_spmv(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *x, double *y)
{
    for (int32_t r = 0; r < n; r++)
    {
        double t = 0.0;
        for (int32_t q = rowptr[r]; q < rowptr[r + 1]; q++) t += val[q] * x[col[q]];
        y[r] = t;
    }
}

static double dot(int32_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int32_t q = 0; q < n; q++) s += a[q] * b[q];
    return s;
}

int oracle_cg(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b, double *x,
              double tol, int maxIters, int precond, double *tol_error)
{
    double *r = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    double *p = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    double *z = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    double *t = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    double *inv = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    /* DiagonalPreconditioner::factorize: 1/diag, or 1 where the row has no (or a zero) diagonal. */
    for (int32_t rr = 0; rr < n; rr++)
    {
        inv[rr] = 1.0;
        if (precond != 0) continue;
        for (int32_t q = rowptr[rr]; q < rowptr[rr + 1]; q++)
            if (col[q] == rr && val[q] != 0.0) inv[rr] = 1.0 / val[q];
    }
    int iters = 0;
    oracle_spmv(n, rowptr, col, val, x, t);
    for (int32_t q = 0; q < n; q++) r[q] = b[q] - t[q];
    const double rhsNorm2 = dot(n, b, b);
    if (rhsNorm2 == 0.0)
    {
        for (int32_t q = 0; q < n; q++) x[q] = 0.0;
        *tol_error = 0.0;
        goto done;
    }
    {
        const double considerAsZero = 2.2250738585072014e-308; /* numeric_limits<double>::min() */
        double threshold = tol * tol * rhsNorm2;
        if (threshold < considerAsZero) threshold = considerAsZero;
        double rn2 = dot(n, r, r);
        if (rn2 < threshold)
        {
            *tol_error = sqrt(rn2 / rhsNorm2);
            goto done;
        }
        for (int32_t q = 0; q < n; q++) p[q] = inv[q] * r[q];
        double absNew = dot(n, r, p);
        int i = 0;
        while (i < maxIters)
        {
            oracle_spmv(n, rowptr, col, val, p, t);
            const double alpha = absNew / dot(n, p, t);
            for (int32_t q = 0; q < n; q++) x[q] += alpha * p[q];
            for (int32_t q = 0; q < n; q++) r[q] -= alpha * t[q];
            rn2 = dot(n, r, r);
            if (rn2 < threshold) break;
            for (int32_t q = 0; q < n; q++) z[q] = inv[q] * r[q];
            const double absOld = absNew;
            absNew = dot(n, r, z);
            const double beta = absNew / absOld;
            for (int32_t q = 0; q < n; q++) p[q] = z[q] + beta * p[q];
            i++;
        }
        *tol_error = sqrt(rn2 / rhsNorm2);
        iters = i;
    }
done:
    free(r); free(p); free(z); free(t); free(inv);
    return iters;
}

/* SolveEigenPressure's floor handling (PressureSettings.h:415-427) for eigen_solver = CG, DIAGONAL. */
int oracle_solve_pressure(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b,
                          double *x, const double *guess, double tol, int maxIters, double referenceNorm)
{
    double s = tol;
    if (referenceNorm > 0.0)
    {
        const double bn = sqrt(dot(n, b, b));
        if (bn <= tol * referenceNorm)
        {
            for (int32_t q = 0; q < n; q++) x[q] = guess ? guess[q] : 0.0;
            return 0;
        }
        s = tol * referenceNorm / bn;
        if (s > 0.5) s = 0.5;
    }
    for (int32_t q = 0; q < n; q++) x[q] = guess ? guess[q] : 0.0;
    double err = 0.0;
    return oracle_cg(n, rowptr, col, val, b, x, s, maxIters, 0, &err);
}

/* SubtractMean (PressureSettings.h:314-321). */
void oracle_subtract_mean(double *v, int64_t n)
{
    if (v == NULL || n <= 0) return;
    double m = 0.0;
    for (int64_t i = 0; i < n; i++) m += v[i];
    m /= (double)n;
    for (int64_t i = 0; i < n; i++) v[i] -= m;
}

double oracle_relres(int32_t n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b,
                     const double *x)
{
    double rr = 0.0, bb = 0.0;
    for (int32_t r = 0; r < n; r++)
    {
        double t = 0.0;
        for (int32_t q = rowptr[r]; q < rowptr[r + 1]; q++) t += val[q] * x[col[q]];
        const double d = b[r] - t;
        rr += d * d;
        bb += b[r] * b[r];
    }
    return bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
}

/* ---- right-hand side and projection: PressureSolver3D.cpp:248-260, :288-298, :399-438;
 *      PressureSolver2D.cpp:218-235, :262-271, :370-396; MAC.cpp:349-367, :443-455; MAC.h:149-155, :363-367 ---- */

static int64_t at3(const oracle_grid *g, int i, int j, int k) { return ((int64_t)i * g->Nx + j) * g->Nz + k; }

/* The known pressure at a Dirichlet neighbour: p_b, or p_b(t^{n+1}) - p_b(t^n) when pOld is given. */
static double known_p(const double *p, const double *pOld, int64_t c) { return pOld ? p[c] - pOld[c] : p[c]; }

void oracle_rhs(const oracle_grid *g, const int32_t *order, int64_t n, const double *u, const double *v,
                const double *w, const double *p, const double *pOld, double rho, double dt, double *b)
{
    const int Nx = g->Nx, Ny = g->Ny, Nz = g->dim == 2 ? 1 : g->Nz;
    const double scale = g->nonUniform ? 1.0 : 1.0 / (g->dh * g->dh);
    const double rhoDt = rho / dt;
    for (int64_t q = 0; q < n; q++)
    {
        const int64_t c = order[q];
        const int k = (int)(c % Nz), j = (int)((c / Nz) % Nx), i = (int)(c / ((int64_t)Nz * Nx));
        double div, vol, ds = 0.0;
        if (g->dim == 3)
        {
            const double u0 = u[((int64_t)i * (Nx + 1) + j) * Nz + k], u1 = u[((int64_t)i * (Nx + 1) + j + 1) * Nz + k];
            const double v0 = v[((int64_t)i * Nx + j) * Nz + k], v1 = v[((int64_t)(i + 1) * Nx + j) * Nz + k];
            const double w0 = w[((int64_t)i * Nx + j) * (Nz + 1) + k], w1 = w[((int64_t)i * Nx + j) * (Nz + 1) + k + 1];
            if (g->nonUniform)
            {
                /* MAC::GetDivergencyAt, stretched: net face flux over the cell volume (MAC.cpp:353-360) */
                const double dx = g->wx[j], dy = g->wy[i], dz = g->wz[k];
                const double fu = (u1 - u0) * dy * dz, fv = (v1 - v0) * dx * dz, fw = (w1 - w0) * dx * dy;
                div = (fu + fv + fw) / (dx * dy * dz);
                vol = g->wx[j] * g->wy[i] * g->wz[k];   /* MAC cellVolume (MAC.cpp:243) */
            }
            else
            {
                const double du = (u1 - u0) / g->dh, dv = (v1 - v0) / g->dh, dw = (w1 - w0) / g->dh;
                div = du + dv + dw;
                vol = 1.0;
            }
            /* DirichletSum (PressureSolver3D.cpp:248-260): PRESSURE neighbours only, FaceCoefs3D x RingFactor3D */
            double cq[6];
            if (!g->nonUniform)
                for (int a = 0; a < 6; a++) cq[a] = 1.0;
            else
            {
                const double ax = g->wy[i] * g->wz[k], ay = g->wx[j] * g->wz[k], az = g->wx[j] * g->wy[i];
                cq[0] = ay / face_gap(g->wy, Ny, i);
                cq[1] = ay / face_gap(g->wy, Ny, i - 1);
                cq[2] = ax / face_gap(g->wx, Nx, j);
                cq[3] = ax / face_gap(g->wx, Nx, j - 1);
                cq[4] = az / face_gap(g->wz, Nz, k);
                cq[5] = az / face_gap(g->wz, Nz, k - 1);
            }
            const int DI[6] = {1, -1, 0, 0, 0, 0}, DJ[6] = {0, 0, 1, -1, 0, 0}, DK[6] = {0, 0, 0, 0, 1, -1};
            for (int a = 0; a < 6; a++)
            {
                const int ni = i + DI[a], nj = j + DJ[a], nk = k + DK[a];
                const int64_t nc = at3(g, ni, nj, nk);
                if (g->types[nc] != ORACLE_PRESSURE) continue;
                double rf = 1.0;
                if (ni == 0 || ni == Ny - 1 || nj == 0 || nj == Nx - 1 || nk == 0 || nk == Nz - 1)
                {
                    const int s = (a % 2 == 0) ? 1 : -1;
                    rf = a < 2 ? ring_factor_axis(g->wy, Ny, i, s)
                               : (a < 4 ? ring_factor_axis(g->wx, Nx, j, s) : ring_factor_axis(g->wz, Nz, k, s));
                }
                ds += cq[a] * rf * known_p(p, pOld, nc);
            }
        }
        else
        {
            const double u0 = u[(int64_t)i * (Nx + 1) + j], u1 = u[(int64_t)i * (Nx + 1) + j + 1];
            const double v0 = v[(int64_t)i * Nx + j], v1 = v[(int64_t)(i + 1) * Nx + j];
            if (g->nonUniform)
            {
                const double dx = g->wx[j], dy = g->wy[i];
                const double fu = (u1 - u0) * dy, fv = (v1 - v0) * dx;
                div = (fu + fv) / (dx * dy);
                vol = g->wx[j] * g->wy[i];   /* MAC.cpp:235 */
            }
            else
            {
                const double du = (u1 - u0) / g->dh, dv = (v1 - v0) / g->dh;
                div = du + dv;
                vol = 1.0;
            }
            /* DirichletSum 2D (PressureSolver2D.cpp:223-235): N, S, E, W */
            const int NI[4] = {i + 1, i - 1, i, i}, NJ[4] = {j, j, j + 1, j - 1};
            for (int a = 0; a < 4; a++)
            {
                const int64_t nc = (int64_t)NI[a] * Nx + NJ[a];
                if (g->types[nc] != ORACLE_PRESSURE) continue;
                const double cf = a < 2 ? g->wx[j] / face_gap(g->wy, Ny, a == 0 ? i : i - 1)
                                        : g->wy[i] / face_gap(g->wx, Nx, a == 2 ? j : j - 1);
                double rf = 1.0;
                if (NI[a] == 0 || NI[a] == Ny - 1 || NJ[a] == 0 || NJ[a] == Nx - 1)
                    rf = a < 2 ? ring_factor_axis(g->wy, Ny, i, a == 0 ? 1 : -1)
                               : ring_factor_axis(g->wx, Nx, j, a == 2 ? 1 : -1);
                ds += cf * rf * known_p(p, pOld, nc);
            }
        }
        b[q] = rhoDt * div * vol - scale * ds;
    }
}

/* ProjectionGradient3D / 2D: the face gradient over FaceGap, times DirichletRingFactor where the face's outer cell
 * is a Dirichlet ring cell. n = the face's index along the axis, N = the axis length, tRing0 / tRingN the types of
 * the ring cells at 0 and N-1 on the face's line. */
static double proj_grad(double pc, double pm, const double *wa, int N, int n, int tRing0, int tRingN)
{
    const double grad = (pc - pm) / face_gap(wa, N, n - 1);
    if (n == 1 && is_dirichlet(tRing0)) return grad * ring_factor_axis(wa, N, 1, -1);
    if (n == N - 1 && is_dirichlet(tRingN)) return grad * ring_factor_axis(wa, N, N - 2, +1);
    return grad;
}

static int wall(int a, int b) { return a == ORACLE_SOLID || b == ORACLE_SOLID; }

void oracle_project(const oracle_grid *g, double *u, double *v, double *w, const double *p, double rho, double dt)
{
    const int Nx = g->Nx, Ny = g->Ny, Nz = g->dim == 2 ? 1 : g->Nz;
    const double f = dt / rho;
    const uint8_t *t = g->types;
    if (g->dim == 3)
    {
        for (int i = 1; i < Ny - 1; i++)
            for (int j = 1; j < Nx; j++)
                for (int k = 1; k < Nz - 1; k++)
                    if (!wall(t[at3(g, i, j, k)], t[at3(g, i, j - 1, k)]))
                    {
                        const double gr = proj_grad(p[at3(g, i, j, k)], p[at3(g, i, j - 1, k)], g->wx, Nx, j,
                                                    t[at3(g, i, 0, k)], t[at3(g, i, Nx - 1, k)]);
                        double *uf = &u[((int64_t)i * (Nx + 1) + j) * Nz + k];
                        *uf = *uf - f * gr;
                    }
        for (int i = 1; i < Ny; i++)
            for (int j = 1; j < Nx - 1; j++)
                for (int k = 1; k < Nz - 1; k++)
                    if (!wall(t[at3(g, i, j, k)], t[at3(g, i - 1, j, k)]))
                    {
                        const double gr = proj_grad(p[at3(g, i, j, k)], p[at3(g, i - 1, j, k)], g->wy, Ny, i,
                                                    t[at3(g, 0, j, k)], t[at3(g, Ny - 1, j, k)]);
                        double *vf = &v[((int64_t)i * Nx + j) * Nz + k];
                        *vf = *vf - f * gr;
                    }
        for (int i = 1; i < Ny - 1; i++)
            for (int j = 1; j < Nx - 1; j++)
                for (int k = 1; k < Nz; k++)
                    if (!wall(t[at3(g, i, j, k)], t[at3(g, i, j, k - 1)]))
                    {
                        const double gr = proj_grad(p[at3(g, i, j, k)], p[at3(g, i, j, k - 1)], g->wz, Nz, k,
                                                    t[at3(g, i, j, 0)], t[at3(g, i, j, Nz - 1)]);
                        double *wf = &w[((int64_t)i * Nx + j) * (Nz + 1) + k];
                        *wf = *wf - f * gr;
                    }
        return;
    }
    for (int i = 1; i < Ny - 1; i++)
        for (int j = 1; j < Nx; j++)
            if (!wall(t[(int64_t)i * Nx + j], t[(int64_t)i * Nx + j - 1]))
            {
                const double gr = proj_grad(p[(int64_t)i * Nx + j], p[(int64_t)i * Nx + j - 1], g->wx, Nx, j,
                                            t[(int64_t)i * Nx], t[(int64_t)i * Nx + Nx - 1]);
                u[(int64_t)i * (Nx + 1) + j] = u[(int64_t)i * (Nx + 1) + j] - f * gr;
            }
    for (int i = 1; i < Ny; i++)
        for (int j = 1; j < Nx - 1; j++)
            if (!wall(t[(int64_t)i * Nx + j], t[(int64_t)(i - 1) * Nx + j]))
            {
                const double gr = proj_grad(p[(int64_t)i * Nx + j], p[(int64_t)(i - 1) * Nx + j], g->wy, Ny, i, t[j],
                                            t[(int64_t)(Ny - 1) * Nx + j]);
                v[(int64_t)i * Nx + j] = v[(int64_t)i * Nx + j] - f * gr;
            }
}

/* ---- pressure boundary conditions, scatter and increment: MAC.cpp:866-871, :1064-1096; MAC.h:482-486;
 *      PressureSolver3D.cpp:181-186, :316-336, :374-393; PressureSolver2D.cpp:156-160, :287-304, :340-357 ------- */

static void ring_entry(const oracle_grid *g, double *p, const double *prescribed, int64_t dst, int64_t src)
{
    const int t = g->types[dst];
    if (t == ORACLE_EMPTY) p[dst] = 0.0;
    else if (t == ORACLE_PRESSURE) { if (prescribed) p[dst] = prescribed[dst]; }
    else p[dst] = p[src];
}

void oracle_enforce_pressure_bc(const oracle_grid *g, double *p, const double *prescribed)
{
    const int Nx = g->Nx, Ny = g->Ny, Nz = g->dim == 2 ? 1 : g->Nz;
    if (g->dim == 2)
    {
        for (int j = 0; j < Nx; j++)
        {
            ring_entry(g, p, prescribed, j, (int64_t)Nx + j);
            ring_entry(g, p, prescribed, (int64_t)(Ny - 1) * Nx + j, (int64_t)(Ny - 2) * Nx + j);
        }
        for (int i = 0; i < Ny; i++)
        {
            ring_entry(g, p, prescribed, (int64_t)i * Nx, (int64_t)i * Nx + 1);
            ring_entry(g, p, prescribed, (int64_t)i * Nx + Nx - 1, (int64_t)i * Nx + Nx - 2);
        }
        return;
    }
    for (int j = 0; j < Nx; j++)
        for (int k = 0; k < Nz; k++)
        {
            ring_entry(g, p, prescribed, at3(g, 0, j, k), at3(g, 1, j, k));
            ring_entry(g, p, prescribed, at3(g, Ny - 1, j, k), at3(g, Ny - 2, j, k));
        }
    for (int i = 0; i < Ny; i++)
        for (int k = 0; k < Nz; k++)
        {
            ring_entry(g, p, prescribed, at3(g, i, 0, k), at3(g, i, 1, k));
            ring_entry(g, p, prescribed, at3(g, i, Nx - 1, k), at3(g, i, Nx - 2, k));
        }
    for (int i = 0; i < Ny; i++)
        for (int j = 0; j < Nx; j++)
        {
            ring_entry(g, p, prescribed, at3(g, i, j, 0), at3(g, i, j, 1));
            ring_entry(g, p, prescribed, at3(g, i, j, Nz - 1), at3(g, i, j, Nz - 2));
        }
}

/* LevelPressure without a scene (PressureSolver3D.cpp:181-186): v -= mean(v). */
static void remove_mean(double *v, int64_t n)
{
    if (n <= 0) return;
    double s = 0.0;
    for (int64_t q = 0; q < n; q++) s += v[q];
    const double mean = s / (double)n;
    for (int64_t q = 0; q < n; q++) v[q] -= mean;
}

void oracle_scatter_pressure(const oracle_grid *g, const int32_t *order, int64_t n, const double *x, double *p,
                             const double *pOld, int level, const double *prescribed)
{
    double *xl = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t q = 0; q < n; q++) xl[q] = x[q];
    if (!pOld && level) remove_mean(xl, n);
    for (int64_t q = 0; q < n; q++) p[order[q]] = xl[q];
    free(xl);
    oracle_enforce_pressure_bc(g, p, prescribed);
    if (pOld)
    {
        const int64_t cells = (int64_t)g->Nx * g->Ny * (g->dim == 2 ? 1 : g->Nz);
        for (int64_t c = 0; c < cells; c++)
            if (g->types[c] == ORACLE_PRESSURE) p[c] = p[c] - pOld[c];
    }
}

void oracle_accumulate_increment(const oracle_grid *g, const int32_t *order, int64_t n, const double *pOld,
                                 double *p, int level, const double *prescribed)
{
    double *v = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t q = 0; q < n; q++) v[q] = pOld[order[q]] + p[order[q]];
    if (level) remove_mean(v, n);
    for (int64_t q = 0; q < n; q++) p[order[q]] = v[q];
    free(v);
    oracle_enforce_pressure_bc(g, p, prescribed);
}

/* LevelPressure's inflow part (PressureSolver3D.cpp:15-30, :181-211; PressureSolver2D.cpp:13-29, :156-178), on x
 * (IDP order) AFTER its mean has been removed: for each ring cell of the x, y (and z) faces in the reference's loop
 * order, an INFLOW ring cell whose two inward neighbours are FLUID unknowns and where the scene gives a level
 * (level[c] not NaN) adds level - (p1 + (p1 - p2) (sR - s1) / (s1 - s2)); x += sum / count when any did. centre:
 * MAC::CentrePos(a, n), n = 0 .. N-1, per axis. Returns the count. */
static int64_t unknown_id(const int32_t *inv, int64_t c) { return inv[c]; }

int oracle_level_inflow(const oracle_grid *g, const int32_t *order, int64_t n, double *x, const double *level,
                        const double *cx, const double *cy, const double *cz)
{
    const int Nx = g->Nx, Ny = g->Ny, Nz = g->dim == 2 ? 1 : g->Nz;
    const int64_t cells = (int64_t)Nx * Ny * Nz;
    int32_t *inv = (int32_t *)malloc(sizeof(int32_t) * (size_t)cells);
    for (int64_t c = 0; c < cells; c++) inv[c] = -1;
    for (int64_t q = 0; q < n; q++) inv[order[q]] = (int32_t)q;
    double sum = 0.0;
    int cnt = 0;
#define AT(i, j, k) (((int64_t)(i) * Nx + (j)) * Nz + (k))
#define ACC(ri, rj, rk, i1, j1, k1, i2, j2, k2, sR, s1, s2)                                                      \
    do                                                                                                          \
    {                                                                                                           \
        const int64_t cR = AT(ri, rj, rk), c1 = AT(i1, j1, k1), c2 = AT(i2, j2, k2);                            \
        if (g->types[cR] == ORACLE_INFLOW && g->types[c1] == ORACLE_FLUID && g->types[c2] == ORACLE_FLUID)      \
        {                                                                                                       \
            const int64_t id1 = unknown_id(inv, c1), id2 = unknown_id(inv, c2);                                 \
            if (id1 >= 0 && id2 >= 0 && level[cR] == level[cR])                                                 \
            {                                                                                                   \
                const double p1 = x[id1], p2 = x[id2];                                                          \
                sum += level[cR] - (p1 + (p1 - p2) * ((sR) - (s1)) / ((s1) - (s2)));                            \
                cnt++;                                                                                          \
            }                                                                                                   \
        }                                                                                                       \
    } while (0)
    if (g->dim == 3)
    {
        for (int i = 1; i < Ny - 1; i++)
            for (int k = 1; k < Nz - 1; k++)
            {
                ACC(i, 0, k, i, 1, k, i, 2, k, cx[0], cx[1], cx[2]);
                ACC(i, Nx - 1, k, i, Nx - 2, k, i, Nx - 3, k, cx[Nx - 1], cx[Nx - 2], cx[Nx - 3]);
            }
        for (int j = 1; j < Nx - 1; j++)
            for (int k = 1; k < Nz - 1; k++)
            {
                ACC(0, j, k, 1, j, k, 2, j, k, cy[0], cy[1], cy[2]);
                ACC(Ny - 1, j, k, Ny - 2, j, k, Ny - 3, j, k, cy[Ny - 1], cy[Ny - 2], cy[Ny - 3]);
            }
        for (int i = 1; i < Ny - 1; i++)
            for (int j = 1; j < Nx - 1; j++)
            {
                ACC(i, j, 0, i, j, 1, i, j, 2, cz[0], cz[1], cz[2]);
                ACC(i, j, Nz - 1, i, j, Nz - 2, i, j, Nz - 3, cz[Nz - 1], cz[Nz - 2], cz[Nz - 3]);
            }
    }
    else
    {
        for (int i = 1; i < Ny - 1; i++)
        {
            ACC(i, 0, 0, i, 1, 0, i, 2, 0, cx[0], cx[1], cx[2]);
            ACC(i, Nx - 1, 0, i, Nx - 2, 0, i, Nx - 3, 0, cx[Nx - 1], cx[Nx - 2], cx[Nx - 3]);
        }
        for (int j = 1; j < Nx - 1; j++)
        {
            ACC(0, j, 0, 1, j, 0, 2, j, 0, cy[0], cy[1], cy[2]);
            ACC(Ny - 1, j, 0, Ny - 2, j, 0, Ny - 3, j, 0, cy[Ny - 1], cy[Ny - 2], cy[Ny - 3]);
        }
    }
#undef ACC
#undef AT
    free(inv);
    if (cnt > 0)
    {
        const double shift = sum / cnt;
        for (int64_t q = 0; q < n; q++) x[q] += shift;
    }
    return cnt;
}

/* BuildWarmStartGuess (PressureSolver3D.cpp:222-244; PressureSolver2D.cpp:191-212) once WarmStartEnabled holds:
 * x0 = p_prev at the unknowns; no guess (returns 0) when that is zero everywhere; the mean removed when `level`
 * (NeedsCompatibility). */
int oracle_warm_start_guess(const oracle_grid *g, const int32_t *order, int64_t n, const double *pPrev, int level,
                            double *x0)
{
    (void)g;
    double sq = 0.0;
    for (int64_t q = 0; q < n; q++)
    {
        x0[q] = pPrev[order[q]];
        sq += x0[q] * x0[q];
    }
    if (sq == 0.0) return 0;
    if (level) remove_mean(x0, n);
    return 1;
}
