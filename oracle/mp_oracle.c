/*
 * mp_oracle.c -- CPU restatement of MantlePar's H/C transport step.
 *
 * TEST INFRASTRUCTURE ONLY: correctness checker and timed CPU baseline for the
 * HIP library.  Loaded only by tests/, the smoke check and bench.py's
 * cpu_baseline leg.  Each block cites the reference file:line it restates.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   phase                golden vectors written by the reference's own
 *                        src/phase compiled unmodified (oracle/_ref), plus the
 *                        closed forms of src/phase/test/test_eutectic.cpp.
 *   gauss, geometry      polynomial exactness, closed forms.
 *   stencil, weno        Q1/Q2 reproduction, analytic indicators, the
 *                        direct-formula weights, extreme weights, boundary
 *                        candidate counts (src/mlweno/test/test_reconstruction.cpp).
 *   flux, operator       LLF closed form, constant-state preservation, global
 *                        balance, brute force on tiny meshes (local and global
 *                        LF speed), the global speed's closed forms.
 *   SSP-RK3              stability polynomial and third-order convergence
 *                        (not in the reference: it uses SSPRK2, evolve.cpp:469-479).
 * The mlweno sources need PETSc and LAPACKE, which this image lacks, so they are
 * unbuildable here; nothing of theirs is compiled.
 */
#include "mp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static _Thread_local char g_error[256];

const char* mpo_last_error(void) { return g_error; }

static int fail(int code, const char* message)
{
    snprintf(g_error, sizeof g_error, "%s", message);
    return code;
}

#define CHECK(cond, code, msg) do { if (!(cond)) return fail((code), (msg)); } while (0)
#define CALL(expr) do { int e_ = (expr); if (e_) return e_; } while (0)

static int finite_d(double v) { return isfinite(v); }

/* =============================================================== geometry */

/* Gauss-Legendre points on [-1,1], ascending.  The reference obtains the rule
 * from PetscDTGaussQuadrature (src/core/integral.cpp:7-20); this is Newton's
 * method on P_n, exact to rounding. */
int mpo_gauss_rule(int n, double* points, double* weights)
{
    CHECK(n >= 1 && n <= 64, MPO_ERR_ARG, "Gauss quadrature requires 1..64 points");
    for (int k = 0; k < n; ++k) {
        long double x = cosl(3.14159265358979323846264338327950288L * (k + 0.75L) / (n + 0.5L));
        long double dp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1, p1 = x;
            for (int m = 2; m <= n; ++m) {
                long double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
                p0 = p1; p1 = p2;
            }
            if (n == 1) { p1 = x; p0 = 1; }
            dp = n * (x * p1 - p0) / (x * x - 1);
            long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-19L) break;
        }
        {
            long double p0 = 1, p1 = x;
            for (int m = 2; m <= n; ++m) {
                long double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
                p0 = p1; p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
        }
        /* descending x for ascending k; store ascending */
        points[n - 1 - k] = (double)x;
        weights[n - 1 - k] = (double)(2.0L / ((1 - x * x) * dp * dp));
    }
    if (n % 2 == 1) points[n / 2] = 0.0;
    for (int k = 0; k < n / 2; ++k) {   /* exact symmetry */
        double a = 0.5 * (points[n - 1 - k] - points[k]);
        points[k] = -a; points[n - 1 - k] = a;
        double w = 0.5 * (weights[k] + weights[n - 1 - k]);
        weights[k] = w; weights[n - 1 - k] = w;
    }
    return MPO_OK;
}

/* src/core/integral.cpp:91-97 */
mpo_point mpo_map_edge_point(double ref, const mpo_point edge[2])
{
    const double a = 0.5 * (1.0 - ref), b = 0.5 * (1.0 + ref);
    mpo_point r = {{a * edge[0].p[0] + b * edge[1].p[0], a * edge[0].p[1] + b * edge[1].p[1]}};
    return r;
}

/* src/core/integral.cpp:99-106 */
mpo_point mpo_map_cell_point(mpo_point ref, const mpo_point c[4])
{
    const mpo_point lower[2] = {c[0], c[1]}, upper[2] = {c[3], c[2]};
    const mpo_point vertical[2] = {mpo_map_edge_point(ref.p[0], lower),
                                   mpo_map_edge_point(ref.p[0], upper)};
    return mpo_map_edge_point(ref.p[1], vertical);
}

/* src/core/integral.cpp:108-121 */
double mpo_cell_jacobian(mpo_point ref, const mpo_point c[4])
{
    const double s = 0.5 * (1.0 + ref.p[0]), t = 0.5 * (1.0 + ref.p[1]);
    double dxi[2], deta[2];
    for (int d = 0; d < 2; ++d) {
        dxi[d] = 0.5 * ((1.0 - t) * (c[1].p[d] - c[0].p[d]) + t * (c[2].p[d] - c[3].p[d]));
        deta[d] = 0.5 * ((1.0 - s) * (c[3].p[d] - c[0].p[d]) + s * (c[2].p[d] - c[1].p[d]));
    }
    return dxi[0] * deta[1] - dxi[1] * deta[0];
}

/* src/core/integral.cpp:62-89 */
int mpo_edge_geometry(const mpo_point e[2], double* length, mpo_point* normal)
{
    for (int k = 0; k < 2; ++k)
        CHECK(finite_d(e[k].p[0]) && finite_d(e[k].p[1]), MPO_ERR_ARG, "Edge coordinates must be finite");
    const double dx = e[1].p[0] - e[0].p[0], dy = e[1].p[1] - e[0].p[1];
    CHECK(finite_d(dx) && finite_d(dy), MPO_ERR_ARG, "Edge displacement must be finite");
    const double scale = fmax(fabs(dx), fabs(dy));
    CHECK(scale > 0.0, MPO_ERR_ARG, "Edge endpoints must be distinct");
    const double ux = dx / scale, uy = dy / scale;
    const double norm = sqrt(ux * ux + uy * uy);
    const double len = scale * norm;
    CHECK(finite_d(len) && len > 0.0, MPO_ERR_ARG, "Edge length must be finite and positive");
    *length = len;
    normal->p[0] = uy / norm;
    normal->p[1] = -ux / norm;
    return MPO_OK;
}

/* src/core/quad_validation.h:15-46 */
int mpo_validate_quad(const mpo_point c[4])
{
    double xmin = c[0].p[0], xmax = xmin, ymin = c[0].p[1], ymax = ymin;
    for (int k = 0; k < 4; ++k) {
        CHECK(finite_d(c[k].p[0]) && finite_d(c[k].p[1]), MPO_ERR_ARG,
              "Quadrilateral coordinates must be finite");
        xmin = fmin(xmin, c[k].p[0]); xmax = fmax(xmax, c[k].p[0]);
        ymin = fmin(ymin, c[k].p[1]); ymax = fmax(ymax, c[k].p[1]);
    }
    const double sx = xmax - xmin, sy = ymax - ymin;
    CHECK(finite_d(sx) && finite_d(sy) && sx > 0.0 && sy > 0.0, MPO_ERR_ARG,
          "Quadrilateral extents must be finite and positive");
    const double tol = 64.0 * DBL_EPSILON;
    for (int k = 0; k < 4; ++k) {
        const mpo_point* p = &c[k]; const mpo_point* nx = &c[(k + 1) % 4];
        const mpo_point* pv = &c[(k + 3) % 4];
        const double ux = (nx->p[0] - p->p[0]) / sx, uy = (nx->p[1] - p->p[1]) / sy;
        const double vx = (pv->p[0] - p->p[0]) / sx, vy = (pv->p[1] - p->p[1]) / sy;
        CHECK(ux * vy - uy * vx > tol, MPO_ERR_ARG,
              "Quadrilateral must be counterclockwise, convex and nonsingular");
    }
    return MPO_OK;
}

typedef struct { int n; double x[16], w[16]; } rule1d;

static int make_rule(int n, rule1d* r)
{
    r->n = n;
    return mpo_gauss_rule(n, r->x, r->w);
}

typedef double (*cell_fn)(mpo_point p, const void* ctx);

/* IntegrateCell, src/core/integral.h:49-78 (double accumulation, j outer). */
static int integrate_cell(const mpo_point c[4], const rule1d* r, cell_fn f, const void* ctx,
                          double* out)
{
    CALL(mpo_validate_quad(c));
    double sum = 0.0;
    for (int j = 0; j < r->n; ++j)
        for (int i = 0; i < r->n; ++i) {
            mpo_point ref = {{r->x[i], r->x[j]}};
            mpo_point p = mpo_map_cell_point(ref, c);
            double J = mpo_cell_jacobian(ref, c);
            CHECK(finite_d(p.p[0]) && finite_d(p.p[1]) && finite_d(J) && J > 0.0, MPO_ERR_FP,
                  "Cell mapping produced an invalid point or Jacobian");
            double v = f(p, ctx);
            CHECK(finite_d(v), MPO_ERR_FP, "Cell integrand returned a non-finite value");
            sum += r->w[i] * r->w[j] * J * v;
        }
    CHECK(finite_d(sum), MPO_ERR_FP, "Cell integral is not finite");
    *out = sum;
    return MPO_OK;
}

static double fn_one(mpo_point p, const void* ctx) { (void)p; (void)ctx; return 1.0; }

typedef struct { int i, j; } mono_ctx;

static double power_int(double x, int d)
{
    double r = 1;
    for (int k = 0; k < d; ++k) r *= x;
    return r;
}

static double fn_monomial(mpo_point p, const void* ctx)
{
    const mono_ctx* m = (const mono_ctx*)ctx;
    return power_int(p.p[0], m->i) * power_int(p.p[1], m->j);
}

/* ====================================================== TensorStencilPoly */

/* src/mlweno/tensorstencilpoly.cpp:51-57 */
static double falling(int degree, int derivative)
{
    double r = 1;
    for (int k = 0; k < derivative; ++k) r *= (double)(degree - k);
    return r;
}

/* Horner in normalized coordinates, src/mlweno/tensorstencilpoly.cpp:61-74 */
static double normalized_derivative(const double* c, int si, int sj, double x, double y,
                                    int dx, int dy)
{
    if (dx >= si || dy >= sj) return 0;
    double result = 0;
    for (int j = sj; j-- > dy;) {
        double row = 0;
        for (int i = si; i-- > dx;) row = row * x + c[j * si + i] * falling(i, dx);
        result = result * y + row * falling(j, dy);
    }
    return result;
}

static int normalize_quad(const mpo_point c[4], mpo_point center, double h, mpo_point out[4])
{
    mpo_point w[4];
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 2; ++d) w[k].p[d] = (c[k].p[d] - center.p[d]) / h;
    CALL(mpo_validate_quad(w));
    memcpy(out, w, sizeof w);
    return MPO_OK;
}

/* Dense LU with partial pivoting and solve, the dgesv contract used at
 * src/mlweno/tensorstencilpoly.cpp:209-215 (row-major, B overwritten by X). */
static int dense_solve(int n, double* A, double* B)
{
    int piv[MPO_MAX_STENCIL];
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); p = i; }
        piv[k] = p;
        if (A[p * n + k] == 0.0) return fail(MPO_ERR_LIB, "Stencil moment solve failed (singular)");
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t;
            }
        const double inv = 1.0 / A[k * n + k];
        for (int i = k + 1; i < n; ++i) {
            A[i * n + k] *= inv;
            const double l = A[i * n + k];
            for (int j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
        }
    }
    for (int k = 0; k < n; ++k)       /* row interchanges of B */
        if (piv[k] != k)
            for (int j = 0; j < n; ++j) {
                double t = B[k * n + j]; B[k * n + j] = B[piv[k] * n + j]; B[piv[k] * n + j] = t;
            }
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {           /* L y = b (unit lower) */
            double s = B[i * n + c];
            for (int k = 0; k < i; ++k) s -= A[i * n + k] * B[k * n + c];
            B[i * n + c] = s;
        }
        for (int i = n - 1; i >= 0; --i) {      /* U x = y */
            double s = B[i * n + c];
            for (int k = i + 1; k < n; ++k) s -= A[i * n + k] * B[k * n + c];
            B[i * n + c] = s / A[i * n + i];
        }
    }
    return MPO_OK;
}

/* BuildSmoothness, src/mlweno/tensorstencilpoly.cpp:318-368 */
static int build_smoothness(const mpo_stencil* s, const mpo_point nc[4], int gp, double* matrix)
{
    rule1d r;
    CALL(make_rule(gp, &r));
    double area = 0;
    CALL(integrate_cell(nc, &r, fn_one, NULL, &area));
    CHECK(area > 0, MPO_ERR_FP, "Smoothness integration area must be positive");
    const int n = s->count;
    double work[MPO_MAX_STENCIL * MPO_MAX_STENCIL];
    double der[MPO_MAX_STENCIL];
    memset(work, 0, sizeof(double) * n * n);
    for (int qy = 0; qy < r.n; ++qy)
        for (int qx = 0; qx < r.n; ++qx) {
            mpo_point ref = {{r.x[qx], r.x[qy]}};
            mpo_point p = mpo_map_cell_point(ref, nc);
            const double weight = r.w[qx] * r.w[qy] * (mpo_cell_jacobian(ref, nc) / area);
            CHECK(finite_d(weight) && weight > 0, MPO_ERR_FP, "Invalid smoothness quadrature weight");
            for (int dy = 0; dy < s->size_j; ++dy)
                for (int dx = 0; dx < s->size_i; ++dx) {
                    if (dx + dy == 0 || dx + dy > s->order) continue;
                    for (int c = 0; c < n; ++c) {
                        der[c] = normalized_derivative(s->coef + c * n, s->size_i, s->size_j,
                                                       p.p[0], p.p[1], dx, dy);
                        CHECK(finite_d(der[c]), MPO_ERR_FP, "Smoothness basis derivative is not finite");
                    }
                    for (int row = 0; row < n; ++row)
                        for (int col = 0; col <= row; ++col)
                            work[row * n + col] += weight * der[row] * der[col];
                }
        }
    for (int row = 0; row < n; ++row)
        for (int col = 0; col <= row; ++col) {
            CHECK(finite_d(work[row * n + col]), MPO_ERR_FP, "Smoothness matrix is not finite");
            work[col * n + row] = work[row * n + col];
        }
    memcpy(matrix, work, sizeof(double) * n * n);
    return MPO_OK;
}

/* TensorStencilPoly::Initialize(corners, ...), src/mlweno/tensorstencilpoly.cpp:146-231 */
int mpo_stencil_init(mpo_stencil* out, const mpo_point (*corners)[4], int start_i, int start_j,
                     int size_i, int size_j, int order, double h)
{
    CHECK(size_i > 0 && size_j > 0 && size_i * size_j <= MPO_MAX_STENCIL, MPO_ERR_ARG,
          "Stencil dimensions must be positive and fit the oracle's stencil limit");
    CHECK(start_i >= 0 && start_j >= 0, MPO_ERR_ARG, "Stencil origin or last cell index is out of range");
    CHECK(order >= 0 && finite_d(h) && h > 0, MPO_ERR_ARG,
          "Smoothness order must be nonnegative and h finite and positive");
    const int n = size_i * size_j;
    for (int c = 0; c < n; ++c) CALL(mpo_validate_quad(corners[c]));
    mpo_stencil* w = (mpo_stencil*)malloc(sizeof(mpo_stencil));
    if (!w) return fail(MPO_ERR_MEM, "out of memory");
    memset(w, 0, sizeof *w);
    w->start_i = start_i; w->start_j = start_j; w->size_i = size_i; w->size_j = size_j;
    w->count = n; w->order = order; w->h = h; w->target_index = -1;
    memcpy(w->corners, corners, sizeof(mpo_point) * 4 * n);

    /* centre: mean of the four outer vertices, as differences from the anchor (:163-176) */
    const mpo_point anchor = corners[0][0];
    const mpo_point outer[4] = {anchor, corners[size_i - 1][1], corners[n - 1][2],
                                corners[(size_j - 1) * size_i][3]};
    for (int d = 0; d < 2; ++d) {
        double offset = 0;
        for (int k = 0; k < 4; ++k) offset += (outer[k].p[d] - anchor.p[d]) / 4;
        w->center.p[d] = anchor.p[d] + offset;
        if (!finite_d(w->center.p[d])) { free(w); return fail(MPO_ERR_FP, "Stencil center is not finite"); }
    }

    const int gp = size_i > size_j ? (size_i > 3 ? size_i : 3) : (size_j > 3 ? size_j : 3);
    rule1d r;
    int e = make_rule(gp, &r);
    double moments[MPO_MAX_STENCIL * MPO_MAX_STENCIL], inverse[MPO_MAX_STENCIL * MPO_MAX_STENCIL];
    memset(inverse, 0, sizeof inverse);
    for (int c = 0; c < n && !e; ++c) {
        mpo_point nc[4];
        double area = 0;
        e = normalize_quad(corners[c], w->center, h, nc);
        if (!e) e = integrate_cell(nc, &r, fn_one, NULL, &area);
        if (!e && !(area > 0)) e = fail(MPO_ERR_FP, "Normalized cell area must be positive");
        for (int j = 0; j < size_j && !e; ++j)
            for (int i = 0; i < size_i && !e; ++i) {
                mono_ctx m = {i, j};
                double moment = 0;
                e = integrate_cell(nc, &r, fn_monomial, &m, &moment);
                moments[c * n + j * size_i + i] = moment / area;
                if (!e && !finite_d(moments[c * n + j * size_i + i]))
                    e = fail(MPO_ERR_FP, "Cell moment is not finite");
            }
        inverse[c * n + c] = 1;
    }
    if (!e) e = dense_solve(n, moments, inverse);
    for (int c = 0; c < n && !e; ++c)
        for (int k = 0; k < n; ++k) {
            const double v = inverse[k * n + c];
            if (!finite_d(v)) { e = fail(MPO_ERR_FP, "Stencil solve produced a non-finite coefficient"); break; }
            w->coef[c * n + k] = v;
        }
    if (!e) {
        const mpo_point ref[4] = {{{-0.5, -0.5}}, {{0.5, -0.5}}, {{0.5, 0.5}}, {{-0.5, 0.5}}};
        e = build_smoothness(w, ref, gp, w->sigma);
    }
    if (!e) memcpy(out, w, sizeof *w);
    free(w);
    return e;
}

/* SetTargetSmoothness, src/mlweno/tensorstencilpoly.cpp:383-398 */
int mpo_stencil_set_target(mpo_stencil* s, int li, int lj)
{
    CHECK(s->count > 0, MPO_ERR_STATE, "Initialize the stencil before selecting a target cell");
    CHECK(li >= 0 && li < s->size_i && lj >= 0 && lj < s->size_j, MPO_ERR_ARG,
          "Target must be a local cell offset within the stencil");
    const int index = lj * s->size_i + li;
    mpo_point nc[4];
    CALL(normalize_quad(s->corners[index], s->center, s->h, nc));
    const int degree = (s->size_i - 1) + (s->size_j - 1);
    double matrix[MPO_MAX_STENCIL * MPO_MAX_STENCIL];
    CALL(build_smoothness(s, nc, degree > 3 ? degree : 3, matrix));
    memcpy(s->target_sigma, matrix, sizeof(double) * s->count * s->count);
    s->target_index = index;
    return MPO_OK;
}

static int check_averages(const mpo_stencil* s, const double* a)
{
    CHECK(s->count > 0, MPO_ERR_STATE, "Initialize the stencil before evaluation");
    for (int k = 0; k < s->count; ++k)
        CHECK(finite_d(a[k]), MPO_ERR_ARG, "Cell averages must be finite");
    return MPO_OK;
}

/* ApplySmoothness, src/mlweno/tensorstencilpoly.cpp:400-459 (value only) */
static int apply_smoothness(const mpo_stencil* s, const double* matrix, const double* a,
                            double* value)
{
    CALL(check_averages(s, a));
    const int n = s->count;
    long double d[MPO_MAX_STENCIL];
    for (int i = 0; i < n; ++i) d[i] = (long double)a[i] - a[0];
    long double sum = 0, magnitude = 0;
    for (int row = 0; row < n; ++row)
        for (int col = 0; col < n; ++col) {
            const long double term = d[row] * matrix[row * n + col] * d[col];
            sum += term;
            magnitude += fabsl(term);
        }
    CHECK(isfinite(sum) && isfinite(magnitude), MPO_ERR_FP, "Smoothness evaluation overflowed");
    const long double tol = 64 * (long double)DBL_EPSILON * magnitude;
    CHECK(sum >= -tol, MPO_ERR_FP, "Smoothness quadratic form is negative beyond roundoff");
    const double result = (double)(sum > 0 ? sum : 0.0L);
    CHECK(finite_d(result), MPO_ERR_FP, "Smoothness indicator is not representable");
    *value = result;
    return MPO_OK;
}

int mpo_stencil_smoothness(const mpo_stencil* s, const double* a, int use_target, double* value)
{
    if (use_target) {
        CHECK(s->target_index >= 0, MPO_ERR_STATE,
              "Call SetTargetSmoothness for this target cell before querying it");
        return apply_smoothness(s, s->target_sigma, a, value);
    }
    return apply_smoothness(s, s->sigma, a, value);
}

static int stencil_eval_at(const mpo_stencil* s, const double* a, mpo_point point,
                           mpo_point center, double h, int dx, int dy, double* value);

/* EvaluateDerivative with PolynomialValue, src/mlweno/tensorstencilpoly.cpp:247-316 */
int mpo_stencil_eval(const mpo_stencil* s, const double* a, mpo_point point, int dx, int dy,
                     double* value)
{
    return stencil_eval_at(s, a, point, s->center, s->h, dx, dy, value);
}

static int stencil_eval_at(const mpo_stencil* s, const double* a, mpo_point point,
                           mpo_point center, double h, int dx, int dy, double* value)
{
    CALL(check_averages(s, a));
    const int n = s->count;
    double poly[MPO_MAX_STENCIL];
    memset(poly, 0, sizeof(double) * n);
    poly[0] = a[0];
    for (int c = 0; c < n; ++c) {
        const double diff = a[c] - a[0];
        CHECK(finite_d(diff), MPO_ERR_FP, "Cell-average difference is not representable");
        for (int k = 0; k < n; ++k) poly[k] += diff * s->coef[c * n + k];
    }
    for (int k = 0; k < n; ++k)
        CHECK(finite_d(poly[k]), MPO_ERR_FP, "Reconstructed coefficient is not finite");
    CHECK(dx >= 0 && dy >= 0 && finite_d(point.p[0]) && finite_d(point.p[1]), MPO_ERR_ARG,
          "Derivative orders must be nonnegative and the point finite");
    const double x = (point.p[0] - center.p[0]) / h;
    const double y = (point.p[1] - center.p[1]) / h;
    CHECK(finite_d(x) && finite_d(y), MPO_ERR_FP, "Normalized evaluation point is not finite");
    double result = normalized_derivative(poly, s->size_i, s->size_j, x, y, dx, dy);
    if (dx < s->size_i && dy < s->size_j) {
        for (int k = 0; k < dx; ++k) result /= h;
        for (int k = 0; k < dy; ++k) result /= h;
    }
    CHECK(finite_d(result), MPO_ERR_FP, "Polynomial value or derivative is not finite");
    *value = result;
    return MPO_OK;
}

/* ========================================================= Reconstruction */

void mpo_production_options(mpo_weno_options* o)
{
    /* src/couple/state_reconstruction.cpp:64-68; defaults src/mlweno/reconstruction.h:28-34 */
    memset(o, 0, sizeof *o);
    o->large.size_i = 3; o->large.size_j = 3; o->large.order = 2;
    o->large.noffsets = 1; o->large.off_i[0] = -1; o->large.off_j[0] = -1;
    o->small.size_i = 2; o->small.size_j = 2; o->small.order = 1;
    o->small.noffsets = 4;
    const int si[4] = {0, 0, -1, -1}, sj[4] = {-1, 0, -1, 0};
    for (int k = 0; k < 4; ++k) { o->small.off_i[k] = si[k]; o->small.off_j[k] = sj[k]; }
    o->small.target_smoothness = 1;
    o->use_constant = 0;
    o->constant_linear_weight = 1e-5;
    o->epsilon = 1e-2;
    o->s = 1;
}

void mpo_cell_corners(const mpo_mesh* m, int i, int j, mpo_point c[4])
{
    const int ci[4] = {i, i + 1, i + 1, i}, cj[4] = {j, j, j + 1, j + 1};
    for (int k = 0; k < 4; ++k) {
        const double* v = m->coords + ((size_t)cj[k] * (size_t)(m->nx + 1) + (size_t)ci[k]) * 2;
        c[k].p[0] = v[0]; c[k].p[1] = v[1];
    }
}

/* MeshInfo::CacheAreas: one-point Gauss rule, src/core/mesh_info.cpp:287-300 */
double mpo_cell_area(const mpo_mesh* m, int i, int j)
{
    mpo_point c[4];
    mpo_cell_corners(m, i, j, c);
    const mpo_point ref = {{0.0, 0.0}};
    return 2.0 * 2.0 * mpo_cell_jacobian(ref, c);
}

/* src/mlweno/reconstruction.cpp:10-41 */
static int check_family(const mpo_family* f)
{
    CHECK(f->size_i > 0 && f->size_j > 0, MPO_ERR_ARG,
          "Stencil dimensions must be positive with a representable product");
    CHECK(f->order >= 0, MPO_ERR_ARG, "Smoothness order must be nonnegative");
    CHECK(f->noffsets >= 0 && f->noffsets <= MPO_MAX_CAND, MPO_ERR_ARG, "Too many offsets");
    for (int k = 0; k < f->noffsets; ++k) {
        CHECK(f->off_i[k] <= 0 && f->off_i[k] >= 1 - f->size_i && f->off_j[k] <= 0 &&
              f->off_j[k] >= 1 - f->size_j, MPO_ERR_ARG,
              "Every prescribed stencil must contain the target cell");
        const double w = f->has_weights ? f->linear_weight[k] : 1.0;
        CHECK(finite_d(w) && w >= 0, MPO_ERR_ARG, "Linear weights must be finite and nonnegative");
    }
    return MPO_OK;
}

int mpo_recon_storage_count(const mpo_weno_options* opt)
{
    return opt->large.noffsets + opt->small.noffsets;
}

/* Reconstruction::Initialize, src/mlweno/reconstruction.cpp:109-210 */
int mpo_recon_init(mpo_recon* out, const mpo_mesh* mesh, int ti, int tj,
                   const mpo_weno_options* opt, double h, mpo_stencil* storage,
                   const mpo_stencil* const* shared)
{
    CHECK(ti >= 0 && ti < mesh->nx && tj >= 0 && tj < mesh->ny, MPO_ERR_ARG,
          "Reconstruction target must be a physical cell");
    CHECK(finite_d(h) && h > 0 && finite_d(opt->epsilon) && opt->epsilon > 0 && opt->s >= 0,
          MPO_ERR_ARG, "h and epsilon must be finite and positive; s must be nonnegative");
    CHECK(finite_d(opt->constant_linear_weight) && opt->constant_linear_weight >= 0, MPO_ERR_ARG,
          "Constant linear weight must be finite and nonnegative");
    CALL(check_family(&opt->large));
    CALL(check_family(&opt->small));
    mpo_recon w;
    memset(&w, 0, sizeof w);
    w.nx = mesh->nx; w.ny = mesh->ny; w.ti = ti; w.tj = tj;
    w.req_bi = ti; w.req_bj = tj; w.req_ei = ti + 1; w.req_ej = tj + 1;
    w.opt = *opt;
    for (int f = 0; f < 2; ++f) {
        const mpo_family* fam = f == 0 ? &opt->large : &opt->small;
        for (int k = 0; k < fam->noffsets; ++k) {
            const double lw = fam->has_weights ? fam->linear_weight[k] : 1.0;
            if (lw == 0) continue;
            const int si = ti + fam->off_i[k], sj = tj + fam->off_j[k];
            if (si < 0 || sj < 0 || fam->size_i > mesh->nx || fam->size_j > mesh->ny ||
                si > mesh->nx - fam->size_i || sj > mesh->ny - fam->size_j)
                continue;   /* discarded, not shifted (:150-153) */
            const int slot = f == 0 ? k : opt->large.noffsets + k;
            mpo_candidate* c = &w.cand[w.ncand];
            c->family = f; c->configured_index = k; c->linear_index = k; c->linear_weight = lw;
            c->start_i = si; c->start_j = sj; c->h = h;
            {
                /* stencil centre: mean of the four outer vertices (tensorstencilpoly.cpp:163-176) */
                mpo_point q0[4], q1[4], q2[4], q3[4];
                mpo_cell_corners(mesh, si, sj, q0);
                mpo_cell_corners(mesh, si + fam->size_i - 1, sj, q1);
                mpo_cell_corners(mesh, si + fam->size_i - 1, sj + fam->size_j - 1, q2);
                mpo_cell_corners(mesh, si, sj + fam->size_j - 1, q3);
                const mpo_point anchor = q0[0];
                const mpo_point outer[4] = {anchor, q1[1], q2[2], q3[3]};
                for (int d = 0; d < 2; ++d) {
                    double offset = 0;
                    for (int m = 0; m < 4; ++m) offset += (outer[m].p[d] - anchor.p[d]) / 4;
                    c->center.p[d] = anchor.p[d] + offset;
                }
            }
            if (shared) {
                CHECK(shared[slot] != NULL, MPO_ERR_STATE, "Missing shared stencil table");
                c->stencil = shared[slot];
            } else {
                mpo_point corners[MPO_MAX_STENCIL][4];
                for (int jj = 0; jj < fam->size_j; ++jj)
                    for (int ii = 0; ii < fam->size_i; ++ii)
                        mpo_cell_corners(mesh, si + ii, sj + jj, corners[jj * fam->size_i + ii]);
                CALL(mpo_stencil_init(&storage[slot], (const mpo_point (*)[4])corners, si, sj,
                                      fam->size_i, fam->size_j, fam->order, h));
                if (fam->target_smoothness)
                    CALL(mpo_stencil_set_target(&storage[slot], -fam->off_i[k], -fam->off_j[k]));
                c->stencil = &storage[slot];
            }
            if (si < w.req_bi) w.req_bi = si;
            if (sj < w.req_bj) w.req_bj = sj;
            if (si + fam->size_i > w.req_ei) w.req_ei = si + fam->size_i;
            if (sj + fam->size_j > w.req_ej) w.req_ej = sj + fam->size_j;
            ++w.ncand;
        }
    }
    CHECK(w.ncand > 0 || (opt->use_constant && opt->constant_linear_weight > 0), MPO_ERR_ARG,
          "No positive-weight candidate remains at this target cell");
    *out = w;
    return MPO_OK;
}

static int eta(int r) { return r == 1 ? 1 : (r == 2 ? 3 : 4); }

/* LogDenominator, src/mlweno/reconstruction.cpp:46-54 */
static long double log_denominator(double sigma, long double log_reg)
{
    if (sigma == 0) return log_reg;
    const long double ls = logl((long double)sigma);
    const long double hi = ls > log_reg ? ls : log_reg, lo = ls > log_reg ? log_reg : ls;
    return hi + log1pl(expl(lo - hi));
}

/* NormalizeWeights, src/mlweno/reconstruction.cpp:212-265 */
static int normalize_weights(const mpo_recon* r, const double* sigma, double area,
                             double* weights /* ncand+1 */)
{
    long double la[MPO_MAX_CAND + 1];
    for (int k = 0; k <= r->ncand; ++k) la[k] = -INFINITY;
    const long double lreg = logl((long double)r->opt.epsilon) + logl((long double)area);
    for (int k = 0; k < r->ncand; ++k) {
        const mpo_family* fam = r->cand[k].family == 0 ? &r->opt.large : &r->opt.small;
        const int rr = fam->order + 1;
        const long double ex = (long double)r->opt.s * rr + eta(rr);
        la[k] = logl((long double)r->cand[k].linear_weight) - ex * log_denominator(sigma[k], lreg);
        CHECK(isfinite(la[k]), MPO_ERR_FP, "Polynomial log weight is not finite");
    }
    if (r->opt.use_constant && r->opt.constant_linear_weight > 0) {
        const long double ex = (long double)r->opt.s + eta(1);
        la[r->ncand] = logl((long double)r->opt.constant_linear_weight) - ex * lreg;
        CHECK(isfinite(la[r->ncand]), MPO_ERR_FP, "Constant log weight is not finite");
    }
    long double mx = la[0];
    for (int k = 1; k <= r->ncand; ++k) if (la[k] > mx) mx = la[k];
    CHECK(isfinite(mx), MPO_ERR_FP, "No finite active log weight");
    long double sum = 0;
    for (int k = 0; k <= r->ncand; ++k) { la[k] = expl(la[k] - mx); sum += la[k]; }
    CHECK(isfinite(sum) && sum > 0, MPO_ERR_FP, "Invalid ML-WENO weight normalization");
    for (int k = 0; k <= r->ncand; ++k) weights[k] = (double)(la[k] / sum);
    return MPO_OK;
}

/* Reconstruction::UpdateImpl (value path), src/mlweno/reconstruction.cpp:275-355 */
int mpo_recon_update(mpo_recon* r, const double* values, int bi, int bj, int ei, int ej,
                     double area)
{
    CHECK(finite_d(area) && area > 0, MPO_ERR_ARG, "Weight area scale must be finite and positive");
    CHECK(bi >= 0 && bj >= 0 && ei > bi && ej > bj && ei <= r->nx && ej <= r->ny, MPO_ERR_ARG,
          "Cell-average range must be a nonempty rectangle of physical cells");
    CHECK(bi <= r->req_bi && bj <= r->req_bj && ei >= r->req_ei && ej >= r->req_ej, MPO_ERR_ARG,
          "Cell averages are unavailable locally; update/increase the solution halo");
    const size_t pitch = (size_t)(ei - bi);
#define PATCH(ii, jj) values[(size_t)((jj) - bj) * pitch + (size_t)((ii) - bi)]
    const double ta = PATCH(r->ti, r->tj);
    CHECK(finite_d(ta), MPO_ERR_ARG, "Target cell average must be finite");
    double centered[MPO_MAX_CAND][MPO_MAX_STENCIL];
    double sigma[MPO_MAX_CAND], weights[MPO_MAX_CAND + 1];
    for (int k = 0; k < r->ncand; ++k) {
        const mpo_stencil* s = r->cand[k].stencil;
        for (int j = 0; j < s->size_j; ++j)
            for (int i = 0; i < s->size_i; ++i) {
                const double avg = PATCH(r->cand[k].start_i + i, r->cand[k].start_j + j);
                CHECK(finite_d(avg), MPO_ERR_ARG, "Referenced cell averages must be finite");
                const double diff = avg - ta;
                CHECK(finite_d(diff), MPO_ERR_FP, "Centered cell average is not representable");
                centered[k][j * s->size_i + i] = diff;
            }
        const mpo_family* fam = r->cand[k].family == 0 ? &r->opt.large : &r->opt.small;
        CALL(mpo_stencil_smoothness(s, centered[k], fam->target_smoothness, &sigma[k]));
    }
#undef PATCH
    CALL(normalize_weights(r, sigma, area, weights));
    for (int k = 0; k < r->ncand; ++k) {
        memcpy(r->cand[k].centered, centered[k], sizeof(double) * r->cand[k].stencil->count);
        r->cand[k].smoothness = sigma[k];
        r->cand[k].weight = weights[k];
    }
    r->target_average = ta;
    r->constant_weight = weights[r->ncand];
    r->area = area;
    r->has_weights = 1;
    return MPO_OK;
}

/* Reconstruction::EvaluateDerivative, src/mlweno/reconstruction.cpp:477-502 */
int mpo_recon_eval(const mpo_recon* r, mpo_point x, int dx, int dy, double* value)
{
    CHECK(r->has_weights, MPO_ERR_STATE, "Update reconstruction before evaluation");
    CHECK(dx >= 0 && dy >= 0 && finite_d(x.p[0]) && finite_d(x.p[1]), MPO_ERR_ARG,
          "Derivative orders must be nonnegative and the point finite");
    long double result = (dx == 0 && dy == 0) ? r->target_average : 0;
    for (int k = 0; k < r->ncand; ++k) {
        if (r->cand[k].weight == 0) continue;
        double residual = 0;
        CALL(stencil_eval_at(r->cand[k].stencil, r->cand[k].centered, x, r->cand[k].center,
                             r->cand[k].h, dx, dy, &residual));
        result += (long double)r->cand[k].weight * residual;
    }
    CHECK(isfinite(result) && fabsl(result) <= DBL_MAX, MPO_ERR_FP,
          "Reconstructed value is not representable");
    *value = (double)result;
    return MPO_OK;
}

/* ========================================================= advective flux */

static int real_value(long double v, double* out)
{
    CHECK(isfinite(v) && fabsl(v) <= DBL_MAX, MPO_ERR_FP, "Advective flux or derivative is not representable");
    *out = (double)v;
    return MPO_OK;
}

/* CheckPoint, src/mlweno/advectiveflux.cpp:20-37 (position/time omitted: the
 * linear law does not read them) */
static int check_point(mpo_point v, mpo_point n)
{
    for (int d = 0; d < 2; ++d)
        CHECK(finite_d(v.p[d]) && finite_d(n.p[d]), MPO_ERR_ARG,
              "Flux position, velocity and normal must be finite");
    const double norm = hypot(n.p[0], n.p[1]);
    CHECK(finite_d(norm) && fabs(norm - 1) <= 128 * DBL_EPSILON, MPO_ERR_ARG,
          "Advective flux requires a unit normal");
    return MPO_OK;
}

/* LinearAdvectiveFlux, src/mlweno/advectiveflux.cpp:272-285 */
static int linear_flux(double u, mpo_point v, mpo_point n, double* value, double* derivative)
{
    CALL(check_point(v, n));
    CHECK(finite_d(u), MPO_ERR_ARG, "Linear-advection state must be finite");
    double vn;
    CALL(real_value((long double)v.p[0] * n.p[0] + (long double)v.p[1] * n.p[1], &vn));
    CALL(real_value((long double)vn * u, value));
    *derivative = vn;
    return MPO_OK;
}

static int bounds_ok(double bound, double required)
{
    return bound >= required || required - bound <= 64 * DBL_EPSILON * fmax(bound, required);
}

/* FluxImpl (value path), src/mlweno/advectiveflux.cpp:129-164 */
int mpo_llf_point(double left, double right, mpo_point v, mpo_point n, int mode,
                  double global_speed, double* flux, double* local_speed)
{
    CALL(check_point(v, n));
    CHECK(mode == 0 || mode == 1, MPO_ERR_ARG, "Unknown LF mode");
    if (mode == 1)
        CHECK(finite_d(global_speed) && global_speed >= 0, MPO_ERR_ARG,
              "Global LF speed must be finite and nonnegative");
    double fl, dl, fr, dr;
    CHECK(finite_d(left), MPO_ERR_ARG, "Advective trace must be finite");
    CALL(linear_flux(left, v, n, &fl, &dl));
    CHECK(finite_d(right), MPO_ERR_ARG, "Advective trace must be finite");
    CALL(linear_flux(right, v, n, &fr, &dr));
    const double a = fabs(dl), b = fabs(dr);
    const double speed = a > b ? a : b;   /* LocalSpeed endpoint estimate (:91-127) */
    const double alpha = mode == 0 ? speed : global_speed;
    CHECK(bounds_ok(alpha, speed), MPO_ERR_ARG,
          "Supplied global LF speed is smaller than the current local bound");
    const long double jump = (long double)right - left;
    CALL(real_value(((long double)fl + fr - (long double)alpha * jump) / 2, flux));
    if (local_speed) *local_speed = speed;
    return MPO_OK;
}

/* EdgeImpl (value path), src/mlweno/advectiveflux.cpp:192-227 */
int mpo_integrate_edge_flux(const mpo_point edge[2], int nq, const double* qp, const double* qw,
                            const mpo_point* velocity, const double* left, const double* right,
                            int mode, double global_speed, double* flux)
{
    (void)qp;
    double length; mpo_point normal;
    CALL(mpo_edge_geometry(edge, &length, &normal));
    long double sum = 0;
    for (int q = 0; q < nq; ++q) {
        CHECK(finite_d(velocity[q].p[0]) && finite_d(velocity[q].p[1]), MPO_ERR_ARG,
              "Edge velocity must be finite");
        double f;
        CALL(mpo_llf_point(left[q], right[q], velocity[q], normal, mode, global_speed, &f, NULL));
        const long double weight = (long double)qw[q] * length / 2;
        sum += weight * f;
    }
    return real_value(sum, flux);
}

/* BuildLinearAdvectionBoundaryData, src/mlweno/advectiveflux.cpp:394-424 */
int mpo_boundary_data(const mpo_point edge[2], int nq, const double* qp, const double* qw,
                      const mpo_point* velocity, const double* inflow_values, int* types,
                      double* values)
{
    (void)qp; (void)qw;
    double length; mpo_point normal;
    CALL(mpo_edge_geometry(edge, &length, &normal));
    for (int q = 0; q < nq; ++q) {
        CALL(check_point(velocity[q], normal));
        double vn;
        CALL(real_value((long double)velocity[q].p[0] * normal.p[0] +
                        (long double)velocity[q].p[1] * normal.p[1], &vn));
        if (vn < 0) {
            CHECK(inflow_values != NULL, MPO_ERR_ARG, "An inflow point requires prescribed boundary data");
            CHECK(finite_d(inflow_values[q]), MPO_ERR_FP, "Inflow boundary callback must return a finite state");
            types[q] = 0; values[q] = inflow_values[q];
        } else if (vn > 0) {
            types[q] = 1; values[q] = 0;
        } else {
            types[q] = 2; values[q] = 0;
        }
    }
    return MPO_OK;
}

/* BoundaryImpl (value path), src/mlweno/advectiveflux.cpp:229-268 */
int mpo_integrate_boundary_flux(const mpo_point edge[2], int nq, const double* qp,
                                const double* qw, const mpo_point* velocity,
                                const double* interior, const int* types, const double* values,
                                double* flux)
{
    (void)qp;
    double length; mpo_point normal;
    CALL(mpo_edge_geometry(edge, &length, &normal));
    long double sum = 0;
    for (int q = 0; q < nq; ++q) {
        CHECK(types[q] >= 0 && types[q] <= 2, MPO_ERR_ARG, "Unknown advection boundary policy");
        CHECK(finite_d(interior[q]), MPO_ERR_ARG, "Interior boundary trace must be finite");
        CALL(check_point(velocity[q], normal));
        if (types[q] == 2) continue;
        double f, d;
        CALL(linear_flux(types[q] == 1 ? interior[q] : values[q], velocity[q], normal, &f, &d));
        const long double weight = (long double)qw[q] * length / 2;
        sum += weight * f;
    }
    return real_value(sum, flux);
}

/* ============================================================ diffusive flux */

/* CreateLagrangeDerivativeWeights, src/mlweno/diffusiveflux.cpp:203-254:
 * first-derivative weights of the Lagrange interpolant through npts equally
 * spaced nodes 0..n (unit spacing), at the middle (n/2, even npts only) or at
 * an end node. */
int mpo_lagrange_weights(int npts, int location, double* w)
{
    CHECK(npts >= 2, MPO_ERR_ARG, "A first-derivative rule requires at least two points");
    CHECK(location >= 0 && location <= 2, MPO_ERR_ARG, "Unknown Lagrange derivative location");
    CHECK(location != MPO_LAGRANGE_MIDDLE || npts % 2 == 0, MPO_ERR_ARG,
          "Legacy midpoint derivative requires an even number of points");
    CHECK(npts <= MPO_MAX_SAMPLES + 1, MPO_ERR_ARG, "Too many Lagrange nodes");
    const int n = npts - 1;
    double work[MPO_MAX_SAMPLES + 1];
    if (location == MPO_LAGRANGE_MIDDLE) {
        /* d/dx of the interpolant at x = n/2: w_i = f_i C(n,i) / (n - 2i)^2,
         * antisymmetric (w_{n-i} = -w_i), |f_i| = 2^(1-n) prod_{odd k<=n} k/(k-1)
         * with the sign alternating from the outermost node inward. */
        long double factor = exp2l(1.0L - n);
        for (int k = 3; k <= n; k += 2) factor *= (long double)k / (k - 1);
        if ((npts / 2) % 2) factor = -factor;
        CHECK(factor != 0 && isfinite(factor), MPO_ERR_FP,
              "Lagrange order exceeds representable coefficient range");
        long double choose = 1;
        for (int i = 0; i < npts / 2; ++i) {
            const long double d = (long double)n - 2.0L * i;
            CALL(real_value(factor * choose / (d * d), &work[i]));
            work[n - i] = -work[i];
            choose *= (long double)(n - i) / (i + 1);
            factor = -factor;
        }
    } else {
        /* end node n: w_n = H_n (harmonic number), w_i = (-1)^(n-i) C(n,i)/(n-i) */
        long double harmonic = 0;
        for (int j = 1; j <= n; ++j) harmonic += 1.0L / j;
        CALL(real_value(harmonic, &work[n]));
        long double choose = 1, sign = n % 2 ? -1 : 1;
        for (int i = 0; i < n; ++i) {
            CALL(real_value(sign * choose / (n - i), &work[i]));
            choose *= (long double)(n - i) / (i + 1);
            sign = -sign;
        }
        if (location == MPO_LAGRANGE_END_LO) {
            for (int i = 0; i < npts / 2; ++i) {
                const double tmp = work[i]; work[i] = work[n - i]; work[n - i] = tmp;
            }
            for (int i = 0; i < npts; ++i) work[i] = -work[i];
        }
    }
    memcpy(w, work, (size_t)npts * sizeof(double));
    return MPO_OK;
}

static int same_point(mpo_point a, mpo_point b) { return a.p[0] == b.p[0] && a.p[1] == b.p[1]; }

/* FindFace, src/mlweno/diffusiveflux.cpp:24-40 */
static int find_face(const mpo_point cell[4], const mpo_point edge[2], int reverse, int* face)
{
    CALL(mpo_validate_quad(cell));
    for (int f = 0; f < 4; ++f)
        if (same_point(cell[f], edge[reverse ? 1 : 0]) && same_point(cell[(f + 1) % 4], edge[reverse ? 0 : 1])) {
            *face = f;
            return MPO_OK;
        }
    return fail(MPO_ERR_ARG, "Edge must match the cell vertices with the required orientation");
}

/* RayDistance, src/mlweno/diffusiveflux.cpp:42-74: distance along a unit ray
 * from a point of face `face` to the first of the cell's other faces. */
static int ray_distance(const mpo_point cell[4], int face, mpo_point o, mpo_point dir, double* out)
{
    long double exit_ = INFINITY;
    for (int f = 0; f < 4; ++f) {
        if (f == face) continue;
        const mpo_point a = cell[f], b = cell[(f + 1) % 4];
        const long double tx = (long double)b.p[0] - a.p[0], ty = (long double)b.p[1] - a.p[1];
        const long double rx = (long double)o.p[0] - a.p[0], ry = (long double)o.p[1] - a.p[1];
        const long double inside = tx * ry - ty * rx;
        CHECK(isfinite(inside) && inside > 0, MPO_ERR_ARG,
              "Edge quadrature point must lie strictly between the other cell faces");
        const long double rate = tx * dir.p[1] - ty * dir.p[0];
        CHECK(isfinite(rate), MPO_ERR_FP, "Normal-ray intersection overflowed");
        if (rate < 0) {
            const long double d = inside / (-rate);
            if (d < exit_) exit_ = d;
        }
    }
    double work;
    CALL(real_value(exit_, &work));
    CHECK(work > 0, MPO_ERR_ARG, "Normal sampling requires positive available distance");
    *out = work;
    return MPO_OK;
}

/* ShiftPoint, src/mlweno/diffusiveflux.cpp:76-86 */
static int shift_point(mpo_point p, mpo_point n, long double d, mpo_point* out)
{
    mpo_point w;
    for (int k = 0; k < 2; ++k) CALL(real_value((long double)p.p[k] + d * n.p[k], &w.p[k]));
    *out = w;
    return MPO_OK;
}

/* DiffusiveSampling::Build, src/mlweno/diffusiveflux.cpp:256-338.  right ==
 * NULL builds the boundary form (n samples inside plus the face node). */
int mpo_diffusive_sampling(const mpo_point edge[2], const mpo_point left[4], const mpo_point* right,
                           int nq, const double* qp, const double* qw, int n, double extent,
                           mpo_sampling* out)
{
    CHECK(n >= 2 && n % 2 == 0 && n <= MPO_MAX_SAMPLES, MPO_ERR_ARG,
          "Diffusion requires an even sample count >= 2 with room for a boundary node");
    CHECK(finite_d(extent) && extent > 0 && extent <= 1, MPO_ERR_ARG,
          "Normal sampling extentFraction must lie in (0,1]");
    CHECK(nq >= 1 && nq <= 16, MPO_ERR_ARG, "Edge quadrature needs 1..16 points");
    mpo_sampling s;
    memset(&s, 0, sizeof s);
    CALL(mpo_edge_geometry(edge, &s.length, &s.normal));
    int lf = 0, rf = 0;
    CALL(find_face(left, edge, 0, &lf));
    if (right) CALL(find_face(right, edge, 1, &rf));
    s.boundary = right == NULL;
    s.nq = nq;
    s.count = s.boundary ? n + 1 : n;
    CALL(mpo_lagrange_weights(s.count, s.boundary ? MPO_LAGRANGE_END_HI : MPO_LAGRANGE_MIDDLE,
                              s.coefficients));
    const mpo_point inward = {{-s.normal.p[0], -s.normal.p[1]}};
    const int m = s.count;
    for (int q = 0; q < nq; ++q) {
        CHECK(qp[q] > -1 && qp[q] < 1, MPO_ERR_ARG, "Normal sampling needs open-edge quadrature points");
        s.points[q] = mpo_map_edge_point(qp[q], edge);
        double available, other;
        CALL(ray_distance(left, lf, s.points[q], inward, &available));
        if (right) {
            CALL(ray_distance(right, rf, s.points[q], s.normal, &other));
            if (other < available) available = other;
        }
        const long double spacing = (long double)extent * available *
                                    (s.boundary ? 1.0L / n : 2.0L / (n - 1));
        CALL(real_value(spacing, &s.spacing[q]));
        CHECK(s.spacing[q] > 0, MPO_ERR_FP, "Normal sample spacing underflowed");
        CALL(real_value((long double)s.length * qw[q] / 2, &s.weights[q]));
        CHECK(s.weights[q] > 0, MPO_ERR_FP, "Physical edge quadrature weight underflowed");
        for (int k = 0; k < m; ++k) {
            const long double offset = s.boundary ? (long double)k - n : (long double)k - (n - 1) / 2.0L;
            const int a = q * m + k;
            CALL(shift_point(s.points[q], s.normal, offset * s.spacing[q], &s.samples[a]));
            s.sides[a] = s.boundary ? (k + 1 == m ? MPO_SIDE_BOUNDARY : MPO_SIDE_LEFT)
                                    : (k < m / 2 ? MPO_SIDE_LEFT : MPO_SIDE_RIGHT);
            long double signed_distance = 0;
            for (int d = 0; d < 2; ++d)
                signed_distance += ((long double)s.samples[a].p[d] - s.points[q].p[d]) * s.normal.p[d];
            CHECK((offset < 0 && signed_distance < 0) || (offset > 0 && signed_distance > 0) ||
                  (offset == 0 && signed_distance == 0), MPO_ERR_FP,
                  "Coordinates cannot resolve the requested normal sample offset");
            if (k > 0) {
                long double step = 0;
                for (int d = 0; d < 2; ++d)
                    step += ((long double)s.samples[a].p[d] - s.samples[a - 1].p[d]) * s.normal.p[d];
                CHECK(step > 0, MPO_ERR_FP, "Coordinates cannot resolve distinct normal samples");
            }
        }
    }
    *out = s;
    return MPO_OK;
}

/* IntegrateImpl, value path with the identity diffused quantity,
 * src/mlweno/diffusiveflux.cpp:102-200 and :364-372.  bc_type/bc_value are
 * per quadrature point for a boundary sampling (NULL for interior);
 * samples[q*count+k] ignores the boundary slot. */
int mpo_diffusive_flux(const mpo_sampling* s, const double* samples, const double* kappa,
                       const int* bc_type, const double* bc_value, double* flux)
{
    CHECK(s->boundary == (bc_type != NULL), MPO_ERR_ARG,
          "Use the matching interior or boundary flux function");
    const int m = s->count;
    double values[MPO_MAX_SAMPLES + 1];
    long double integral = 0;
    for (int q = 0; q < s->nq; ++q) {
        const double weight = s->weights[q];
        const int type = bc_type ? bc_type[q] : MPO_DIFF_STATE;
        if (bc_type) {
            CHECK(type >= MPO_DIFF_STATE && type <= MPO_DIFF_ZERO_FLUX, MPO_ERR_ARG,
                  "Unknown diffusion boundary condition");
            if (type == MPO_DIFF_ZERO_FLUX) continue;
            CHECK(finite_d(bc_value[q]), MPO_ERR_ARG, "Boundary data must be finite");
            if (type == MPO_DIFF_FLUX) {
                integral += (long double)weight * bc_value[q];
                continue;
            }
        }
        CHECK(finite_d(kappa[q]) && kappa[q] >= 0, MPO_ERR_ARG,
              "Diffusivity must be finite and nonnegative");
        for (int k = 0; k < m; ++k) {
            const double u = (bc_type && k + 1 == m) ? bc_value[q] : samples[q * m + k];
            CHECK(finite_d(u), MPO_ERR_ARG, "Diffusion sample state must be finite");
            values[k] = u;
        }
        long double gradient = 0;
        for (int k = 0; k + 1 < m; ++k)
            gradient += (long double)s->coefficients[k] * ((long double)values[k] - values[m - 1]);
        gradient /= s->spacing[q];
        CHECK(isfinite(gradient), MPO_ERR_FP, "Sampled normal derivative overflowed");
        integral -= (long double)weight * kappa[q] * gradient;
    }
    return real_value(integral, flux);
}

/* =================================================================== phase */

void mpo_material_defaults(mpo_material* p)
{
    /* src/phase/eutectic_rescaled.h:10-26 */
    p->Tm0 = 2000.0; p->Te0 = 1480.0; p->nu = 6.5e6; p->L = 5.0e5; p->cp = 1200.0; p->Xe = 0.25;
    p->rho = 3000.0; p->rhor = 500.0; p->mus = 1.0e19; p->mul = 1.0; p->k0 = 1.0e-8;
    p->g = 10.0; p->alpha0 = 3.0e-5;
}

static int checked_double(long double v, double* out)
{
    CHECK(isfinite(v) && fabsl(v) <= (long double)DBL_MAX, MPO_ERR_FP, "value is not representable");
    *out = (double)v;
    return MPO_OK;
}

static int positive_scale(long double v, double* out)
{
    CALL(checked_double(v, out));
    CHECK(*out > 0.0, MPO_ERR_FP, "value is not a positive scale");
    return MPO_OK;
}

#define REQ_FINITE(v) CHECK(finite_d(v), MPO_ERR_ARG, #v " must be finite")
#define REQ_POSITIVE(v) do { REQ_FINITE(v); CHECK((v) > 0.0, MPO_ERR_ARG, #v " must be positive"); } while (0)

/* derive, src/phase/eutectic_rescaled.cpp:42-86 */
int mpo_phase_set_parameters(mpo_phase_model* m, const mpo_material* p)
{
    REQ_POSITIVE(p->Tm0);
    REQ_POSITIVE(p->Te0);
    CHECK(p->Tm0 > p->Te0, MPO_ERR_ARG, "Tm0 must be greater than Te0");
    REQ_FINITE(p->nu);
    CHECK(p->nu != 0.0, MPO_ERR_ARG, "nu must be nonzero");
    REQ_POSITIVE(p->L);
    REQ_POSITIVE(p->cp);
    REQ_FINITE(p->Xe);
    CHECK(p->Xe > 0.0 && p->Xe < 1.0, MPO_ERR_ARG, "Xe must lie strictly between 0 and 1");
    REQ_POSITIVE(p->rho);
    REQ_POSITIVE(p->rhor);
    REQ_POSITIVE(p->mus);
    REQ_POSITIVE(p->mul);
    REQ_POSITIVE(p->k0);
    REQ_POSITIVE(p->g);
    REQ_FINITE(p->alpha0);
    mpo_derived d;
    const long double deltaT = (long double)p->Tm0 - p->Te0;
    const long double h0 = (long double)p->cp * deltaT;
    CALL(positive_scale(deltaT, &d.dT));
    CALL(positive_scale(h0, &d.h0));
    CALL(checked_double(1.0L / p->nu, &d.gamma));
    CHECK(d.gamma != 0.0, MPO_ERR_FP, "gamma underflows to zero");
    CALL(positive_scale(p->L / h0, &d.LD));
    CALL(positive_scale(p->Tm0 / deltaT, &d.TDm0));
    CALL(positive_scale(p->Te0 / deltaT, &d.TDe0));
    CALL(positive_scale(1.0L / p->k0, &d.invk0));
    const long double l0 = sqrtl((long double)p->mus * p->k0 / p->mul);
    const long double u0 = (long double)p->k0 * p->rhor * p->g / p->mul;
    CALL(positive_scale(l0, &d.l0));
    CALL(positive_scale(u0, &d.u0));
    CALL(positive_scale((long double)p->rhor * p->g * l0, &d.p0));
    CALL(positive_scale(l0 / u0, &d.t0));
    m->p = *p;
    m->d = d;
    return MPO_OK;
}

/* boundaries, src/phase/eutectic_rescaled.cpp:143-149, 163-185 (Xe passed in so
 * the hypereutectic mirror can reuse it) */
static int phase_bounds_xe(const mpo_phase_model* m, double Xe, double C, double P,
                           mpo_phase_bounds* b)
{
    CHECK(finite_d(C), MPO_ERR_ARG, "CD must be finite");
    CHECK(C >= 0.0 && C <= Xe, MPO_ERR_ARG, "CD must lie in [0, Xe]");
    CHECK(finite_d(m->d.TDe0) && finite_d(P), MPO_ERR_ARG, "TD and P must be finite");
    const long double shift = (long double)P / m->p.nu / m->d.dT;
    mpo_phase_bounds w;
    CALL(checked_double((long double)m->d.TDe0 + shift, &w.Tep));
    CALL(checked_double((long double)w.Tep + 1.0L, &w.Tmp));
    const long double ratio = (long double)C / Xe;
    CALL(checked_double((long double)w.Tep + m->d.LD * ratio, &w.H_eutectic_end));
    CALL(checked_double((long double)w.Tep + (1.0L - ratio) + m->d.LD, &w.H_liquidus));
    if (C == 0.0) CALL(checked_double((long double)w.Tmp + m->d.LD, &w.H_liquidus));
    CHECK(!(w.Tmp <= w.Tep || w.H_liquidus <= w.Tep || (C == 0.0 && w.H_liquidus <= w.Tmp)),
          MPO_ERR_FP, "Phase temperature/enthalpy intervals are not resolvable");
    *b = w;
    return MPO_OK;
}

int mpo_phase_boundaries(const mpo_phase_model* m, double C, double P, mpo_phase_bounds* b)
{
    return phase_bounds_xe(m, m->p.Xe, C, P, b);
}

/* selectRegion, src/phase/eutectic_rescaled.cpp:95-105 */
static int select_region(double H, double C, const mpo_phase_bounds* b)
{
    if (C == 0.0) {
        if (H <= b->Tmp) return 1;
        if (H < b->H_liquidus) return 6;
        return 5;
    }
    if (H <= b->Tep) return 2;
    if (H >= b->H_liquidus) return 5;
    if (H <= b->H_eutectic_end) return 3;
    return 4;
}

/* unitFraction, src/phase/eutectic_rescaled.cpp:108-114 */
static int unit_fraction(long double v, double* out)
{
    const long double tol = 64.0L * (long double)DBL_EPSILON;
    CHECK(isfinite(v) && v >= -tol && v <= 1.0L + tol, MPO_ERR_RUNTIME,
          "Computed phase fraction is outside [0, 1]");
    *out = (double)(v < 0 ? 0.0L : (v > 1 ? 1.0L : v));
    return MPO_OK;
}

static double negative_derivative(long double mag)
{
    if (mag > (long double)DBL_MAX) return -INFINITY;
    return -(double)mag;
}

/* EutecticModel::evaluate, src/phase/eutectic_rescaled.cpp:192-282 */
static int phase_eval_xe(const mpo_phase_model* m, double Xe_d, double H, double C, double P,
                         mpo_phase_state* out)
{
    CHECK(finite_d(H), MPO_ERR_ARG, "HD must be finite");
    mpo_phase_bounds b;
    CALL(phase_bounds_xe(m, Xe_d, C, P, &b));
    const long double latent = m->d.LD;
    const long double Xe = Xe_d;
    mpo_phase_state s;
    memset(&s, 0, sizeof s);
    s.finite_temperature_derivatives = 1;
    s.region = select_region(H, C, &b);
    s.Tep = b.Tep;
    switch (s.region) {
    case 1:
        s.phi1 = 1.0; s.TDp = H; s.dTD_dHD = 1.0;
        break;
    case 6:
        CALL(unit_fraction(((long double)H - b.Tmp) / latent, &s.phil));
        s.phi1 = 1.0 - s.phil;
        s.TDp = b.Tmp;
        break;
    case 2:
        s.phi1 = 1.0 - C; s.phi2 = C; s.TDp = H; s.dTD_dHD = 1.0;
        break;
    case 3: {
        const long double maxl = (long double)C / Xe;
        long double liquid;
        if (H == b.H_eutectic_end) liquid = maxl;
        else {
            const long double e = ((long double)H - b.Tep) / latent;
            liquid = maxl < e ? maxl : e;
        }
        long double solid2 = (long double)C - liquid * Xe;
        if (solid2 < 0) solid2 = 0;
        CALL(unit_fraction(liquid, &s.phil));
        CALL(unit_fraction(solid2, &s.phi2));
        CALL(unit_fraction(1.0L - liquid - solid2, &s.phi1));
        s.cl = Xe_d;
        s.TDp = b.Tep;
        break;
    }
    case 4: {
        const long double a = (long double)H - b.Tmp;
        const long double ratio = (long double)C / Xe;
        const long double root = hypotl(a, 2.0L * sqrtl(latent) * sqrtl(ratio));
        const long double liquid = a >= 0.0L ? (a + root) / (2.0L * latent) : 2.0L * ratio / (root - a);
        CALL(unit_fraction(liquid, &s.phil));
        CALL(unit_fraction(1.0L - liquid, &s.phi1));
        CALL(unit_fraction((long double)C / liquid, &s.cl));
        CALL(checked_double(a >= 0.0L ? (long double)b.Tmp - ratio / liquid
                                      : (long double)H - latent * liquid, &s.TDp));
        CALL(unit_fraction(a >= 0.0L ? 2.0L * (latent / root) * (ratio / (root + a))
                                     : 0.5L * (1.0L - a / root), &s.dTD_dHD));
        s.dTD_dCD = negative_derivative((latent / root) / Xe);
        break;
    }
    case 5:
        s.phil = 1.0; s.cl = C;
        CALL(checked_double((long double)H - latent, &s.TDp));
        s.dTD_dHD = 1.0;
        break;
    }
    if (C == 0.0 && H > b.Tep && s.region != 5) {
        const long double gap = fabsl((long double)H - b.Tmp);
        s.dTD_dCD = gap == 0.0L ? -INFINITY : negative_derivative((latent / gap) / Xe);
    }
    const double solid = s.phi1 + s.phi2;
    s.has_solid = solid > 0.0;
    s.has_liquid = s.phil > 0.0;
    s.cs = s.has_solid ? s.phi2 / solid : 0.0;
    if (!s.has_liquid) s.cl = 0.0;
    s.finite_temperature_derivatives = isfinite(s.dTD_dCD) && isfinite(s.dTD_dHD);
    *out = s;
    return MPO_OK;
}

int mpo_phase_evaluate(const mpo_phase_model* m, double H, double C, double P, mpo_phase_state* s)
{
    return phase_eval_xe(m, m->p.Xe, H, C, P, s);
}

int mpo_phase_classify(const mpo_phase_model* m, double H, double C, double P, int* region)
{
    CHECK(finite_d(H), MPO_ERR_ARG, "HD must be finite");
    mpo_phase_bounds b;
    CALL(mpo_phase_boundaries(m, C, P, &b));
    *region = select_region(H, C, &b);
    return MPO_OK;
}

/* staticPressure, src/phase/eutectic_rescaled.cpp:151-161 (the one-argument
 * form passes the model's l0; initialization.cpp:155-162 feeds it
 * surface_y - y and rejects a non-finite result). */
int mpo_static_pressure_scaled(const mpo_phase_model* m, double zD, double length, double* P)
{
    CHECK(finite_d(zD), MPO_ERR_ARG, "zD must be finite");
    CHECK(finite_d(length), MPO_ERR_ARG, "length_scale must be finite");
    CHECK(length > 0.0, MPO_ERR_ARG, "length_scale must be positive");
    const long double v = (long double)m->p.rho * m->p.g * zD * length;
    CHECK(isfinite(v) && fabsl(v) <= (long double)DBL_MAX, MPO_ERR_FP,
          "static pressure is not representable");
    *P = (double)v;
    return MPO_OK;
}

int mpo_static_pressure(const mpo_phase_model* m, double zD, double* P)
{
    return mpo_static_pressure_scaled(m, zD, m->d.l0, P);
}

/* Hypereutectic mirror (extension; the reference rejects C > Xe,
 * src/phase/eutectic_rescaled.cpp:88-93).  Component 2 plays the role of
 * component 1 on the mirrored axis C' = 1 - C with eutectic Xe' = 1 - Xe. */
int mpo_phase_evaluate_ext(const mpo_phase_model* m, double H, double C, double P,
                           mpo_phase_state* out)
{
    if (!(C > m->p.Xe)) return phase_eval_xe(m, m->p.Xe, H, C, P, out);
    CHECK(finite_d(C) && C <= 1.0, MPO_ERR_ARG, "CD must lie in [0, 1]");
    mpo_phase_state s;
    CALL(phase_eval_xe(m, 1.0 - m->p.Xe, H, 1.0 - C, P, &s));
    mpo_phase_state r = s;
    r.phi1 = s.phi2;
    r.phi2 = s.phi1;
    r.cl = s.has_liquid ? 1.0 - s.cl : 0.0;
    const double solid = r.phi1 + r.phi2;
    r.cs = s.has_solid ? r.phi2 / solid : 0.0;
    r.dTD_dCD = -s.dTD_dCD;
    *out = r;
    return MPO_OK;
}

int mpo_phase_batch(const mpo_phase_model* m, int64_t n, const double* H, const double* C,
                    double P, int hyper, double* T, double* phi, double* Cl, double* Cs,
                    uint8_t* region)
{
    int err = 0;
    char msg[256] = {0};
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        mpo_phase_state s;
        int e = hyper ? mpo_phase_evaluate_ext(m, H[k], C[k], P, &s)
                      : mpo_phase_evaluate(m, H[k], C[k], P, &s);
        if (e) {
#pragma omp critical
            { if (!err) { err = e; snprintf(msg, sizeof msg, "%s", g_error); } }
            continue;
        }
        if (T) T[k] = s.TDp;
        if (phi) phi[k] = s.phil;
        if (Cl) Cl[k] = s.cl;
        if (Cs) Cs[k] = s.cs;
        if (region) region[k] = (uint8_t)s.region;
    }
    if (err) return fail(err, msg);
    return MPO_OK;
}

/* ====================================================== transport operator */

struct mpo_transport {
    mpo_mesh mesh;
    double* coords;
    int nx, ny, q;
    double qp[16], qw[16];
    mpo_weno_options opt;
    int table_mode, threads, slots;
    mpo_stencil* stencils;            /* mode 0: [cell][slot]; mode 1: [slot] */
    double* area;                     /* [cell] */
    double* vnx;                      /* [q][ny][nx+1] */
    double* vny;                      /* [q][ny+1][nx] */
    uint8_t* bc[4];                   /* bottom, right, top, left */
    double* inflow[2][4];             /* field, side -> [k][q] canonical order */
    double* traces;                   /* [cell][4][q] */
    double* fv;                       /* vertical face flux  [ny][nx+1] (+x) */
    double* fh;                       /* horizontal face flux [ny+1][nx] (+y) */
    /* bound limiter (off unless set): cell rule, diffusion sampling, C bound */
    int limit, lim_cell_points, lim_samples;
    double lim_extent, lim_cupper;
    double* theta[2];                 /* last limiter factors per field [cell] */
    /* thermal diffusion of H (off unless set): kappa, sampling, boundary data */
    int diffuse, dif_samples;
    double kappa, dif_extent;
    uint8_t* tbc[4];                  /* [face][q]: 0 zero flux, 1 temperature, 2 flux */
    double* tval[4];                  /* [face][q] canonical order */
    uint8_t* bcC[4];                  /* advection policies of C (NULL: those of H) */
    int lf_mode;                      /* interior faces: 0 local, 1 global LF speed */
    double global_speed;
};

static size_t cidx(const mpo_transport* t, int i, int j) { return (size_t)j * t->nx + i; }

void mpo_transport_destroy(mpo_transport* t)
{
    if (!t) return;
    free(t->coords); free(t->stencils); free(t->area); free(t->vnx); free(t->vny);
    for (int s = 0; s < 4; ++s) { free(t->bc[s]); free(t->inflow[0][s]); free(t->inflow[1][s]); }
    free(t->traces); free(t->fv); free(t->fh); free(t->theta[0]); free(t->theta[1]);
    for (int s = 0; s < 4; ++s) { free(t->tbc[s]); free(t->tval[s]); free(t->bcC[s]); }
    free(t);
}

static void set_threads(const mpo_transport* t)
{
#ifdef _OPENMP
    if (t->threads > 0) omp_set_num_threads(t->threads);
#else
    (void)t;
#endif
}

static int side_len(const mpo_transport* t, int side) { return (side == 0 || side == 2) ? t->nx : t->ny; }

int mpo_transport_create(mpo_transport** out, const mpo_mesh* mesh, int q,
                         const mpo_weno_options* opt, const double* vnx, const double* vny,
                         const uint8_t* bc_bottom, const uint8_t* bc_right,
                         const uint8_t* bc_top, const uint8_t* bc_left,
                         const double* inflow_H, const double* inflow_C, int table_mode,
                         int threads)
{
    CHECK(mesh->nx >= 2 && mesh->ny >= 2, MPO_ERR_ARG, "Transport requires at least 2x2 cells");
    CHECK(q >= 1 && q <= 16, MPO_ERR_ARG, "Edge quadrature needs 1..16 points");
    CHECK(table_mode == 0 || table_mode == 1, MPO_ERR_ARG, "Unknown table mode");
    mpo_transport* t = (mpo_transport*)calloc(1, sizeof *t);
    if (!t) return fail(MPO_ERR_MEM, "out of memory");
    t->nx = mesh->nx; t->ny = mesh->ny; t->q = q; t->opt = *opt;
    t->table_mode = table_mode; t->threads = threads;
    t->slots = mpo_recon_storage_count(opt);
    const size_t N = (size_t)t->nx * t->ny;
    const size_t nv = (size_t)(t->nx + 1) * (t->ny + 1) * 2;
    t->coords = (double*)malloc(nv * sizeof(double));
    t->area = (double*)malloc(N * sizeof(double));
    t->vnx = (double*)malloc((size_t)q * t->ny * (t->nx + 1) * sizeof(double));
    t->vny = (double*)malloc((size_t)q * (t->ny + 1) * t->nx * sizeof(double));
    t->traces = (double*)malloc(N * 4 * q * sizeof(double));
    t->fv = (double*)malloc((size_t)t->ny * (t->nx + 1) * sizeof(double));
    t->fh = (double*)malloc((size_t)(t->ny + 1) * t->nx * sizeof(double));
    t->theta[0] = (double*)malloc(N * sizeof(double));
    t->theta[1] = (double*)malloc(N * sizeof(double));
    if (!t->coords || !t->area || !t->vnx || !t->vny || !t->traces || !t->fv || !t->fh ||
        !t->theta[0] || !t->theta[1]) {
        mpo_transport_destroy(t); return fail(MPO_ERR_MEM, "out of memory");
    }
    memcpy(t->coords, mesh->coords, nv * sizeof(double));
    t->mesh.nx = t->nx; t->mesh.ny = t->ny; t->mesh.coords = t->coords;
    memcpy(t->vnx, vnx, (size_t)q * t->ny * (t->nx + 1) * sizeof(double));
    memcpy(t->vny, vny, (size_t)q * (t->ny + 1) * t->nx * sizeof(double));
    const uint8_t* bcs[4] = {bc_bottom, bc_right, bc_top, bc_left};
    size_t inflow_offset = 0;
    for (int s = 0; s < 4; ++s) {
        const int n = side_len(t, s);
        t->bc[s] = (uint8_t*)malloc((size_t)n);
        for (int k = 0; k < n; ++k) t->bc[s][k] = bcs[s] ? bcs[s][k] : MPO_BC_ZERO;
        for (int f = 0; f < 2; ++f) {
            const double* src = f == 0 ? inflow_H : inflow_C;
            if (src) {
                t->inflow[f][s] = (double*)malloc((size_t)n * q * sizeof(double));
                memcpy(t->inflow[f][s], src + inflow_offset, (size_t)n * q * sizeof(double));
            }
        }
        inflow_offset += (size_t)n * q;
    }
    int e = mpo_gauss_rule(q, t->qp, t->qw);
    for (size_t c = 0; c < N && !e; ++c) {
        t->area[c] = mpo_cell_area(&t->mesh, (int)(c % t->nx), (int)(c / t->nx));
        if (!(t->area[c] > 0) || !finite_d(t->area[c])) e = fail(MPO_ERR_FP, "Cell area underflowed to zero");
    }
    if (!e && table_mode == 0) {
        t->stencils = (mpo_stencil*)malloc(N * t->slots * sizeof(mpo_stencil));
        if (!t->stencils) e = fail(MPO_ERR_MEM, "out of memory for per-cell stencil tables");
    } else if (!e) {
        t->stencils = (mpo_stencil*)calloc((size_t)t->slots, sizeof(mpo_stencil));
        /* one table per candidate type, from the first cell where it exists */
        int have[2 * MPO_MAX_CAND] = {0};
        int found = 0;
        for (size_t c = 0; c < N && !e && found < t->slots; ++c) {
            mpo_recon r;
            mpo_stencil tmp[2 * MPO_MAX_CAND];
            const int i = (int)(c % t->nx), j = (int)(c / t->nx);
            e = mpo_recon_init(&r, &t->mesh, i, j, opt, sqrt(t->area[c]), tmp, NULL);
            for (int k = 0; k < r.ncand && !e; ++k) {
                const int slot = r.cand[k].family == 0 ? r.cand[k].configured_index
                                                       : opt->large.noffsets + r.cand[k].configured_index;
                if (!have[slot]) { t->stencils[slot] = tmp[slot]; have[slot] = 1; ++found; }
            }
        }
    }
    if (!e && table_mode == 0) {
        int err = 0;
        char msg[256] = {0};
        set_threads(t);
#pragma omp parallel for schedule(dynamic, 64)
        for (size_t c = 0; c < N; ++c) {
            mpo_recon r;
            const int i = (int)(c % t->nx), j = (int)(c / t->nx);
            int ee = mpo_recon_init(&r, &t->mesh, i, j, opt, sqrt(t->area[c]),
                                    t->stencils + c * t->slots, NULL);
            if (ee) {
#pragma omp critical
                { if (!err) { err = ee; snprintf(msg, sizeof msg, "%s", g_error); } }
            }
        }
        if (err) e = fail(err, msg);
    }
    if (e) { mpo_transport_destroy(t); return e; }
    *out = t;
    return MPO_OK;
}

/* Build one cell's reconstruction over its production candidates. */
static int cell_recon(const mpo_transport* t, int i, int j, mpo_recon* r)
{
    const mpo_stencil* ptr[2 * MPO_MAX_CAND];
    const size_t c = cidx(t, i, j);
    for (int s = 0; s < t->slots; ++s)
        ptr[s] = t->table_mode == 0 ? t->stencils + c * t->slots + s : t->stencils + s;
    return mpo_recon_init(r, &t->mesh, i, j, &t->opt, sqrt(t->area[c]), NULL, ptr);
}

/* Face f of cell (i,j) in canonical orientation: 0 bottom (+x), 1 right (+y),
 * 2 top (+x), 3 left (+y). */
static void canonical_face(const mpo_transport* t, int i, int j, int f, mpo_point e[2])
{
    mpo_point c[4];
    mpo_cell_corners(&t->mesh, i, j, c);
    switch (f) {
    case 0: e[0] = c[0]; e[1] = c[1]; break;
    case 1: e[0] = c[1]; e[1] = c[2]; break;
    case 2: e[0] = c[3]; e[1] = c[2]; break;
    default: e[0] = c[0]; e[1] = c[3]; break;
    }
}

/* Lax-Friedrichs speed of the interior faces (FluxImpl's mode,
 * src/mlweno/advectiveflux.cpp:129-164): 0 local |v.n| per point, 1 the
 * global speed, which every point checks against its own bound. */
int mpo_transport_set_lf(mpo_transport* t, int mode, double global_speed)
{
    CHECK(mode == 0 || mode == 1, MPO_ERR_ARG, "Unknown LF mode");
    if (mode == 1)
        CHECK(finite_d(global_speed) && global_speed >= 0, MPO_ERR_ARG,
              "Global LF speed must be finite and nonnegative");
    t->lf_mode = mode;
    t->global_speed = mode == 1 ? global_speed : 0.0;
    return MPO_OK;
}

int mpo_transport_set_limiter(mpo_transport* t, int enable, int cell_points, int samples,
                              double extent, double c_upper)
{
    if (!enable) { t->limit = 0; return MPO_OK; }
    CHECK(cell_points >= 1 && cell_points <= 16, MPO_ERR_ARG, "Cell quadrature needs 1..16 points per axis");
    CHECK(samples >= 2 && samples % 2 == 0 && samples <= MPO_MAX_SAMPLES, MPO_ERR_ARG,
          "Diffusion requires an even sample count >= 2 with room for a boundary node");
    CHECK(finite_d(extent) && extent > 0 && extent <= 1, MPO_ERR_ARG,
          "Normal sampling extentFraction must lie in (0,1]");
    CHECK(finite_d(c_upper) && c_upper > 0, MPO_ERR_ARG, "Composition bound must be positive");
    t->limit = 1; t->lim_cell_points = cell_points; t->lim_samples = samples;
    t->lim_extent = extent; t->lim_cupper = c_upper;
    return MPO_OK;
}

/* StateReconstruction::Limit, src/couple/state_reconstruction.cpp:136-192, for
 * one field of one cell (upper = +inf for H, Xe for C): the largest factor
 * theta <= 1 such that mean + theta*(R(p) - mean) stays within the 3x3
 * neighbourhood's bounds, clamped to [0, upper], at the cell centre, the cell
 * Gauss points, the edge Gauss points and the in-cell normal-line diffusion
 * samples of every edge; rounded inward by 1e-13. */
static int cell_limit(const mpo_transport* t, const mpo_recon* r, const double* u, int i, int j,
                      double upper, double* theta_out)
{
    const int nx = t->nx, ny = t->ny;
    const double mean = u[cidx(t, i, j)];
    CHECK(finite_d(mean) && mean >= 0 && mean <= upper, MPO_ERR_ARG,
          "Inadmissible cell average in coupled reconstruction");
    double lo = mean, hi = mean;
    for (int y = j > 0 ? j - 1 : 0; y < (j + 2 < ny ? j + 2 : ny); ++y)
        for (int x = i > 0 ? i - 1 : 0; x < (i + 2 < nx ? i + 2 : nx); ++x) {
            const double v = u[cidx(t, x, y)];
            lo = fmin(lo, v); hi = fmax(hi, v);
        }
    lo = fmax(0.0, lo);
    hi = fmin(upper, hi);
    rule1d cell;
    CALL(make_rule(t->lim_cell_points, &cell));
    mpo_point c[4];
    mpo_cell_corners(&t->mesh, i, j, c);
    double theta = 1.0;
#define LIMIT_AT(P) do { \
        double v_; CALL(mpo_recon_eval(r, (P), 0, 0, &v_)); \
        CHECK(finite_d(v_), MPO_ERR_FP, "Nonfinite reconstructed state"); \
        const double d_ = v_ - mean; \
        if (v_ > hi) theta = fmin(theta, (hi - mean) / d_); \
        if (v_ < lo) theta = fmin(theta, (lo - mean) / d_); \
    } while (0)
    const mpo_point centre = {{0.0, 0.0}};
    LIMIT_AT(mpo_map_cell_point(centre, c));
    for (int a = 0; a < cell.n; ++a)
        for (int b = 0; b < cell.n; ++b) {
            const mpo_point ref = {{cell.x[a], cell.x[b]}};
            LIMIT_AT(mpo_map_cell_point(ref, c));
        }
    /* neighbour across edge e (e = 0 bottom, 1 right, 2 top, 3 left) */
    const int ni[4] = {i, i + 1, i, i - 1}, nj[4] = {j - 1, j, j + 1, j};
    for (int e = 0; e < 4; ++e) {
        const mpo_point edge[2] = {c[e], c[(e + 1) % 4]};
        for (int q = 0; q < t->q; ++q) LIMIT_AT(mpo_map_edge_point(t->qp[q], edge));
        const int inside = ni[e] >= 0 && ni[e] < nx && nj[e] >= 0 && nj[e] < ny;
        mpo_point other[4];
        if (inside) mpo_cell_corners(&t->mesh, ni[e], nj[e], other);
        mpo_sampling s;
        CALL(mpo_diffusive_sampling(edge, c, inside ? other : NULL, t->q, t->qp, t->qw,
                                    t->lim_samples, t->lim_extent, &s));
        for (int a = 0; a < s.nq * s.count; ++a)
            if (s.sides[a] == MPO_SIDE_LEFT) LIMIT_AT(s.samples[a]);
    }
#undef LIMIT_AT
    *theta_out = theta < 1 ? fmax(0.0, theta * (1 - 1e-13)) : 1.0;
    return MPO_OK;
}

/* StateReconstruction::Update + trace evaluation,
 * src/couple/state_reconstruction.cpp:101-134: a trace is
 * mean + theta*(R(p) - mean), theta = 1 unless the limiter is on. */
static int field_traces(mpo_transport* t, int field, const double* u, double* sigma, double* omega)
{
    const size_t N = (size_t)t->nx * t->ny;
    int err = 0;
    char msg[256] = {0};
    set_threads(t);
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t c = 0; c < N; ++c) {
        if (err) continue;
        const int i = (int)(c % t->nx), j = (int)(c / t->nx);
        mpo_recon r;
        int e = cell_recon(t, i, j, &r);
        if (!e) e = mpo_recon_update(&r, u, 0, 0, t->nx, t->ny, t->area[c]);
        double theta = 1.0;
        if (!e && t->limit)
            e = cell_limit(t, &r, u, i, j, field == 0 ? INFINITY : t->lim_cupper, &theta);
        t->theta[field][c] = theta;
        for (int f = 0; f < 4 && !e; ++f) {
            mpo_point edge[2];
            canonical_face(t, i, j, f, edge);
            for (int q = 0; q < t->q && !e; ++q) {
                double v;
                e = mpo_recon_eval(&r, mpo_map_edge_point(t->qp[q], edge), 0, 0, &v);
                t->traces[(c * 4 + f) * t->q + q] = u[c] + theta * (v - u[c]);
            }
        }
        if (!e && sigma) {
            for (int k = 0; k < 5; ++k) { sigma[c * 5 + k] = 0; omega[c * 5 + k] = 0; }
            for (int k = 0; k < r.ncand; ++k) {
                const int slot = r.cand[k].family == 0 ? r.cand[k].configured_index
                                                       : t->opt.large.noffsets + r.cand[k].configured_index;
                if (slot < 5) { sigma[c * 5 + slot] = r.cand[k].smoothness; omega[c * 5 + slot] = r.cand[k].weight; }
            }
        }
        if (e) {
#pragma omp critical
            { if (!err) { err = e; snprintf(msg, sizeof msg, "%s", g_error); } }
        }
    }
    if (err) return fail(err, msg);
    return MPO_OK;
}

int mpo_transport_limits(mpo_transport* t, const double* H, const double* C, double* thetaH,
                         double* thetaC)
{
    const size_t N = (size_t)t->nx * t->ny;
    CALL(field_traces(t, 0, H, NULL, NULL));
    memcpy(thetaH, t->theta[0], N * sizeof(double));
    CALL(field_traces(t, 1, C, NULL, NULL));
    memcpy(thetaC, t->theta[1], N * sizeof(double));
    return MPO_OK;
}

int mpo_transport_traces(mpo_transport* t, int field, const double* u, double* traces)
{
    CHECK(field == 0 || field == 1, MPO_ERR_ARG, "field must be 0 (H) or 1 (C)");
    CALL(field_traces(t, field, u, NULL, NULL));
    memcpy(traces, t->traces, (size_t)t->nx * t->ny * 4 * t->q * sizeof(double));
    return MPO_OK;
}

int mpo_transport_reconstruct(mpo_transport* t, const double* u, double* sigma, double* omega,
                              double* traces)
{
    CALL(field_traces(t, 0, u, sigma, omega));
    if (traces) memcpy(traces, t->traces, (size_t)t->nx * t->ny * 4 * t->q * sizeof(double));
    return MPO_OK;
}

static mpo_point axis_velocity(double vn, int vertical)
{
    mpo_point v = {{vertical ? vn : 0.0, vertical ? 0.0 : vn}};
    return v;
}

int mpo_transport_set_diffusion(mpo_transport* t, double kappa, int samples, double extent,
                                const uint8_t* types, const double* values)
{
    CHECK(finite_d(kappa) && kappa >= 0, MPO_ERR_ARG, "diffusivity must be nonnegative");
    CHECK(samples >= 2 && samples % 2 == 0 && samples <= MPO_MAX_SAMPLES, MPO_ERR_ARG,
          "Diffusion requires an even sample count >= 2 with room for a boundary node");
    CHECK(finite_d(extent) && extent > 0 && extent <= 1, MPO_ERR_ARG,
          "Normal sampling extentFraction must lie in (0,1]");
    size_t to = 0, vo = 0;
    for (int s = 0; s < 4; ++s) {
        const int n = side_len(t, s);
        free(t->tbc[s]); free(t->tval[s]);
        t->tbc[s] = (uint8_t*)calloc((size_t)n * t->q, 1);
        t->tval[s] = (double*)calloc((size_t)n * t->q, sizeof(double));
        if (!t->tbc[s] || !t->tval[s]) return fail(MPO_ERR_MEM, "out of memory");
        for (int k = 0; k < n * t->q; ++k) {
            t->tbc[s][k] = types ? types[to + k] : 0;
            CHECK(t->tbc[s][k] <= 2, MPO_ERR_ARG, "Unknown thermal boundary policy");
        }
        if (values) memcpy(t->tval[s], values + vo, (size_t)n * t->q * sizeof(double));
        to += (size_t)n * t->q;
        vo += (size_t)n * t->q;
    }
    t->diffuse = 1; t->kappa = kappa; t->dif_samples = samples; t->dif_extent = extent;
    return MPO_OK;
}

/* Legacy preheat diffusion of one face (src/couple/preheat.cpp:176-190): the
 * normal-line samples take the constant cell averages of H of their side;
 * a boundary face takes its thermal data (EvaluateThermalBoundary,
 * src/couple/initialization.cpp:490-506: temperature -> PrescribedQuantity,
 * flux -> PrescribedFlux).  edge is CCW for cell L; R < 0 marks a boundary
 * face of `side` (face index k, Gauss order reversed when rev).  Also returns
 * the rate terms |sum of sample partials| per side
 * (src/couple/preheat.cpp:137-148). */
static int face_diffusion(const mpo_transport* t, const mpo_point edge[2], int li, int lj, int ri,
                          int rj, const double* u, int side, int k, int rev, double* flux,
                          double* rate_left, double* rate_right)
{
    mpo_point lc[4], rc[4];
    mpo_cell_corners(&t->mesh, li, lj, lc);
    const int boundary = ri < 0;
    if (!boundary) mpo_cell_corners(&t->mesh, ri, rj, rc);
    mpo_sampling s;
    CALL(mpo_diffusive_sampling(edge, lc, boundary ? NULL : rc, t->q, t->qp, t->qw, t->dif_samples,
                                t->dif_extent, &s));
    double samples[16 * (MPO_MAX_SAMPLES + 1)], kappa[16], bv[16];
    int bt[16];
    const double uL = u[cidx(t, li, lj)], uR = boundary ? 0.0 : u[cidx(t, ri, rj)];
    for (int a = 0; a < s.nq * s.count; ++a)
        samples[a] = s.sides[a] == MPO_SIDE_LEFT ? uL : (s.sides[a] == MPO_SIDE_RIGHT ? uR : 0.0);
    for (int q = 0; q < s.nq; ++q) {
        kappa[q] = t->kappa;
        if (boundary) {
            const int qc = rev ? s.nq - 1 - q : q;
            const int pol = t->tbc[side] ? t->tbc[side][(size_t)k * s.nq + qc] : 0;
            bt[q] = pol == 1 ? MPO_DIFF_QUANTITY : (pol == 2 ? MPO_DIFF_FLUX : MPO_DIFF_ZERO_FLUX);
            bv[q] = t->tval[side] ? t->tval[side][(size_t)k * s.nq + qc] : 0.0;
        }
    }
    CALL(mpo_diffusive_flux(&s, samples, kappa, boundary ? bt : NULL, boundary ? bv : NULL, flux));
    /* sample partials of IntegrateImpl's derivative path (identity law):
     * -W_q kappa / ds_q times the coefficient, the last slot taking minus the
     * sum of the others; skipped points (zero or prescribed flux) give none */
    long double dl = 0, dr = 0;
    long double last = 0;
    for (int kk = 0; kk + 1 < s.count; ++kk) last -= s.coefficients[kk];
    for (int q = 0; q < s.nq; ++q) {
        if (boundary && bt[q] != MPO_DIFF_QUANTITY && bt[q] != MPO_DIFF_STATE) continue;
        const long double scale = -(long double)s.weights[q] * kappa[q] / s.spacing[q];
        for (int kk = 0; kk < s.count; ++kk) {
            const int a = q * s.count + kk;
            const long double coeff = kk + 1 == s.count ? last : s.coefficients[kk];
            if (s.sides[a] == MPO_SIDE_LEFT) dl += scale * coeff;
            else if (s.sides[a] == MPO_SIDE_RIGHT) dr += scale * coeff;
        }
    }
    if (rate_left) *rate_left = fabs((double)dl);
    if (rate_right) *rate_right = fabs((double)dr);
    return MPO_OK;
}

/* Interior and boundary face fluxes for one field from the stored traces.
 * Face loop and LLF integration: src/mlweno/test/transport_test_support.h:278-306,
 * src/mlweno/advectiveflux.cpp:192-227; boundary policy
 * src/couple/initialization.cpp:469-489 with src/mlweno/advectiveflux.cpp:229-268. */
static int face_fluxes(mpo_transport* t, int field, const double* u)
{
    const int diffuse = field == 0 && t->diffuse;
    const int nx = t->nx, ny = t->ny, nq = t->q;
    int err = 0;
    char msg[256] = {0};
    set_threads(t);
    /* vertical faces: x_i, i = 0..nx; the edge is cell (i-1,j)'s right side (CCW) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i <= nx && !err; ++i) {
            mpo_point vel[16];
            double left[16], right[16], flux = 0;
            int e = 0;
            if (i > 0 && i < nx) {
                mpo_point edge[2];
                canonical_face(t, i - 1, j, 1, edge);
                for (int q = 0; q < nq; ++q) {
                    vel[q] = axis_velocity(t->vnx[((size_t)q * ny + j) * (nx + 1) + i], 1);
                    left[q] = t->traces[(cidx(t, i - 1, j) * 4 + 1) * nq + q];
                    right[q] = t->traces[(cidx(t, i, j) * 4 + 3) * nq + q];
                }
                e = mpo_integrate_edge_flux(edge, nq, t->qp, t->qw, vel, left, right, t->lf_mode,
                                            t->global_speed, &flux);
                if (!e && diffuse) {
                    double d;
                    e = face_diffusion(t, edge, i - 1, j, i, j, u, 0, 0, 0, &d, NULL, NULL);
                    flux += d;
                }
            } else {
                /* boundary: interior cell's CCW edge, outward normal */
                const int side = i == 0 ? 3 : 1;
                const int ci = i == 0 ? 0 : nx - 1;
                mpo_point c4[4], edge[2];
                mpo_cell_corners(&t->mesh, ci, j, c4);
                if (side == 1) { edge[0] = c4[1]; edge[1] = c4[2]; }
                else { edge[0] = c4[3]; edge[1] = c4[0]; }
                const int rev = side == 3;     /* left edge runs along -y */
                int types[16]; double vals[16], inflow[16];
                const double* src = t->inflow[field][side];
                for (int q = 0; q < nq; ++q) {
                    const int qc = rev ? nq - 1 - q : q;
                    vel[q] = axis_velocity(t->vnx[((size_t)qc * ny + j) * (nx + 1) + i], 1);
                    left[q] = t->traces[(cidx(t, ci, j) * 4 + side) * nq + qc];
                    inflow[q] = src ? src[(size_t)j * nq + qc] : NAN;
                }
                double length; mpo_point normal;
                e = mpo_edge_geometry(edge, &length, &normal);
                const int policy = t->bc[side][j];
                for (int q = 0; q < nq && !e; ++q) {
                    const double speed = vel[q].p[0] * normal.p[0] + vel[q].p[1] * normal.p[1];
                    if (policy == MPO_BC_ZERO || speed == 0) { types[q] = 2; vals[q] = 0; }
                    else if (speed < 0) {
                        if (policy != MPO_BC_INFLOW_OUTFLOW) e = fail(MPO_ERR_BACKFLOW, "Backflow encountered on an outflow-only boundary.");
                        else if (!src) e = fail(MPO_ERR_ARG, "Missing inflow profile.");
                        else { types[q] = 0; vals[q] = inflow[q]; }
                    } else { types[q] = 1; vals[q] = 0; }
                }
                if (!e) e = mpo_integrate_boundary_flux(edge, nq, t->qp, t->qw, vel, left, types, vals, &flux);
                if (!e && diffuse) {
                    double d;
                    e = face_diffusion(t, edge, ci, j, -1, -1, u, side, j, rev, &d, NULL, NULL);
                    flux += d;
                }
                if (!e && side == 3) flux = -flux;   /* store along +x */
            }
            if (e) {
#pragma omp critical
                { if (!err) { err = e; snprintf(msg, sizeof msg, "%s", g_error); } }
            }
            t->fv[(size_t)j * (nx + 1) + i] = flux;
        }
    }
    /* horizontal faces: y_j, j = 0..ny; the edge is cell (i,j-1)'s top side (CCW, -x) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= ny; ++j) {
        for (int i = 0; i < nx && !err; ++i) {
            mpo_point vel[16];
            double left[16], right[16], flux = 0;
            int e = 0;
            if (j > 0 && j < ny) {
                mpo_point c4[4], edge[2];
                mpo_cell_corners(&t->mesh, i, j - 1, c4);
                edge[0] = c4[2]; edge[1] = c4[3];
                for (int q = 0; q < nq; ++q) {
                    const int qc = nq - 1 - q;
                    vel[q] = axis_velocity(t->vny[((size_t)qc * (ny + 1) + j) * nx + i], 0);
                    left[q] = t->traces[(cidx(t, i, j - 1) * 4 + 2) * nq + qc];
                    right[q] = t->traces[(cidx(t, i, j) * 4 + 0) * nq + qc];
                }
                e = mpo_integrate_edge_flux(edge, nq, t->qp, t->qw, vel, left, right, t->lf_mode,
                                            t->global_speed, &flux);
                if (!e && diffuse) {
                    double d;
                    e = face_diffusion(t, edge, i, j - 1, i, j, u, 0, 0, 0, &d, NULL, NULL);
                    flux += d;
                }
            } else {
                const int side = j == 0 ? 0 : 2;
                const int cj = j == 0 ? 0 : ny - 1;
                mpo_point c4[4], edge[2];
                mpo_cell_corners(&t->mesh, i, cj, c4);
                if (side == 0) { edge[0] = c4[0]; edge[1] = c4[1]; }
                else { edge[0] = c4[2]; edge[1] = c4[3]; }
                const int rev = side == 2;     /* top edge runs along -x */
                int types[16]; double vals[16], inflow[16];
                const double* src = t->inflow[field][side];
                for (int q = 0; q < nq; ++q) {
                    const int qc = rev ? nq - 1 - q : q;
                    vel[q] = axis_velocity(t->vny[((size_t)qc * (ny + 1) + j) * nx + i], 0);
                    left[q] = t->traces[(cidx(t, i, cj) * 4 + side) * nq + qc];
                    inflow[q] = src ? src[(size_t)i * nq + qc] : NAN;
                }
                double length; mpo_point normal;
                e = mpo_edge_geometry(edge, &length, &normal);
                const int policy = t->bc[side][i];
                for (int q = 0; q < nq && !e; ++q) {
                    const double speed = vel[q].p[0] * normal.p[0] + vel[q].p[1] * normal.p[1];
                    if (policy == MPO_BC_ZERO || speed == 0) { types[q] = 2; vals[q] = 0; }
                    else if (speed < 0) {
                        if (policy != MPO_BC_INFLOW_OUTFLOW) e = fail(MPO_ERR_BACKFLOW, "Backflow encountered on an outflow-only boundary.");
                        else if (!src) e = fail(MPO_ERR_ARG, "Missing inflow profile.");
                        else { types[q] = 0; vals[q] = inflow[q]; }
                    } else { types[q] = 1; vals[q] = 0; }
                }
                if (!e) e = mpo_integrate_boundary_flux(edge, nq, t->qp, t->qw, vel, left, types, vals, &flux);
                if (!e && diffuse) {
                    double d;
                    e = face_diffusion(t, edge, i, cj, -1, -1, u, side, i, rev, &d, NULL, NULL);
                    flux += d;
                }
                if (!e && side == 0) flux = -flux;   /* store along +y */
            }
            if (e) {
#pragma omp critical
                { if (!err) { err = e; snprintf(msg, sizeof msg, "%s", g_error); } }
            }
            t->fh[(size_t)j * nx + i] = flux;
        }
    }
    if (err) return fail(err, msg);
    return MPO_OK;
}

/* Gather form of the +-flux/area scatter, src/couple/preheat.cpp:193-196 */
static void gather_rhs(const mpo_transport* t, double* rhs, double* bflux)
{
    const int nx = t->nx, ny = t->ny;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double in = t->fv[(size_t)j * (nx + 1) + i] - t->fv[(size_t)j * (nx + 1) + i + 1]
                            + t->fh[(size_t)j * nx + i] - t->fh[(size_t)(j + 1) * nx + i];
            rhs[cidx(t, i, j)] = in / t->area[cidx(t, i, j)];
        }
    /* outward boundary total, fixed order */
    long double b = 0;
    for (int j = 0; j < ny; ++j) b += t->fv[(size_t)j * (nx + 1) + nx] - t->fv[(size_t)j * (nx + 1)];
    for (int i = 0; i < nx; ++i) b += t->fh[(size_t)ny * nx + i] - t->fh[i];
    *bflux = (double)b;
}

int mpo_transport_rhs(mpo_transport* t, const double* H, const double* C, double* rhsH,
                      double* rhsC, double* bfluxH, double* bfluxC)
{
    double bh = 0, bc = 0;
    CALL(field_traces(t, 0, H, NULL, NULL));
    CALL(face_fluxes(t, 0, H));
    gather_rhs(t, rhsH, &bh);
    CALL(field_traces(t, 1, C, NULL, NULL));
    CALL(face_fluxes(t, 1, C));
    gather_rhs(t, rhsC, &bc);
    if (bfluxH) *bfluxH = bh;
    if (bfluxC) *bfluxC = bc;
    return MPO_OK;
}

/* Per-cell rate, src/couple/preheat.cpp:125-152 (advection part) */
int mpo_transport_rates(mpo_transport* t, double* rate, double* max_rate)
{
    const int nx = t->nx, ny = t->ny, nq = t->q;
    const size_t N = (size_t)nx * ny;
    double* r = (double*)calloc(N, sizeof(double));
    double* dummy_u = (double*)calloc(N, sizeof(double));   /* rates do not depend on H */
    if (!r || !dummy_u) { free(r); free(dummy_u); return fail(MPO_ERR_MEM, "out of memory"); }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            mpo_point e[2];
            const int ci = i < nx ? i : nx - 1;
            canonical_face(t, ci, j, i < nx ? 3 : 1, e);
            double length; mpo_point n;
            CALL(mpo_edge_geometry(e, &length, &n));
            double speed = 0;
            for (int q = 0; q < nq; ++q)
                speed += .5 * length * t->qw[q] * fabs(t->vnx[((size_t)q * ny + j) * (nx + 1) + i]);
            double dl = 0, dr = 0;
            if (t->diffuse) {
                double f;
                mpo_point ce[2];
                if (i > 0 && i < nx) {
                    canonical_face(t, i - 1, j, 1, ce);
                    CALL(face_diffusion(t, ce, i - 1, j, i, j, dummy_u, 0, 0, 0, &f, &dl, &dr));
                } else {
                    mpo_point c4[4];
                    mpo_cell_corners(&t->mesh, ci, j, c4);
                    if (i == nx) { ce[0] = c4[1]; ce[1] = c4[2]; } else { ce[0] = c4[3]; ce[1] = c4[0]; }
                    CALL(face_diffusion(t, ce, ci, j, -1, -1, dummy_u, i == nx ? 1 : 3, j, i == 0,
                                        &f, i == 0 ? &dr : &dl, NULL));
                }
            }
            if (i > 0) r[cidx(t, i - 1, j)] += (speed + dl) / t->area[cidx(t, i - 1, j)];
            if (i < nx) r[cidx(t, i, j)] += (speed + dr) / t->area[cidx(t, i, j)];
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            mpo_point e[2];
            const int cj = j < ny ? j : ny - 1;
            canonical_face(t, i, cj, j < ny ? 0 : 2, e);
            double length; mpo_point n;
            CALL(mpo_edge_geometry(e, &length, &n));
            double speed = 0;
            for (int q = 0; q < nq; ++q)
                speed += .5 * length * t->qw[q] * fabs(t->vny[((size_t)q * (ny + 1) + j) * nx + i]);
            double dl = 0, dr = 0;
            if (t->diffuse) {
                double f;
                mpo_point ce[2];
                if (j > 0 && j < ny) {
                    mpo_point c4[4];
                    mpo_cell_corners(&t->mesh, i, j - 1, c4);
                    ce[0] = c4[2]; ce[1] = c4[3];
                    CALL(face_diffusion(t, ce, i, j - 1, i, j, dummy_u, 0, 0, 0, &f, &dl, &dr));
                } else {
                    mpo_point c4[4];
                    mpo_cell_corners(&t->mesh, i, cj, c4);
                    if (j == ny) { ce[0] = c4[2]; ce[1] = c4[3]; } else { ce[0] = c4[0]; ce[1] = c4[1]; }
                    CALL(face_diffusion(t, ce, i, cj, -1, -1, dummy_u, j == ny ? 2 : 0, i, j == ny,
                                        &f, j == 0 ? &dr : &dl, NULL));
                }
            }
            if (j > 0) r[cidx(t, i, j - 1)] += (speed + dl) / t->area[cidx(t, i, j - 1)];
            if (j < ny) r[cidx(t, i, j)] += (speed + dr) / t->area[cidx(t, i, j)];
        }
    double mx = 0;
    for (size_t c = 0; c < N; ++c) if (r[c] > mx) mx = r[c];
    if (rate) memcpy(rate, r, N * sizeof(double));
    free(r);
    free(dummy_u);
    if (max_rate) *max_rate = mx;
    return MPO_OK;
}

/* Neumaier compensated sum of area*u, fixed order. */
double mpo_transport_integral(const mpo_transport* t, const double* u)
{
    const size_t N = (size_t)t->nx * t->ny;
    double s = 0, c = 0;
    for (size_t k = 0; k < N; ++k) {
        const double x = t->area[k] * u[k];
        const double y = s + x;
        c += fabs(s) >= fabs(x) ? (s - y) + x : (x - y) + s;
        s = y;
    }
    return s + c;
}

/* Stage combination u_out = a*u_n + b*(u_k + dt*L).  Forward Euler and SSPRK2
 * follow src/couple/preheat.cpp:267 and src/couple/evolve.cpp:469-479; SSPRK3
 * (Shu-Osher) has no counterpart in the reference, which steps with SSPRK2. */
static void combine(size_t N, double a, const double* un, double b, const double* uk,
                    double dt, const double* L, double* out)
{
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < N; ++k) {
        const double w = uk[k] + dt * L[k];
        out[k] = a == 0.0 ? b * w : a * un[k] + b * w;
    }
}

int mpo_transport_advance(mpo_transport* t, double* H, double* C, int nsteps, double dt,
                          int scheme, double* totals)
{
    CHECK(scheme >= 1 && scheme <= 3, MPO_ERR_ARG, "scheme must be 1, 2 or 3");
    CHECK(finite_d(dt) && dt > 0, MPO_ERR_ARG, "dt must be finite and positive");
    const size_t N = (size_t)t->nx * t->ny;
    double* buf = (double*)malloc(N * 8 * sizeof(double));
    if (!buf) return fail(MPO_ERR_MEM, "out of memory");
    double *H1 = buf, *C1 = buf + N, *H2 = buf + 2 * N, *C2 = buf + 3 * N;
    double *LH = buf + 4 * N, *LC = buf + 5 * N, *Hn = buf + 6 * N, *Cn = buf + 7 * N;
    static const double A[3][3] = {{0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
    static const double B[3][3] = {{1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};
    static const double W[3][3] = {{1, 0, 0}, {0.5, 0.5, 0}, {1.0 / 6, 1.0 / 6, 2.0 / 3}};
    int e = 0;
    set_threads(t);
    for (int s = 0; s < nsteps && !e; ++s) {
        memcpy(Hn, H, N * sizeof(double)); memcpy(Cn, C, N * sizeof(double));
        const double* hk = Hn; const double* ck = Cn;
        double bh_total = 0, bc_total = 0;
        for (int st = 0; st < scheme && !e; ++st) {
            double bh, bc;
            e = mpo_transport_rhs(t, hk, ck, LH, LC, &bh, &bc);
            if (e) break;
            double* ho = st == scheme - 1 ? H : (st % 2 == 0 ? H1 : H2);
            double* co = st == scheme - 1 ? C : (st % 2 == 0 ? C1 : C2);
            combine(N, A[scheme - 1][st], Hn, B[scheme - 1][st], hk, dt, LH, ho);
            combine(N, A[scheme - 1][st], Cn, B[scheme - 1][st], ck, dt, LC, co);
            bh_total += W[scheme - 1][st] * bh; bc_total += W[scheme - 1][st] * bc;
            hk = ho; ck = co;
        }
        if (!e && totals) {
            totals[4 * s + 0] = mpo_transport_integral(t, H);
            totals[4 * s + 1] = mpo_transport_integral(t, C);
            totals[4 * s + 2] = bh_total;
            totals[4 * s + 3] = bc_total;
        }
    }
    free(buf);
    return e;
}

/* ================================================== coupled transport law */

/* Separate advection policies for C (c.boundary.composition); NULL side: the H
 * policy of that side. */
int mpo_transport_set_composition_boundary(mpo_transport* t, const uint8_t* bottom,
                                           const uint8_t* right, const uint8_t* top,
                                           const uint8_t* left)
{
    const uint8_t* src[4] = {bottom, right, top, left};
    for (int s = 0; s < 4; ++s) {
        free(t->bcC[s]);
        t->bcC[s] = NULL;
        if (!src[s]) continue;
        const int n = side_len(t, s);
        t->bcC[s] = (uint8_t*)malloc((size_t)n);
        if (!t->bcC[s]) return fail(MPO_ERR_MEM, "out of memory");
        memcpy(t->bcC[s], src[s], (size_t)n);
    }
    return MPO_OK;
}

/* PhaseTransportVelocities, src/couple/evolve.cpp:36-52 */
static void transport_velocities(mpo_point stokes, mpo_point darcy, double phi, double theta,
                                 const mpo_phase_state* a, const mpo_phase_state* b,
                                 mpo_point* mixture, mpo_point* effective);

void mpo_phase_transport_velocities(const double solid[2], const double relative[2], double phi,
                                    double theta, double cl_a, double cs_a, double cl_b,
                                    double cs_b, double mixture[2], double effective[2])
{
    mpo_phase_state a, b;
    memset(&a, 0, sizeof a);
    memset(&b, 0, sizeof b);
    a.cl = cl_a; a.cs = cs_a; b.cl = cl_b; b.cs = cs_b;
    const mpo_point st = {{solid[0], solid[1]}}, dr = {{relative[0], relative[1]}};
    mpo_point m, e;
    transport_velocities(st, dr, phi, theta, &a, &b, &m, &e);
    mixture[0] = m.p[0]; mixture[1] = m.p[1];
    effective[0] = e.p[0]; effective[1] = e.p[1];
}

static void transport_velocities(mpo_point stokes, mpo_point darcy, double phi, double theta,
                                 const mpo_phase_state* a, const mpo_phase_state* b,
                                 mpo_point* mixture, mpo_point* effective)
{
    const double cl = .5 * (a->cl + b->cl), cs = .5 * (a->cs + b->cs);
    const double bulk = phi * cl + (1 - phi) * cs;
    const double factor = phi == 0 ? 0 : pow(phi, 1 + theta);
    for (int k = 0; k < 2; ++k) {
        const double flux = factor * darcy.p[k];
        mixture->p[k] = stokes.p[k] + flux;
        effective->p[k] = stokes.p[k] + (bulk > 0 ? cl / bulk * flux : 0);
    }
}

typedef struct {
    const mpo_transport* t;
    const mpo_phase_model* pm;
    const mpo_coupled_params* cp;
    mpo_recon* rh;                /* [cell] limited reconstructions of H and C */
    mpo_recon* rc;
    const double *H, *C;
    double* thH;                  /* [cell] limiter factors */
    double* thC;
} coupled_ctx;

/* c.pressure(p, t): constant, or lithostatic staticPressure(surface_y - y)
 * with a finite result (src/couple/initialization.cpp:148-163) */
static int coupled_pressure(const coupled_ctx* x, mpo_point p, double* P)
{
    if (!x->cp->lithostatic) {
        *P = x->cp->pressure;
        return MPO_OK;
    }
    CALL(mpo_static_pressure(x->pm, x->cp->surface_y - p.p[1], P));
    CHECK(finite_d(*P), MPO_ERR_FP, "static pressure is not finite");
    return MPO_OK;
}

static int coupled_phase_at(const coupled_ctx* x, mpo_point p, double H, double C,
                            mpo_phase_state* s)
{
    double P;
    CALL(coupled_pressure(x, p, &P));
    return x->cp->hypereutectic ? mpo_phase_evaluate_ext(x->pm, H, C, P, s)
                                : mpo_phase_evaluate(x->pm, H, C, P, s);
}

/* Phase of the limited reconstruction at p, src/couple/evolve.cpp:270-277 with
 * StateReconstruction::Evaluate, src/couple/state_reconstruction.cpp:127-134 */
static int coupled_phase(const coupled_ctx* x, int i, int j, mpo_point p, mpo_phase_state* s,
                         double* Hp, double* Cp)
{
    const size_t c = cidx(x->t, i, j);
    double h, cc;
    CALL(mpo_recon_eval(&x->rh[c], p, 0, 0, &h));
    CALL(mpo_recon_eval(&x->rc[c], p, 0, 0, &cc));
    h = x->H[c] + x->thH[c] * (h - x->H[c]);
    cc = x->C[c] + x->thC[c] * (cc - x->C[c]);
    CALL(coupled_phase_at(x, p, h, cc, s));
    CHECK(s->TDp >= 0 && s->phil < 1, MPO_ERR_ARG,
          "Coupled Darcy-Stokes requires nonnegative temperature and porosity < 1");
    if (Hp) *Hp = h;
    if (Cp) *Cp = cc;
    return MPO_OK;
}

static size_t edge_of(const mpo_transport* t, int vertical, int i, int j)
{
    return vertical ? (size_t)t->nx * (t->ny + 1) + (size_t)j * (t->nx + 1) + i
                    : (size_t)j * t->nx + i;
}

/* One face of the coupled RHS, src/couple/evolve.cpp:291-354.  edge runs CCW
 * for cell (li, lj); (ri, rj) = (-1, -1) on the boundary (side, face k, Gauss
 * order reversed against the canonical one when rev).  eid: reference edge id. */
static int coupled_face(const coupled_ctx* x, const mpo_flow* flow, const mpo_point edge[2],
                        int li, int lj, int ri, int rj, int side, int k, int rev, size_t eid,
                        double out[4])
{
    const mpo_transport* t = x->t;
    const mpo_coupled_params* cp = x->cp;
    const int nq = t->q, boundary = ri < 0;
    double length; mpo_point normal;
    CALL(mpo_edge_geometry(edge, &length, &normal));
    double fluxH = 0, fluxC = 0, fluxMix = 0, rate = 0;
    for (int q = 0; q < nq; ++q) {
        const int qc = rev ? nq - 1 - q : q;
        const size_t a_ = eid * nq + qc;
        const mpo_point stokes = {{flow->edge_stokes[2 * a_], flow->edge_stokes[2 * a_ + 1]}};
        const mpo_point darcy = {{flow->edge_darcy[2 * a_], flow->edge_darcy[2 * a_ + 1]}};
        const double ephi = flow->edge_phi[a_];
        const mpo_point p = mpo_map_edge_point(t->qp[q], edge);
        mpo_phase_state a, b;
        double hL, cL, hR, cR;
        CALL(coupled_phase(x, li, lj, p, &a, &hL, &cL));
        if (boundary) { b = a; hR = hL; cR = cL; }
        else CALL(coupled_phase(x, ri, rj, p, &b, &hR, &cR));
        (void)hR;
        mpo_point mix, eff;
        transport_velocities(stokes, darcy, ephi, cp->theta, &a, &b, &mix, &eff);
        const double vn = mix.p[0] * normal.p[0] + mix.p[1] * normal.p[1];
        const double ve = eff.p[0] * normal.p[0] + eff.p[1] * normal.p[1];
        const double vs = stokes.p[0] * normal.p[0] + stokes.p[1] * normal.p[1];
        const double alphaH = cp->lf_mode ? cp->global_speed : fabs(vn);
        const double alphaC = cp->lf_mode ? cp->global_speed : fabs(ve);
        CHECK(alphaH + 1e-14 >= fabs(vn) && alphaC + 1e-14 >= fabs(ve), MPO_ERR_ARG,
              "Global LF speed is below the stage flow speed");
        double fh = .5 * (vn * (a.TDp + b.TDp) - alphaH * (b.TDp - a.TDp)) - cp->LD * (1 - ephi) * vs;
        double fc = .5 * (ve * (cL + cR) - alphaC * (cR - cL));
        if (boundary) {
            /* EvaluateAdvectionBoundary, src/couple/initialization.cpp:469-489 */
            const int polH = t->bc[side][k];
            const int polC = t->bcC[side] ? t->bcC[side][k] : polH;
            int typeH, typeC;       /* 0 prescribed, 1 extrapolated, 2 zero flux */
            double valH = 0, valC = 0;
            for (int f = 0; f < 2; ++f) {
                const int pol = f ? polC : polH;
                const double speed = f ? ve : vn;
                int* type = f ? &typeC : &typeH;
                double* val = f ? &valC : &valH;
                CHECK(isfinite(speed), MPO_ERR_ARG, "normal speed must be finite");
                if (pol == MPO_BC_ZERO || speed == 0) *type = 2;
                else if (speed < 0) {
                    CHECK(pol == MPO_BC_INFLOW_OUTFLOW, MPO_ERR_BACKFLOW,
                          "Backflow encountered on an outflow-only boundary.");
                    const double* src = t->inflow[f][side];
                    CHECK(src != NULL, MPO_ERR_ARG, "Missing inflow profile.");
                    *type = 0;
                    *val = src[(size_t)k * nq + qc];
                } else *type = 1;
            }
            const double inflowC = typeC == 0 ? valC : cL;
            double T = a.TDp;
            if (typeH == 0) {
                mpo_phase_state sb;
                CALL(coupled_phase_at(x, p, valH, inflowC, &sb));
                T = sb.TDp;
            }
            fh = typeH == 2 ? 0 : vn * T - cp->LD * (1 - ephi) * vs;
            fc = typeC == 2 ? 0 : ve * (typeC == 0 ? valC : cL);
        }
        const double weight = .5 * length * t->qw[q];
        fluxH += weight * fh;
        fluxC += weight * fc;
        fluxMix += weight * vn;
        const double sensitivity = 1 + fmax(fabs(a.dTD_dCD), fabs(b.dTD_dCD));
        CHECK(isfinite(sensitivity), MPO_ERR_FP, "Singular phase derivative in CFL estimate");
        rate += weight * (alphaC + alphaH + 2 * fabs(vs)) * sensitivity;
    }
    /* normal-line diffusion of T, src/couple/evolve.cpp:334-347 */
    mpo_point lc[4], rc[4];
    mpo_cell_corners(&t->mesh, li, lj, lc);
    if (!boundary) mpo_cell_corners(&t->mesh, ri, rj, rc);
    mpo_sampling smp;
    CALL(mpo_diffusive_sampling(edge, lc, boundary ? NULL : rc, nq, t->qp, t->qw, cp->samples,
                                cp->extent, &smp));
    double samples[16 * (MPO_MAX_SAMPLES + 1)], kappa[16], bv[16];
    int bt[16];
    for (int a = 0; a < smp.nq * smp.count; ++a) {
        samples[a] = 0;
        if (smp.sides[a] == MPO_SIDE_BOUNDARY) continue;
        mpo_phase_state ps;
        if (smp.sides[a] == MPO_SIDE_LEFT) CALL(coupled_phase(x, li, lj, smp.samples[a], &ps, NULL, NULL));
        else CALL(coupled_phase(x, ri, rj, smp.samples[a], &ps, NULL, NULL));
        samples[a] = ps.TDp;
    }
    for (int q = 0; q < nq; ++q) {
        kappa[q] = cp->kappa;
        if (boundary) {
            const int qc = rev ? nq - 1 - q : q;
            const int pol = t->tbc[side] ? t->tbc[side][(size_t)k * nq + qc] : 0;
            bt[q] = pol == 1 ? MPO_DIFF_QUANTITY : (pol == 2 ? MPO_DIFF_FLUX : MPO_DIFF_ZERO_FLUX);
            bv[q] = t->tval[side] ? t->tval[side][(size_t)k * nq + qc] : 0.0;
        }
    }
    double dflux;
    CALL(mpo_diffusive_flux(&smp, samples, kappa, boundary ? bt : NULL, boundary ? bv : NULL, &dflux));
    fluxH += dflux;
    /* |sample partials| (identity quantity), src/couple/evolve.cpp:348-350 */
    long double last = 0;
    for (int kk = 0; kk + 1 < smp.count; ++kk) last -= smp.coefficients[kk];
    for (int q = 0; q < nq; ++q) {
        if (boundary && bt[q] != MPO_DIFF_QUANTITY) continue;
        const long double scale = -(long double)smp.weights[q] * kappa[q] / smp.spacing[q];
        for (int kk = 0; kk < smp.count; ++kk) {
            const int a = q * smp.count + kk;
            if (smp.sides[a] == MPO_SIDE_BOUNDARY) continue;
            const long double coeff = kk + 1 == smp.count ? last : smp.coefficients[kk];
            double d;
            CALL(real_value(scale * coeff, &d));
            rate += fabs(d);
        }
    }
    out[0] = fluxH; out[1] = fluxC; out[2] = fluxMix; out[3] = rate;
    return MPO_OK;
}

int mpo_coupled_rhs(mpo_transport* t, const mpo_phase_model* pm, const mpo_coupled_params* cp,
                    const mpo_flow* flow, const double* H, const double* C, double* rhsH,
                    double* rhsC, double* rate, double* div, double* T, double* phi,
                    mpo_coupled_rates* totals)
{
    CHECK(t->limit, MPO_ERR_STATE, "the coupled law needs the bound limiter (set_limiter)");
    const int nx = t->nx, ny = t->ny;
    const size_t N = (size_t)nx * ny;
    if (cp->lithostatic) {
        /* src/couple/initialization.cpp:155-158 */
        CHECK(finite_d(cp->surface_y), MPO_ERR_ARG, "surface_y must be finite");
        for (int i = 0; i < nx; ++i) {
            mpo_point c4[4];
            mpo_cell_corners(&t->mesh, i, ny - 1, c4);
            CHECK(cp->surface_y >= c4[2].p[1] && cp->surface_y >= c4[3].p[1], MPO_ERR_ARG,
                  "surface_y must be at or above the mesh top");
        }
    }
    coupled_ctx x = {t, pm, cp, NULL, NULL, H, C, NULL, NULL};
    x.rh = (mpo_recon*)malloc(N * sizeof(mpo_recon));
    x.rc = (mpo_recon*)malloc(N * sizeof(mpo_recon));
    x.thH = (double*)malloc(N * sizeof(double));
    x.thC = (double*)malloc(N * sizeof(double));
    double* acc = (double*)calloc(N * 6, sizeof(double));  /* rhsH rhsC rate div T phi */
    int e = 0;
    if (!x.rh || !x.rc || !x.thH || !x.thC || !acc) e = fail(MPO_ERR_MEM, "out of memory");
    /* Update + Limit, src/couple/state_reconstruction.cpp:101-192 */
    for (size_t c = 0; c < N && !e; ++c) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        e = cell_recon(t, i, j, &x.rh[c]);
        if (!e) e = mpo_recon_update(&x.rh[c], H, 0, 0, nx, ny, t->area[c]);
        if (!e) e = cell_recon(t, i, j, &x.rc[c]);
        if (!e) e = mpo_recon_update(&x.rc[c], C, 0, 0, nx, ny, t->area[c]);
        if (!e) e = cell_limit(t, &x.rh[c], H, i, j, INFINITY, &x.thH[c]);
        if (!e) e = cell_limit(t, &x.rc[c], C, i, j, t->lim_cupper, &x.thC[c]);
    }
    double bH = 0, bC = 0, sH = 0;
    /* faces: vertical x_i (left cell (i-1, j)), then horizontal y_j (left (i, j-1)) */
    for (int j = 0; j < ny && !e; ++j)
        for (int i = 0; i <= nx && !e; ++i) {
            double o[4];
            mpo_point c4[4], edge[2];
            const size_t eid = edge_of(t, 1, i, j);
            if (i > 0 && i < nx) {
                mpo_cell_corners(&t->mesh, i - 1, j, c4);
                edge[0] = c4[1]; edge[1] = c4[2];
                e = coupled_face(&x, flow, edge, i - 1, j, i, j, 0, 0, 0, eid, o);
                if (e) break;
                const size_t L = cidx(t, i - 1, j), R = cidx(t, i, j);
                acc[L * 6] -= o[0] / t->area[L]; acc[R * 6] += o[0] / t->area[R];
                acc[L * 6 + 1] -= o[1] / t->area[L]; acc[R * 6 + 1] += o[1] / t->area[R];
                acc[L * 6 + 2] += o[3] / t->area[L]; acc[R * 6 + 2] += o[3] / t->area[R];
                acc[L * 6 + 3] += o[2] / t->area[L]; acc[R * 6 + 3] -= o[2] / t->area[R];
            } else {
                const int ci = i == 0 ? 0 : nx - 1, side = i == 0 ? 3 : 1;
                mpo_cell_corners(&t->mesh, ci, j, c4);
                if (side == 1) { edge[0] = c4[1]; edge[1] = c4[2]; }
                else { edge[0] = c4[3]; edge[1] = c4[0]; }
                e = coupled_face(&x, flow, edge, ci, j, -1, -1, side, j, side == 3, eid, o);
                if (e) break;
                const size_t L = cidx(t, ci, j);
                acc[L * 6] -= o[0] / t->area[L]; acc[L * 6 + 1] -= o[1] / t->area[L];
                acc[L * 6 + 2] += o[3] / t->area[L]; acc[L * 6 + 3] += o[2] / t->area[L];
                bH += o[0]; bC += o[1];
            }
        }
    for (int j = 0; j <= ny && !e; ++j)
        for (int i = 0; i < nx && !e; ++i) {
            double o[4];
            mpo_point c4[4], edge[2];
            const size_t eid = edge_of(t, 0, i, j);
            if (j > 0 && j < ny) {
                /* the reference's edge runs from (i, j) to (i+1, j) with the
                 * lower cell on its right; orient it CCW for the lower cell */
                mpo_cell_corners(&t->mesh, i, j - 1, c4);
                edge[0] = c4[2]; edge[1] = c4[3];
                e = coupled_face(&x, flow, edge, i, j - 1, i, j, 0, 0, 1, eid, o);
                if (e) break;
                const size_t L = cidx(t, i, j - 1), R = cidx(t, i, j);
                acc[L * 6] -= o[0] / t->area[L]; acc[R * 6] += o[0] / t->area[R];
                acc[L * 6 + 1] -= o[1] / t->area[L]; acc[R * 6 + 1] += o[1] / t->area[R];
                acc[L * 6 + 2] += o[3] / t->area[L]; acc[R * 6 + 2] += o[3] / t->area[R];
                acc[L * 6 + 3] += o[2] / t->area[L]; acc[R * 6 + 3] -= o[2] / t->area[R];
            } else {
                const int cj = j == 0 ? 0 : ny - 1, side = j == 0 ? 0 : 2;
                mpo_cell_corners(&t->mesh, i, cj, c4);
                if (side == 0) { edge[0] = c4[0]; edge[1] = c4[1]; }
                else { edge[0] = c4[2]; edge[1] = c4[3]; }
                e = coupled_face(&x, flow, edge, i, cj, -1, -1, side, i, side == 2, eid, o);
                if (e) break;
                const size_t L = cidx(t, i, cj);
                acc[L * 6] -= o[0] / t->area[L]; acc[L * 6 + 1] -= o[1] / t->area[L];
                acc[L * 6 + 2] += o[3] / t->area[L]; acc[L * 6 + 3] += o[2] / t->area[L];
                bH += o[0]; bC += o[1];
            }
        }
    /* cell quadrature: cooling source and phase diagnostics, src/couple/evolve.cpp:357-370 */
    rule1d cell;
    if (!e) e = make_rule(cp->cell_points, &cell);
    const int np = cell.n * cell.n;
    for (size_t c = 0; c < N && !e; ++c) {
        const int i = (int)(c % nx), j = (int)(c / nx);
        mpo_point c4[4];
        mpo_cell_corners(&t->mesh, i, j, c4);
        for (int b = 0; b < cell.n && !e; ++b)
            for (int a = 0; a < cell.n && !e; ++a) {
                const size_t k = c * np + (size_t)b * cell.n + a;
                const mpo_point ref = {{cell.x[a], cell.x[b]}};
                const mpo_point p = mpo_map_cell_point(ref, c4);
                const double weight = cell.w[a] * cell.w[b] * mpo_cell_jacobian(ref, c4);
                mpo_phase_state ps;
                e = coupled_phase(&x, i, j, p, &ps, NULL, NULL);
                if (e) break;
                const mpo_point stokes = {{flow->cell_stokes[2 * k], flow->cell_stokes[2 * k + 1]}};
                const mpo_point darcy = {{flow->cell_darcy[2 * k], flow->cell_darcy[2 * k + 1]}};
                mpo_point mix, eff;
                transport_velocities(stokes, darcy, flow->cell_phi[k], cp->theta, &ps, &ps, &mix, &eff);
                const double source = -cp->cooling * ps.TDp * mix.p[1];
                const double wa = weight / t->area[c];
                acc[c * 6] += wa * source;
                acc[c * 6 + 4] += wa * ps.TDp;
                acc[c * 6 + 5] += wa * ps.phil;
                acc[c * 6 + 2] += wa * cp->cooling * fabs(mix.p[1]);
                sH += weight * source;
            }
    }
    double mr = 0, md = 0;
    for (size_t c = 0; c < N && !e; ++c) {
        mr = fmax(mr, acc[c * 6 + 2]);
        md = fmax(md, fabs(acc[c * 6 + 3]));
    }
    if (!e && !isfinite(mr)) e = fail(MPO_ERR_FP, "Nonfinite coupled stage rate");
    if (!e) {
        for (size_t c = 0; c < N; ++c) {
            if (rhsH) rhsH[c] = acc[c * 6];
            if (rhsC) rhsC[c] = acc[c * 6 + 1];
            if (rate) rate[c] = acc[c * 6 + 2];
            if (div) div[c] = acc[c * 6 + 3];
            if (T) T[c] = acc[c * 6 + 4];
            if (phi) phi[c] = acc[c * 6 + 5];
        }
        totals->boundaryH = bH; totals->boundaryC = bC; totals->sourceH = sH;
        totals->maximumRate = mr; totals->divergence = md;
    }
    free(x.rh); free(x.rc); free(x.thH); free(x.thC); free(acc);
    return e;
}

/* Admissible, src/couple/evolve.cpp:386-393 */
static int coupled_admissible(size_t N, const double* H, const double* C, double Xe)
{
    double hmin = INFINITY, hmax = -INFINITY, cmin = INFINITY, cmax = -INFINITY;
    for (size_t k = 0; k < N; ++k) {
        hmin = fmin(hmin, H[k]); hmax = fmax(hmax, H[k]);
        cmin = fmin(cmin, C[k]); cmax = fmax(cmax, C[k]);
        if (!isfinite(H[k]) || !isfinite(C[k])) return 0;
    }
    return isfinite(hmin) && isfinite(hmax) && isfinite(cmin) && isfinite(cmax) && hmin >= 0 &&
           cmin >= 0 && cmax <= Xe;
}

static double coupled_integral(const mpo_transport* t, const double* u)
{
    long double s = 0;
    for (size_t k = 0; k < (size_t)t->nx * t->ny; ++k) s += (long double)u[k] * t->area[k];
    return (double)s;
}

/* Run, src/couple/evolve.cpp:428-511: SSPRK2 steps with admissibility and CFL
 * retries (dt halved, at most 40) and the conservation balance check; the flow
 * samples stay fixed (the flow solve of every RHS is outside this path).
 * H, C advance in place over nsteps accepted steps from time *time. */
int mpo_coupled_run(mpo_transport* t, const mpo_phase_model* pm, const mpo_coupled_params* cp,
                    const mpo_flow* flow, const mpo_coupled_time* tc, double* H, double* C,
                    double* time, int nsteps, mpo_coupled_step* steps)
{
    const size_t N = (size_t)t->nx * t->ny;
    CHECK(tc->cfl > 0 && tc->cfl <= 1, MPO_ERR_ARG, "Coupled CFL must be in (0,1]");
    double* buf = (double*)malloc(N * 8 * sizeof(double));
    if (!buf) return fail(MPO_ERR_MEM, "out of memory");
    double *oldH = buf, *oldC = buf + N, *rH = buf + 2 * N, *rC = buf + 3 * N;
    double *fH = buf + 4 * N, *fC = buf + 5 * N;
    mpo_coupled_rates current, second;
    int e = mpo_coupled_rhs(t, pm, cp, flow, H, C, rH, rC, NULL, NULL, NULL, NULL, &current);
    memcpy(oldH, H, N * sizeof(double)); memcpy(oldC, C, N * sizeof(double));
    const double Xe = cp->admissible_c_max;
    int accepted = 0;
    for (int s = 0; s < nsteps && !e; ++s) {
        const double remaining = tc->end - *time;
        if (remaining <= 32 * DBL_EPSILON * fmax(1.0, fabs(tc->end))) { *time = tc->end; break; }
        const double oldHi = coupled_integral(t, H), oldCi = coupled_integral(t, C);
        const double start = *time;
        double dt = fmin(remaining, tc->fixed ? tc->initial : (accepted == 0 ? tc->initial : tc->maximum));
        const double limit = current.maximumRate > 0 ? tc->cfl / current.maximumRate : tc->maximum;
        if (tc->fixed && !(dt <= limit * (1 + 1e-12))) {
            e = fail(MPO_ERR_ARG, "Fixed time step exceeds the phase-coupled CFL bound; use control: cfl");
            break;
        }
        if (!tc->fixed) dt = fmin(dt, limit);
        memcpy(fH, rH, N * sizeof(double)); memcpy(fC, rC, N * sizeof(double));
        const mpo_coupled_rates first = current;
        int retries = 0;
        double courant = 0;
        for (;;) {
            if (!(dt >= fmin(remaining, tc->minimum) && retries < 40)) {
                e = fail(MPO_ERR_RUNTIME, "Coupled step cannot satisfy admissibility/CFL above minimum step");
                break;
            }
            for (size_t k = 0; k < N; ++k) { H[k] = oldH[k] + dt * fH[k]; C[k] = oldC[k] + dt * fC[k]; }
            *time = start + dt;
            int valid = coupled_admissible(N, H, C, Xe);
            if (valid) {
                e = mpo_coupled_rhs(t, pm, cp, flow, H, C, rH, rC, NULL, NULL, NULL, NULL, &second);
                if (e) break;
                courant = dt * fmax(current.maximumRate, second.maximumRate);
                valid = courant <= tc->cfl * (1 + 1e-12);
            }
            if (valid) {
                for (size_t k = 0; k < N; ++k) {
                    H[k] = .5 * (H[k] + dt * rH[k]) + .5 * oldH[k];
                    C[k] = .5 * (C[k] + dt * rC[k]) + .5 * oldC[k];
                }
                valid = coupled_admissible(N, H, C, Xe);
            }
            if (valid) break;
            if (tc->fixed) { e = fail(MPO_ERR_ARG, "Fixed coupled step failed stage admissibility/CFL"); break; }
            dt *= .5;
            ++retries;
        }
        double balH = 0, balC = 0;
        if (!e) {
            const double Hi = coupled_integral(t, H), Ci = coupled_integral(t, C);
            balH = Hi - oldHi + .5 * dt * (current.boundaryH + second.boundaryH - current.sourceH - second.sourceH);
            balC = Ci - oldCi + .5 * dt * (current.boundaryC + second.boundaryC);
            if (!(fabs(balH) < 1e-10 * fmax(1.0, fabs(oldHi)) && fabs(balC) < 1e-10))
                e = fail(MPO_ERR_RUNTIME, "Coupled finite-volume conservation balance failed");
        }
        if (!e) e = mpo_coupled_rhs(t, pm, cp, flow, H, C, rH, rC, NULL, NULL, NULL, NULL, &current);
        if (e) {
            *time = start;
            memcpy(H, oldH, N * sizeof(double)); memcpy(C, oldC, N * sizeof(double));
            break;
        }
        ++accepted;
        memcpy(oldH, H, N * sizeof(double)); memcpy(oldC, C, N * sizeof(double));
        if (steps) {
            mpo_coupled_step* r = &steps[s];
            r->dt = dt; r->courant = courant; r->retries = retries;
            r->balanceH = balH; r->balanceC = balC;
            r->boundaryH = .5 * (first.boundaryH + second.boundaryH);
            r->boundaryC = .5 * (first.boundaryC + second.boundaryC);
            r->sourceH = .5 * (first.sourceH + second.sourceH);
            r->maximumRate = current.maximumRate;
        }
    }
    free(buf);
    return e;
}

uint64_t mpo_mix64(uint64_t v)
{
    v += UINT64_C(0x9e3779b97f4a7c15);
    v = (v ^ (v >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
    v = (v ^ (v >> 27)) * UINT64_C(0x94d049bb133111eb);
    return v ^ (v >> 31);
}
