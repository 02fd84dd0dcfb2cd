// Host construction of per-mesh device records: Gauss rule, derived phase
// scales and the closed-form reconstruction tables.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "mp_host.h"

namespace mp {

static thread_local char g_error[512];

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
}

mp_status fail(mp_status code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    return code;
}

const char* last_error() { return g_error; }

mp_status status_from_bits(uint32_t bits)
{
    // a stalled one-launch step left items unwritten: nothing else is reliable
    if (bits & kErrStall) return MP_ERR_DEVICE;
    if (bits & kErrBackflow) return MP_ERR_BACKFLOW;
    // the limiter runs before any phase evaluation of the stage
    if (bits & kErrInadmissible) return MP_ERR_ARG;
    if (bits & (kErrCoupledRange | kErrLFSpeed)) return MP_ERR_ARG;
    if (bits & kErrComposition) return MP_ERR_COMPOSITION;
    if (bits & kErrFraction) return MP_ERR_PHASE;
    if (bits & kErrInflow) return MP_ERR_ARG;
    if (bits & (kErrNonFinite | kErrWeights | kErrPhaseRange)) return MP_ERR_FP;
    // last: a non-finite state fails the balance too, and is the cause
    if (bits & kErrBalance) return MP_ERR_BALANCE;
    return MP_OK;
}

const char* describe_bits(uint32_t bits)
{
    if (bits & kErrStall) return "A work item's inputs were not completed in time; the step was abandoned";
    if (bits & kErrBackflow) return "Backflow encountered on an outflow-only boundary";
    if (bits & kErrInadmissible) return "Inadmissible cell average in coupled reconstruction";
    if (bits & kErrCoupledRange) return "Coupled Darcy-Stokes requires nonnegative temperature and porosity < 1";
    if (bits & kErrLFSpeed) return "Global LF speed is below the stage flow speed";
    if (bits & kErrComposition) return "CD must lie in [0, Xe]";
    if (bits & kErrFraction) return "Computed phase fraction is outside [0, 1]";
    if (bits & kErrInflow) return "Missing inflow profile";
    if (bits & kErrNonFinite) return "Non-finite state, trace or flux";
    if (bits & kErrWeights) return "No finite active log weight";
    if (bits & kErrPhaseRange) return "Phase temperature/enthalpy intervals are not resolvable";
    if (bits & kErrBalance) return "Preheat heat balance failed; candidate rejected";
    return "ok";
}

void gauss_legendre(int n, double* x, double* w)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int k = 0; k < n; ++k) {
        long double z = cosl(pi * (k + 0.75L) / (n + 0.5L)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1, p1 = z;
            for (int m = 2; m <= n; ++m) {
                const long double p2 = ((2 * m - 1) * z * p1 - (m - 1) * p0) / m;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p0 = 1; p1 = z; }
            dp = n * (z * p1 - p0) / (z * z - 1);
            const long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        long double p0 = 1, p1 = z;
        for (int m = 2; m <= n; ++m) {
            const long double p2 = ((2 * m - 1) * z * p1 - (m - 1) * p0) / m;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1);
        x[n - 1 - k] = double(z);
        w[n - 1 - k] = double(2.0L / ((1 - z * z) * dp * dp));
    }
    if (n % 2) x[n / 2] = 0.0;
    for (int k = 0; k < n / 2; ++k) {
        const double a = 0.5 * (x[n - 1 - k] - x[k]);
        x[k] = -a;
        x[n - 1 - k] = a;
        const double ww = 0.5 * (w[k] + w[n - 1 - k]);
        w[k] = ww;
        w[n - 1 - k] = ww;
    }
}

static bool finite_ld(long double v) { return std::isfinite(v) && fabsl(v) <= (long double)DBL_MAX; }

mp_status make_phase(const mp_phase_params& p, PhaseHost& out)
{
    // derive(), src/phase/eutectic_rescaled.cpp:42-86 (validation and scales
    // the transport path uses).
    auto pos = [](double v) { return std::isfinite(v) && v > 0.0; };
    if (!pos(p.Tm0)) return fail(MP_ERR_ARG, "Tm0 must be positive");
    if (!pos(p.Te0)) return fail(MP_ERR_ARG, "Te0 must be positive");
    if (p.Tm0 <= p.Te0) return fail(MP_ERR_ARG, "Tm0 must be greater than Te0");
    if (!std::isfinite(p.nu) || p.nu == 0.0) return fail(MP_ERR_ARG, "nu must be finite and nonzero");
    if (!pos(p.L)) return fail(MP_ERR_ARG, "L must be positive");
    if (!pos(p.cp)) return fail(MP_ERR_ARG, "cp must be positive");
    if (!std::isfinite(p.Xe) || p.Xe <= 0.0 || p.Xe >= 1.0)
        return fail(MP_ERR_ARG, "Xe must lie strictly between 0 and 1");
    if (!pos(p.rho) || !pos(p.rhor) || !pos(p.mus) || !pos(p.mul) || !pos(p.k0) || !pos(p.g))
        return fail(MP_ERR_ARG, "rho, rhor, mus, mul, k0 and g must be positive");
    if (!std::isfinite(p.alpha0)) return fail(MP_ERR_ARG, "alpha0 must be finite");
    if (!std::isfinite(p.pressure)) return fail(MP_ERR_ARG, "P must be finite");
    const long double deltaT = (long double)p.Tm0 - p.Te0;
    const long double h0 = (long double)p.cp * deltaT;
    const long double LD = p.L / h0;
    const long double TDe0 = p.Te0 / deltaT;
    if (!finite_ld(deltaT) || !finite_ld(h0) || !finite_ld(LD) || !finite_ld(TDe0) ||
        (double)LD <= 0.0 || (double)TDe0 <= 0.0 || (double)deltaT <= 0.0)
        return fail(MP_ERR_FP, "phase scales are not representable");
    const long double l0 = sqrtl((long double)p.mus * p.k0 / p.mul);
    if (!finite_ld(l0) || (double)l0 <= 0.0) return fail(MP_ERR_FP, "l0 is not a positive scale");
    PhaseHost h;
    std::memset(&h, 0, sizeof h);
    h.dT = (double)deltaT;
    h.TDe0 = (double)TDe0;
    h.nu = p.nu;
    h.rho = p.rho;
    h.g = p.g;
    h.l0 = (double)l0;
    h.dev.LD = (double)LD;
    h.dev.Xe = p.Xe;
    // boundaries(): Tep and Tmp rounded once (eutectic_rescaled.cpp:143-149, 166-169)
    const long double shift = (long double)p.pressure / p.nu / h.dT;
    const long double tep = (long double)h.TDe0 + shift;
    if (!finite_ld(tep)) return fail(MP_ERR_FP, "pressure-corrected temperature is not representable");
    h.dev.Tep = (double)tep;
    const long double tmp = (long double)h.dev.Tep + 1.0L;
    if (!finite_ld(tmp)) return fail(MP_ERR_FP, "Tmp is not representable");
    h.dev.Tmp = (double)tmp;
    if (h.dev.Tmp <= h.dev.Tep) return fail(MP_ERR_FP, "Phase temperature intervals are not resolvable");
    h.dev.hyper = p.hypereutectic ? 1 : 0;
    // 1/x as hi + lo with lo = (1 - hi*x)/x: exact to a few units of 2^-100
    auto recip = [](double x, double& hi, double& lo) {
        hi = 1.0 / x;
        lo = std::fma(-hi, x, 1.0) / x;
    };
    recip(h.dev.Xe, h.dev.ixh, h.dev.ixl);
    recip(1.0 - h.dev.Xe, h.dev.imh, h.dev.iml);
    recip(h.dev.LD, h.dev.ilh, h.dev.ill);
    out = h;
    return MP_OK;
}

mp_status lithostatic_pair(const PhaseHost& h, double surface_y, double y, TepPair& out)
{
    const double zD = surface_y - y;
    if (!std::isfinite(zD)) return fail(MP_ERR_ARG, "zD must be finite");
    const long double P = (long double)h.rho * h.g * zD * h.l0;
    if (!finite_ld(P)) return fail(MP_ERR_FP, "static pressure is not representable");
    const long double shift = (long double)(double)P / h.nu / h.dT;
    const long double tep = (long double)h.TDe0 + shift;
    if (!finite_ld(tep)) return fail(MP_ERR_FP, "pressure-corrected temperature is not representable");
    const long double tmp = (long double)(double)tep + 1.0L;
    if (!finite_ld(tmp)) return fail(MP_ERR_FP, "Tmp is not representable");
    out.tep = (double)tep;
    out.tmp = (double)tmp;
    if (!(out.tmp > out.tep)) return fail(MP_ERR_FP, "Phase temperature intervals are not resolvable");
    return MP_OK;
}

mp_status lithostatic_rows(const mp_mesh_desc& mesh, const PhaseHost& h, double surface_y,
                           std::vector<TepPair>& rows)
{
    if (!std::isfinite(surface_y)) return fail(MP_ERR_ARG, "surface_y must be finite");
    if (surface_y < mesh.y1) return fail(MP_ERR_ARG, "surface_y must be at or above the mesh top");
    std::vector<double> y(size_t(mesh.ny) + 1);
    uniform_vertices(mesh.y0, mesh.y1, mesh.ny, y.data());
    rows.assign(size_t(mesh.ny), TepPair{});
    for (int j = 0; j < mesh.ny; ++j) {
        const double lower = map_edge(0.0, y[j], y[j]), upper = map_edge(0.0, y[j + 1], y[j + 1]);
        MP_CALL(lithostatic_pair(h, surface_y, map_edge(0.0, lower, upper), rows[j]));
    }
    return MP_OK;
}

void uniform_vertices(double a0, double a1, int n, double* v)
{
    const double step = (a1 - a0) / n;
    for (int k = 0; k <= n; ++k) v[k] = a0 + k * step;
    v[0] = a0;
    v[n] = a1;
}

mp_status validate_mesh(const mp_mesh_desc& m)
{
    if (m.nx < 2 || m.ny < 2) return fail(MP_ERR_ARG, "Transport requires at least 2x2 cells");
    if ((int64_t)m.nx * (int64_t)m.ny > (int64_t)1 << 31)
        return fail(MP_ERR_ARG, "mesh exceeds 2^31 cells");
    if (m.q < 1 || m.q > kMaxQ) return fail(MP_ERR_ARG, "edge Gauss points must be 1..%d", kMaxQ);
    if (!std::isfinite(m.x0) || !std::isfinite(m.x1) || !std::isfinite(m.y0) || !std::isfinite(m.y1) ||
        !(m.x1 > m.x0) || !(m.y1 > m.y0))
        return fail(MP_ERR_ARG, "Mesh origin must be finite and lengths finite and positive");
    return MP_OK;
}

mp_status validate_weno(const mp_weno_options& o)
{
    // src/mlweno/reconstruction.cpp:123-131, 35-38
    if (!std::isfinite(o.epsilon) || o.epsilon <= 0 || o.s < 0)
        return fail(MP_ERR_ARG, "h and epsilon must be finite and positive; s must be nonnegative");
    if (!std::isfinite(o.constant_linear_weight) || o.constant_linear_weight < 0)
        return fail(MP_ERR_ARG, "Constant linear weight must be finite and nonnegative");
    if (!std::isfinite(o.large_weight) || o.large_weight < 0)
        return fail(MP_ERR_ARG, "Linear weights must be finite and nonnegative");
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(o.small_weight[k]) || o.small_weight[k] < 0)
            return fail(MP_ERR_ARG, "Linear weights must be finite and nonnegative");
    if (o.lf_mode != 0 && o.lf_mode != 1) return fail(MP_ERR_ARG, "Unknown LF mode");
    if (o.lf_mode == 1 && (!std::isfinite(o.global_speed) || o.global_speed < 0))
        return fail(MP_ERR_ARG, "Global LF speed must be finite and nonnegative");
    return MP_OK;
}

static double log_weight(double w)
{
    return w > 0 ? std::log(w) : -std::numeric_limits<double>::infinity();
}

static int eta(int r) { return r == 1 ? 1 : (r == 2 ? 3 : 4); }

mp_status make_mesh_dev(const mp_mesh_desc& mesh, const mp_weno_options& weno,
                        const PhaseHost& phase, MeshDev& out)
{
    MP_CALL(validate_mesh(mesh));
    MP_CALL(validate_weno(weno));
    MeshDev m;
    std::memset(&m, 0, sizeof m);
    m.nx = mesh.nx;
    m.ny = mesh.ny;
    m.q = mesh.q;
    m.lf_mode = weno.lf_mode;
    m.global_speed = weno.global_speed;
    m.dx = (mesh.x1 - mesh.x0) / mesh.nx;
    m.dy = (mesh.y1 - mesh.y0) / mesh.ny;
    m.area = m.dx * m.dy;
    if (!(m.area > 0) || !std::isfinite(m.area)) return fail(MP_ERR_FP, "Cell area underflowed to zero");
    m.inv_area = 1.0 / m.area;
    // h = sqrt(area) (src/couple/state_reconstruction.cpp:74)
    const long double h = sqrtl((long double)m.area);
    const double aw = (double)(m.dx / h), bw = (double)(m.dy / h);
    m.aw = aw;
    m.bw = bw;
    m.hx = 0.5 * aw;
    m.hy = 0.5 * bw;
    m.kx1 = (double)(1.0L / (2.0L * aw));
    m.kx2 = (double)(1.0L / (2.0L * aw * aw));
    m.kx0 = (double)((long double)aw * aw / 12.0L);
    m.ky1 = (double)(1.0L / (2.0L * bw));
    m.ky2 = (double)(1.0L / (2.0L * bw * bw));
    m.ky0 = (double)((long double)bw * bw / 12.0L);
    m.inv_aw = (double)(1.0L / aw);
    m.inv_bw = (double)(1.0L / bw);
    m.inv_awbw = (double)(1.0L / ((long double)aw * bw));
    m.q11w = (double)(((long double)aw * aw + (long double)bw * bw) / 12.0L);
    {
        const long double rw = sqrtl(((long double)aw * aw + (long double)bw * bw) / 12.0L);
        m.q11s = (double)(rw / ((long double)aw * bw));
        m.q11u = (double)(1.0L / rw);
    }
    m.epsA = weno.epsilon * m.area;
    if (!(m.epsA > 0)) return fail(MP_ERR_FP, "epsilon*area underflows to zero");
    // exponents s*r + eta(r): r = order + 1 (src/mlweno/reconstruction.cpp:43-44, 225-226)
    m.pL = double(weno.s * 3 + eta(3));
    m.pS = double(weno.s * 2 + eta(2));
    m.pC = double(weno.s * 1 + eta(1));
    m.ldL = log_weight(weno.large_weight);
    for (int k = 0; k < 4; ++k) m.ldS[k] = log_weight(weno.small_weight[k]);
    m.use_constant = (weno.use_constant && weno.constant_linear_weight > 0) ? 1 : 0;
    m.ldC = log_weight(weno.constant_linear_weight);
    {
        // Direct form: with linear weights in [2^-64, 2^64] and denominators in
        // [2^-B, 2^B], every alpha lies in [2^-(pL*B + 64), 2^(pL*B + 64)]
        // (pL is the largest exponent), so B = (1000 - 64) / pL keeps the
        // alphas and their sum of at most six normal and finite.
        m.dL = weno.large_weight;
        for (int k = 0; k < 4; ++k) m.dS[k] = weno.small_weight[k];
        m.ipL = weno.s <= 20 ? weno.s * 3 + eta(3) : 0;
        m.ipS = weno.s <= 20 ? weno.s * 2 + eta(2) : 0;
        const int B = m.ipL > 0 ? (1000 - 64) / m.ipL : 0;
        m.wlo = std::ldexp(1.0, -B);
        m.whi = std::ldexp(1.0, B);
        auto weight_ok = [](double d) { return d == 0 || (d >= 0x1p-64 && d <= 0x1p64); };
        bool ok = m.ipL > 0 && weight_ok(m.dL);
        for (int k = 0; k < 4; ++k) ok = ok && weight_ok(m.dS[k]);
        m.aC = 0;
        if (ok && m.use_constant) {
            const long double a = (long double)weno.constant_linear_weight *
                                  powl((long double)m.epsA, -(long double)(weno.s + eta(1)));
            ok = a >= 0x1p-1000L && a <= 0x1p1000L;
            m.aC = ok ? (double)a : 0;
        }
        m.wdirect = ok ? 1 : 0;
    }
    double xq[kMaxQ], wq[kMaxQ];
    gauss_legendre(mesh.q, xq, wq);
    for (int q = 0; q < mesh.q; ++q) {
        m.sv[q] = m.hy * xq[q];
        m.sh[q] = m.hx * xq[q];
        m.wv[q] = wq[q] * m.dy / 2;
        m.wh[q] = wq[q] * m.dx / 2;
    }
    m.ph = phase.dev;
    out = m;
    return MP_OK;
}

mp_status set_limiter_tables(MeshDev& m, int cell_points, int samples, double extent,
                             double c_upper)
{
    if (cell_points < 1 || cell_points > kMaxQ)
        return fail(MP_ERR_ARG, "cell_points_per_axis must be 1..%d", kMaxQ);
    if (samples < 2 || samples % 2 != 0 || samples > kMaxLimS)
        return fail(MP_ERR_ARG, "Diffusion requires an even sample count >= 2 (at most %d here)", kMaxLimS);
    if (!(std::isfinite(extent) && extent > 0 && extent <= 1))
        return fail(MP_ERR_ARG, "Normal sampling extentFraction must lie in (0,1]");
    double xc[kMaxQ], wc[kMaxQ];
    gauss_legendre(cell_points, xc, wc);
    m.lim_ncp = cell_points;
    for (int k = 0; k < cell_points; ++k) {
        m.lim_cx[k] = m.hx * xc[k];
        m.lim_cy[k] = m.hy * xc[k];
        m.lim_cw[k] = wc[k];
    }
    // A face's in-cell samples sit at depth d_k inward from the face along its
    // normal.  The available normal-ray distance of a uniform rectangle is its
    // width (the same on both sides of an interior face), so with n samples and
    // extent f: interior spacing f*w*2/(n-1), offsets k - (n-1)/2 for the n/2
    // in-cell ones; boundary spacing f*w/n, offsets k - n for k < n.
    const long double f = extent;
    const int n = samples;
    m.lim_ni = n / 2;
    m.lim_nb = n;
    for (int k = 0; k < n / 2; ++k) {
        const long double depth = -((long double)k - (n - 1) / 2.0L) * f * 2.0L / (n - 1);
        m.lim_xi[k] = (double)((long double)m.hx - depth * m.aw);
        m.lim_yi[k] = (double)((long double)m.hy - depth * m.bw);
    }
    for (int k = 0; k < n; ++k) {
        const long double depth = -((long double)k - n) * f / n;
        m.lim_xb[k] = (double)((long double)m.hx - depth * m.aw);
        m.lim_yb[k] = (double)((long double)m.hy - depth * m.bw);
    }
    for (int k = 0; k < kMaxQ; ++k) {
        m.lim32_cx[k] = float(m.lim_cx[k]);
        m.lim32_cy[k] = float(m.lim_cy[k]);
        m.lim32_sv[k] = float(m.sv[k]);
        m.lim32_sh[k] = float(m.sh[k]);
    }
    for (int k = 0; k < kMaxLimS; ++k) {
        m.lim32_xi[k] = float(m.lim_xi[k]);
        m.lim32_xb[k] = float(m.lim_xb[k]);
        m.lim32_yi[k] = float(m.lim_yi[k]);
        m.lim32_yb[k] = float(m.lim_yb[k]);
    }
    m.lim_cupper = c_upper;
    m.limit = 1;
    return MP_OK;
}

// First-derivative weights at x of the Lagrange interpolant through the
// nodes 0..npts-1 (unit spacing), each rounded to double as the reference
// stores them (src/mlweno/diffusiveflux.cpp:203-254): w_i = L_i'(x).
void lagrange_weights(int npts, long double x, double* w)
{
    for (int i = 0; i < npts; ++i) {
        long double d = 0;
        for (int k = 0; k < npts; ++k) {
            if (k == i) continue;
            long double p = 1.0L / (i - k);
            for (int l = 0; l < npts; ++l)
                if (l != i && l != k) p *= (x - l) / (long double)(i - l);
            d += p;
        }
        w[i] = (double)d;
    }
}

mp_status set_diffusion_tables(MeshDev& m, double kappa, int samples, double extent)
{
    if (!(std::isfinite(kappa) && kappa >= 0)) return fail(MP_ERR_ARG, "diffusivity must be nonnegative");
    if (samples < 2 || samples % 2 != 0 || samples > kMaxLimS)
        return fail(MP_ERR_ARG, "Diffusion requires an even sample count >= 2 (at most %d here)", kMaxLimS);
    if (!(std::isfinite(extent) && extent > 0 && extent <= 1))
        return fail(MP_ERR_ARG, "Normal sampling extentFraction must lie in (0,1]");
    const int n = samples;
    // Interior face: n samples at (k - (n-1)/2)*ds along the normal, ds =
    // 2 f w/(n-1) with w the width across the face; with the cell averages as
    // samples only the in-cell half differs from the last sample, so
    // F = -kappa * length * S / ds * (H_L - H_R), S = sum_{k<n/2} w_k (< 0).
    double wm[kMaxLimS + 1], we[kMaxLimS + 1];
    lagrange_weights(n, (n - 1) / 2.0L, wm);
    long double S = 0;
    for (int k = 0; k < n / 2; ++k) S += wm[k];
    // Boundary face: n in-cell samples and the face node at the end, ds = f w/n;
    // F = kappa * Hn / ds * (H_in - b) per unit length, Hn = -sum_{k<n} w_k.
    lagrange_weights(n + 1, (long double)n, we);
    long double Hn = 0;
    for (int k = 0; k < n; ++k) Hn -= we[k];
    const double dsx = (double)((long double)extent * m.dx * 2.0L / (n - 1));
    const double dsy = (double)((long double)extent * m.dy * 2.0L / (n - 1));
    const double dbx = (double)((long double)extent * m.dx / n);
    const double dby = (double)((long double)extent * m.dy / n);
    m.dif_x = (double)(-(long double)kappa * m.dy * S / dsx);
    m.dif_y = (double)(-(long double)kappa * m.dx * S / dsy);
    m.dif_bx = (double)((long double)kappa * Hn / dbx);
    m.dif_by = (double)((long double)kappa * Hn / dby);
    return MP_OK;
}

mp_status DeviceBuffer::alloc(size_t n)
{
    reset();
    if (n == 0) return MP_OK;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(MP_ERR_MEMORY, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return MP_OK;
}

void DeviceBuffer::reset()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

bool all_finite(const double* v, size_t n)
{
    bool ok = true;
    for (size_t k = 0; k < n; ++k) ok &= std::isfinite(v[k]);
    return ok;
}

ItemLayout item_layout(int nx, int ny, int rows, int32_t precision)
{
    ItemLayout L;
    L.strips = (nx + kStripCells - 1) / kStripCells;
    L.rows = rows;
    L.split = ny;
    L.tail_rows = rows;
    constexpr int kTailRows = 4;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const long long slots = (long long)cus * 4 * (precision == MP_PREC_MIXED ? 3 : 2);
    const long long uniform = (long long)L.strips * ((ny + rows - 1) / rows);
    if (rows > kTailRows && uniform >= 3 * slots) {
        const long long bands = (2 * slots + L.strips - 1) / L.strips;
        const long long split = ny - bands * kTailRows;
        if (split >= ny / 2) {
            L.split = int(split);
            L.tail_rows = kTailRows;
        }
    }
    L.nlong = L.strips * ((L.split + rows - 1) / rows);
    L.nitems = L.nlong + L.strips * ((ny - L.split + L.tail_rows - 1) / L.tail_rows);
    return L;
}

void apply_layout(const ItemLayout& L, StagePtrs& sp)
{
    sp.rows = L.rows;
    sp.strips = L.strips;
    sp.split = L.split;
    sp.tail_rows = L.tail_rows;
    sp.nlong = L.nlong;
    sp.pad = 0;
}

}  // namespace mp
