// C entry point around the reference's own EutecticModel (src/phase), compiled
// unmodified from where it lies by `make -C oracle ref`.  Test infrastructure:
// it writes golden vectors and cross-checks the C restatement in
// mp_oracle.c; nothing in the product links it.
#include "eutectic_rescaled.h"

#include <cstdint>
#include <cstring>
#include <stdexcept>

using mantle::phase::EutecticModel;
using mantle::phase::MaterialParameters;

extern "C" {

// params: Tm0, Te0, nu, L, cp, Xe, rho, rhor, mus, mul, k0, g, alpha0
// out[k*11 + ...] = phi1, phi2, phil, cl, cs, region, TDp, Tep, dTD_dCD, dTD_dHD, flags
// status[k]: 0 ok, 1 invalid_argument, 2 overflow_error, 3 runtime_error
int ref_phase_evaluate(const double* params, long n, const double* H, const double* C,
                       const double* P, double* out, int* status)
{
    MaterialParameters p;
    p.Tm0 = params[0]; p.Te0 = params[1]; p.nu = params[2]; p.L = params[3];
    p.cp = params[4]; p.Xe = params[5]; p.rho = params[6]; p.rhor = params[7];
    p.mus = params[8]; p.mul = params[9]; p.k0 = params[10]; p.g = params[11];
    p.alpha0 = params[12];
    try {
        const EutecticModel model(p);
        for (long k = 0; k < n; ++k) {
            double* o = out + k * 11;
            try {
                const auto s = model.evaluate(H[k], C[k], P[k]);
                o[0] = s.phi1; o[1] = s.phi2; o[2] = s.phil; o[3] = s.cl; o[4] = s.cs;
                o[5] = static_cast<double>(static_cast<int>(s.region));
                o[6] = s.TDp; o[7] = s.Tep; o[8] = s.dTD_dCD; o[9] = s.dTD_dHD;
                o[10] = (s.has_solid ? 1 : 0) + (s.has_liquid ? 2 : 0) +
                        (s.finite_temperature_derivatives ? 4 : 0);
                status[k] = 0;
            } catch (const std::invalid_argument&) { status[k] = 1; }
            catch (const std::overflow_error&) { status[k] = 2; }
            catch (const std::runtime_error&) { status[k] = 3; }
        }
    } catch (const std::exception&) {
        return -1;
    }
    return 0;
}

// staticPressure(zD[k]) (scaled == 0) or staticPressure(zD[k], length[k]);
// status[k] as in ref_phase_evaluate
int ref_static_pressure(const double* params, long n, const double* zD, const double* length,
                        int scaled, double* out, int* status)
{
    MaterialParameters p;
    p.Tm0 = params[0]; p.Te0 = params[1]; p.nu = params[2]; p.L = params[3];
    p.cp = params[4]; p.Xe = params[5]; p.rho = params[6]; p.rhor = params[7];
    p.mus = params[8]; p.mul = params[9]; p.k0 = params[10]; p.g = params[11];
    p.alpha0 = params[12];
    try {
        const EutecticModel model(p);
        for (long k = 0; k < n; ++k) {
            try {
                out[k] = scaled ? model.staticPressure(zD[k], length[k]) : model.staticPressure(zD[k]);
                status[k] = 0;
            } catch (const std::invalid_argument&) { status[k] = 1; }
            catch (const std::overflow_error&) { status[k] = 2; }
            catch (const std::runtime_error&) { status[k] = 3; }
        }
    } catch (const std::exception&) {
        return -1;
    }
    return 0;
}

// derived: dT, h0, gamma, LD, TDm0, TDe0, invk0, l0, u0, p0, t0
int ref_phase_derived(const double* params, double* derived)
{
    MaterialParameters p;
    p.Tm0 = params[0]; p.Te0 = params[1]; p.nu = params[2]; p.L = params[3];
    p.cp = params[4]; p.Xe = params[5]; p.rho = params[6]; p.rhor = params[7];
    p.mus = params[8]; p.mul = params[9]; p.k0 = params[10]; p.g = params[11];
    p.alpha0 = params[12];
    try {
        const EutecticModel model(p);
        const auto& d = model.derived();
        const double v[11] = {d.dT, d.h0, d.gamma, d.LD, d.TDm0, d.TDe0, d.invk0,
                              d.l0, d.u0, d.p0, d.t0};
        std::memcpy(derived, v, sizeof v);
    } catch (const std::invalid_argument&) { return 1; }
    catch (const std::overflow_error&) { return 2; }
    return 0;
}

}
