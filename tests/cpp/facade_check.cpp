// A compiled C++ caller of include/mantlepar_amd.hpp: the reference's call
// signatures (mantle::phase::EutecticModel, mantle::couple::StateReconstruction)
// driven the way the reference's own callers drive them, checked against
// fixtures that tests/test_gpu_cpp_facade.py writes from the committed golden
// vectors.  Exit status 0 when every check holds.
//
//   facade_check phase <dir>      golden EutecticModel cases (reference build)
//   facade_check pressure <dir>   golden staticPressure cases
//   facade_check transport <dir>  the 20x20 SSP-RK3 run against the oracle
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mantlepar_amd.hpp"

using mantle::phase::EutecticModel;
using mantle::phase::MaterialParameters;
using mantle::phase::PhaseState;

namespace {

int failures = 0;

void check(bool ok, const std::string& what)
{
    if (!ok) {
        if (failures < 20) std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        ++failures;
    }
}

template <typename T>
std::vector<T> load(const std::string& dir, const std::string& name)
{
    std::ifstream f(dir + "/" + name, std::ios::binary);
    if (!f) throw std::runtime_error("missing fixture " + name);
    std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<T> v(bytes.size() / sizeof(T));
    std::memcpy(v.data(), bytes.data(), v.size() * sizeof(T));
    return v;
}

MaterialParameters material(const std::vector<double>& params, std::size_t set)
{
    const double* r = params.data() + set * 13;
    MaterialParameters p;
    p.Tm0 = r[0]; p.Te0 = r[1]; p.nu = r[2]; p.L = r[3]; p.cp = r[4]; p.Xe = r[5];
    p.rho = r[6]; p.rhor = r[7]; p.mus = r[8]; p.mul = r[9]; p.k0 = r[10]; p.g = r[11];
    p.alpha0 = r[12];
    return p;
}

// 0 no exception, 1 invalid_argument, 2 overflow_error, 3 runtime_error
template <typename F>
int thrown(F&& f)
{
    try {
        f();
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (const std::overflow_error&) {
        return 2;
    } catch (const std::runtime_error&) {
        return 3;
    }
    return 0;
}

bool near_boundary(const EutecticModel& m, double H, double C, double P)
{
    const auto b = m.boundaries(C, P);
    for (double v : {b.Tep, b.Tmp, b.H_eutectic_end, b.H_liquidus})
        if (std::fabs(H - v) <= 8 * (std::nextafter(std::fabs(v), INFINITY) - std::fabs(v))) return true;
    return false;
}

// Golden columns: phi1 phi2 phil cl cs region TDp Tep dTD_dCD dTD_dHD flags.
int run_phase(const std::string& dir)
{
    const auto params = load<double>(dir, "params.f64");
    const auto derived = load<double>(dir, "derived.f64");
    const auto set = load<int32_t>(dir, "set.i32");
    const auto H = load<double>(dir, "H.f64");
    const auto C = load<double>(dir, "C.f64");
    const auto P = load<double>(dir, "P.f64");
    const auto status = load<int32_t>(dir, "status.i32");
    const auto out = load<double>(dir, "out.f64");
    const std::size_t nsets = params.size() / 13, n = H.size();
    std::size_t evaluated = 0, band = 0, rejected = 0;
    for (std::size_t s = 0; s < nsets; ++s) {
        const EutecticModel model(material(params, s));
        const auto& d = model.derived();
        const double dv[11] = {d.dT, d.h0, d.gamma, d.LD, d.TDm0, d.TDe0, d.invk0, d.l0, d.u0, d.p0, d.t0};
        for (int k = 0; k < 11; ++k) check(dv[k] == derived[s * 11 + k], "derived scale " + std::to_string(k));
        std::vector<double> bh, bc, bp;
        std::vector<PhaseState> single;
        for (std::size_t k = 0; k < n; ++k) {
            if (std::size_t(set[k]) != s) continue;
            const std::string tag = "set " + std::to_string(s) + " case " + std::to_string(k);
            PhaseState st;
            const int e = thrown([&] { st = model.evaluate(H[k], C[k], P[k]); });
            check(e == status[k], tag + ": exception kind " + std::to_string(e));
            if (status[k] != 0) { ++rejected; continue; }
            if (e) continue;
            ++evaluated;
            const double* r = out.data() + k * 11;
            const bool inband = near_boundary(model, H[k], C[k], P[k]);
            band += inband;
            if (!inband) check(int(st.region) == int(r[5]), tag + ": region");
            if (int(st.region) == int(r[5])) {
                const double v[5] = {st.phi1, st.phi2, st.phil, st.TDp, st.Tep};
                const int col[5] = {0, 1, 2, 6, 7};
                for (int q = 0; q < 5; ++q)
                    check(std::fabs(v[q] - r[col[q]]) <= 2e-14 + 2e-14 * std::fabs(r[col[q]]),
                          tag + ": column " + std::to_string(col[q]));
                const double solid = r[0] + r[1];
                check(std::fabs(st.cs - r[4]) <= 2e-14 + 4e-19 / std::fmax(solid, 1e-300), tag + ": cs");
                check(std::fabs(st.cl - r[3]) <= 2e-14 + 4e-19 / std::fmax(r[2], 1e-300), tag + ": cl");
                check(st.has_solid == (solid > 0.0) && st.has_liquid == (r[2] > 0.0), tag + ": presence");
            }
            bh.push_back(H[k]); bc.push_back(C[k]); bp.push_back(P[k]);
            single.push_back(st);
        }
        // the array overload gives the single-state calls' results
        std::vector<PhaseState> many(bh.size());
        model.evaluate(bh.size(), bh.data(), bc.data(), bp.data(), many.data());
        for (std::size_t k = 0; k < many.size(); ++k)
            check(many[k].TDp == single[k].TDp && many[k].region == single[k].region &&
                      many[k].phil == single[k].phil, "array overload case " + std::to_string(k));
    }
    // classify and setParameters' commit-on-success (src/phase/eutectic_rescaled.h:85-87)
    EutecticModel m(material(params, 0));
    const double Tm0 = m.parameters().Tm0;
    MaterialParameters bad = m.parameters();
    bad.Xe = 1.5;
    check(thrown([&] { m.setParameters(bad); }) == 1 && m.parameters().Tm0 == Tm0 && m.parameters().Xe != 1.5,
          "rejected setParameters leaves the model unchanged");
    check(m.classify(0.0, 0.0, 0.0) == mantle::phase::PhaseRegion::PureSolid, "classify");
    std::ostringstream info;
    m.printInfo(info);
    check(info.str().find("Dimensionless latent heat") != std::string::npos, "printInfo");
    std::printf("facade_check phase: %zu states evaluated, %zu in the boundary band, %zu rejected, %d failures\n",
                evaluated, band, rejected, failures);
    return failures ? 1 : 0;
}

int run_pressure(const std::string& dir)
{
    const auto params = load<double>(dir, "params.f64");
    const auto set = load<int64_t>(dir, "set.i64");
    const auto zD = load<double>(dir, "zD.f64");
    const auto length = load<double>(dir, "length.f64");
    const auto scaled = load<int8_t>(dir, "scaled.i8");
    const auto P = load<double>(dir, "P.f64");
    const auto status = load<int32_t>(dir, "status.i32");
    std::size_t ok = 0;
    for (std::size_t k = 0; k < zD.size(); ++k) {
        const EutecticModel model(material(params, std::size_t(set[k])));
        double v = 0.0;
        const int e = thrown([&] { v = scaled[k] ? model.staticPressure(zD[k], length[k]) : model.staticPressure(zD[k]); });
        check(e == status[k], "pressure case " + std::to_string(k) + ": exception kind");
        if (e == 0 && status[k] == 0) {
            check(v == P[k], "pressure case " + std::to_string(k) + ": value");
            ++ok;
        }
    }
    std::printf("facade_check pressure: %zu pressures bitwise, %d failures\n", ok, failures);
    return failures ? 1 : 0;
}

int run_transport(const std::string& dir)
{
    std::ifstream meta(dir + "/meta.txt");
    mp_mesh_desc mesh{};
    double dt = 0.0;
    int nsteps = 0;
    meta >> mesh.nx >> mesh.ny >> mesh.q >> mesh.x0 >> mesh.x1 >> mesh.y0 >> mesh.y1 >> dt >> nsteps;
    const auto H = load<double>(dir, "H.f64");
    const auto C = load<double>(dir, "C.f64");
    const auto vx = load<double>(dir, "vnx.f64");
    const auto vy = load<double>(dir, "vny.f64");
    const auto Href = load<double>(dir, "Href.f64");
    const auto Cref = load<double>(dir, "Cref.f64");
    mp_phase_params phase;
    mp_default_phase_params(&phase);
    phase.Tm0 = 1600.0; phase.Te0 = 600.0; phase.L = 1.2e6; phase.Xe = 0.4; phase.hypereutectic = 1;
    double worst = 0.0;
    for (int32_t precision : {MP_PREC_FP64, MP_PREC_MIXED}) {
        mantle::couple::StateReconstruction rec;
        mantle::couple::InitialState s;
        s.mesh = mesh;
        s.H = H.data();
        s.C = C.data();
        check(rec.Initialize(s, nullptr, &phase, precision) == MP_OK, std::string("Initialize: ") + mp_last_error());
        check(rec.Initialize(s, nullptr, &phase, precision) == MP_ERR_STATE, "second Initialize is rejected");
        check(rec.SetFaceVelocity(vx.data(), vy.data()) == MP_OK, "SetFaceVelocity");
        double dt_dev = 0.0;
        check(rec.StableStep(0.4, dt_dev) == MP_OK && std::fabs(dt_dev - dt) <= 1e-14 * dt, "StableStep");
        mp_step_report rep{};
        check(rec.Step(nsteps, MP_SCHEME_SSPRK3, dt, &rep) == MP_OK, std::string("Step: ") + mp_last_error());
        double mh = 0, mc = 0, eh = 0, ec = 0;
        for (std::size_t k = 0; k < Href.size(); ++k) {
            mh = std::fmax(mh, std::fabs(Href[k]));
            mc = std::fmax(mc, std::fabs(Cref[k]));
            eh = std::fmax(eh, std::fabs(rec.AveragesH()[k] - Href[k]));
            ec = std::fmax(ec, std::fabs(rec.AveragesC()[k] - Cref[k]));
        }
        const double err = std::fmax(eh / mh, ec / mc);
        const double tol = precision == MP_PREC_FP64 ? 1e-12 : 1e-6;
        check(err <= tol, "stepped state vs oracle: " + std::to_string(err));
        if (precision == MP_PREC_FP64) worst = err;
        // Evaluate at a cell centre, AverageH, and the reference's rejections
        const mantle::couple::MeshIndex c{3, 4};
        mantle::couple::Point p;
        p.p[0] = mesh.x0 + 3.5 * (mesh.x1 - mesh.x0) / mesh.nx;
        p.p[1] = mesh.y0 + 4.5 * (mesh.y1 - mesh.y0) / mesh.ny;
        double eH = 0, eC = 0, eH2 = 0;
        check(rec.Evaluate(c, p, eH, eC) == MP_OK && rec.Enthalpy(c, p, eH2) == MP_OK && eH == eH2,
              "Evaluate / Enthalpy");
        check(std::fabs(eH - rec.AverageH(c)) <= 0.5 * mh, "reconstruction near the average");
        double keep = -1.0;
        check(rec.Evaluate({mesh.nx, 0}, p, keep, keep) == MP_ERR_ARG && keep == -1.0,
              "Evaluate outside the mesh is rejected, outputs untouched");
        bool out_of_range = false;
        try { (void)rec.AverageC({-1, 0}); } catch (const std::out_of_range&) { out_of_range = true; }
        check(out_of_range, "AverageC outside the mesh throws std::out_of_range");
        mantle::couple::InitialState wrong = s;
        wrong.mesh.nx += 1;
        check(rec.Update(wrong) == MP_ERR_ARG, "Update with another mesh is rejected");
        // Update of H only keeps the stepped C
        const double c_before = rec.AverageC(c);
        std::vector<double> H2(H.size(), 0.5), C2(C.size(), 0.1);
        s.H = H2.data();
        s.C = C2.data();
        check(rec.Update(s, false) == MP_OK && rec.AverageH(c) == 0.5 && rec.AverageC(c) == c_before,
              "Update(updateComposition = false)");
        check(rec.Destroy() == MP_OK && rec.Step(1, MP_SCHEME_SSPRK3, dt) == MP_ERR_STATE,
              "Destroy, then Step is rejected");
    }
    std::printf("facade_check transport: %dx%d, %d SSP-RK3 steps, FP64 max rel err %.3e, %d failures\n",
                mesh.nx, mesh.ny, nsteps, worst, failures);
    return failures ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: facade_check phase|pressure|transport <fixture dir>\n");
        return 2;
    }
    const std::string mode = argv[1], dir = argv[2];
    try {
        if (mode == "phase") return run_phase(dir);
        if (mode == "pressure") return run_pressure(dir);
        if (mode == "transport") return run_transport(dir);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_check %s: %s\n", mode.c_str(), e.what());
        return 1;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
