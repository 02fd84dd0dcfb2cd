/*
 * mantlepar_amd.hpp -- the reference's C++ call signatures over the C ABI.
 *
 * A C++ caller of aband/MantlePar's phase model and state reconstruction
 * includes this header instead of the reference's and links
 * libmantlepar_amd.so; the work runs on the GPU through mantlepar_amd.h.
 *
 *   mantle::phase::EutecticModel      src/phase/eutectic_rescaled.h:103-136
 *     with MaterialParameters, DerivedParameters, PhaseRegion,
 *     PhaseBoundaries and PhaseState as declared at :10-76 (same members,
 *     defaults and numbering).  Errors throw what the reference throws:
 *     std::invalid_argument for invalid or non-finite inputs and compositions
 *     outside [0, Xe], std::overflow_error for unrepresentable scales and
 *     temperatures, std::runtime_error for a phase fraction outside [0, 1].
 *   mantle::couple::StateReconstruction  src/couple/state_reconstruction.h:11-22
 *     Initialize / Update / Limit / Destroy / Evaluate / Enthalpy / AverageH /
 *     AverageC over one rank's uniform rectangular mesh, plus the transport
 *     step the reference builds around it (preheat.cpp:160-284,
 *     evolve.cpp:449-507): SetFaceVelocity, StableStep, Step, Phase.  PETSc
 *     types are replaced by plain ones: methods return mp_status (0 on
 *     success, a PetscErrorCode-like category otherwise, the message in
 *     mp_last_error()), fields are row-major j*nx+i host arrays instead of
 *     Vecs, and there is no communicator (one handle per rank).
 *
 * EutecticModel::evaluate(HD, CD, P) evaluates one state with one device
 * call, as a drop-in; callers with many states use the array overload, one
 * launch for all of them.
 */
#ifndef MANTLEPAR_AMD_HPP
#define MANTLEPAR_AMD_HPP

#include <cmath>
#include <cstddef>
#include <limits>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mantlepar_amd.h"

namespace mantle::phase {

struct MaterialParameters {
    double Tm0 = 2000.0;
    double Te0 = 1480.0;
    double nu = 6.5e6;
    double L = 5.0e5;
    double cp = 1200.0;
    double Xe = 0.25;
    double rho = 3000.0;
    double rhor = 500.0;
    double mus = 1.0e19;
    double mul = 1.0;
    double k0 = 1.0e-8;
    double g = 10.0;
    double alpha0 = 3.0e-5;
};

struct DerivedParameters {
    double dT = 0.0, h0 = 0.0, gamma = 0.0, LD = 0.0, TDm0 = 0.0, TDe0 = 0.0;
    double invk0 = 0.0, l0 = 0.0, u0 = 0.0, p0 = 0.0, t0 = 0.0;
};

enum class PhaseRegion {
    PureSolid = 1,
    TwoSolids = 2,
    Eutectic = 3,
    SolidLiquid = 4,
    Liquid = 5,
    PureMelting = 6
};

struct PhaseBoundaries {
    double Tep = 0.0;
    double Tmp = 0.0;
    double H_eutectic_end = 0.0;
    double H_liquidus = 0.0;
};

struct PhaseState {
    double phi1 = 0.0;
    double phi2 = 0.0;
    double phil = 0.0;
    double cl = 0.0;
    double cs = 0.0;
    PhaseRegion region = PhaseRegion::PureSolid;
    double TDp = 0.0;
    double Tep = 0.0;
    double dTD_dCD = 0.0;
    double dTD_dHD = 0.0;
    bool has_solid = false;
    bool has_liquid = false;
    bool finite_temperature_derivatives = true;
};

namespace detail {

// The reference's exception for a status of the library.
inline void raise(mp_status s)
{
    if (s == MP_OK) return;
    const std::string msg = mp_last_error();
    switch (s) {
    case MP_ERR_ARG:
    case MP_ERR_COMPOSITION:
        throw std::invalid_argument(msg);
    case MP_ERR_FP:
        throw std::overflow_error(msg);
    default:
        throw std::runtime_error(msg);
    }
}

inline void require_finite(double v, const char* name)
{
    if (!std::isfinite(v)) throw std::invalid_argument(std::string(name) + " must be finite");
}

}  // namespace detail

// EutecticModel (src/phase/eutectic_rescaled.h:103-136).  setParameters
// validates and derives before committing; a rejected update leaves the
// model unchanged.  evaluate/classify/boundaries are const and may run
// concurrently unless setParameters runs at the same time.
class EutecticModel {
public:
    explicit EutecticModel(const MaterialParameters& parameters = {}) { setParameters(parameters); }

    const MaterialParameters& parameters() const noexcept { return parameters_; }
    const DerivedParameters& derived() const noexcept { return derived_; }

    void setParameters(const MaterialParameters& p)
    {
        const mp_phase_params c = to_c(p);
        mp_phase_derived d;
        detail::raise(mp_phase_derive(&c, &d));
        parameters_ = p;
        derived_ = DerivedParameters{d.dT, d.h0, d.gamma, d.LD, d.TDm0, d.TDe0,
                                     d.invk0, d.l0, d.u0, d.p0, d.t0};
        params_ = c;
    }

    PhaseBoundaries boundaries(double CD, double P) const
    {
        // the reference checks the composition, then the pressure
        detail::require_finite(CD, "CD");
        if (CD < 0.0 || CD > parameters_.Xe) throw std::invalid_argument("CD must lie in [0, Xe]");
        PhaseBoundaries b;
        detail::raise(mp_phase_boundaries(&params_, CD, P, &b.Tep, &b.Tmp, &b.H_eutectic_end,
                                          &b.H_liquidus));
        return b;
    }

    PhaseRegion classify(double HD, double CD, double P) const { return evaluate(HD, CD, P).region; }

    PhaseState evaluate(double HD, double CD, double P) const
    {
        PhaseState s;
        evaluate(1, &HD, &CD, &P, &s);
        return s;
    }

    // n states in one device launch; P may be null (P = 0 everywhere).  The
    // inputs are checked in order first, as n calls of evaluate would, so the
    // first invalid state decides the exception and out is left untouched.
    void evaluate(std::size_t n, const double* HD, const double* CD, const double* P,
                  PhaseState* out) const
    {
        std::vector<double> tep(n);
        for (std::size_t k = 0; k < n; ++k) {
            detail::require_finite(HD[k], "HD");
            tep[k] = boundaries(CD[k], P ? P[k] : 0.0).Tep;
        }
        std::vector<double> T(n), phi(n), Cl(n), Cs(n), phi1(n), phi2(n), dTdC(n), dTdH(n);
        std::vector<uint8_t> region(n);
        detail::raise(mp_phase_evaluate(&params_, int64_t(n), HD, CD, P, T.data(),
                                        phi.data(), Cl.data(), Cs.data(), region.data(),
                                        phi1.data(), phi2.data(), dTdC.data(), dTdH.data()));
        for (std::size_t k = 0; k < n; ++k) {
            PhaseState s;
            s.phi1 = phi1[k];
            s.phi2 = phi2[k];
            s.phil = phi[k];
            s.cl = Cl[k];
            s.cs = Cs[k];
            s.region = static_cast<PhaseRegion>(region[k]);
            s.TDp = T[k];
            s.Tep = tep[k];
            s.dTD_dCD = dTdC[k];
            s.dTD_dHD = dTdH[k];
            s.has_solid = s.phi1 + s.phi2 > 0.0;
            s.has_liquid = s.phil > 0.0;
            s.finite_temperature_derivatives = std::isfinite(s.dTD_dCD) && std::isfinite(s.dTD_dHD);
            out[k] = s;
        }
    }

    double pressureCorrectedTemperature(double TD, double P) const
    {
        double v = 0.0;
        detail::raise(mp_phase_pressure_corrected_temperature(&params_, TD, P, &v));
        return v;
    }

    double staticPressure(double zD) const
    {
        double v = 0.0;
        detail::raise(mp_phase_static_pressure(&params_, zD, 0.0, 0, &v));
        return v;
    }

    double staticPressure(double zD, double length_scale) const
    {
        double v = 0.0;
        detail::raise(mp_phase_static_pressure(&params_, zD, length_scale, 1, &v));
        return v;
    }

    void printInfo(std::ostream& out) const
    {
        out << "Dimensionless melting temperature: " << derived_.TDm0 << '\n'
            << "Dimensionless eutectic temperature: " << derived_.TDe0 << '\n'
            << "Dimensionless latent heat: " << derived_.LD << '\n'
            << "Temperature scale [K]: " << derived_.dT << '\n'
            << "Enthalpy scale [J/kg]: " << derived_.h0 << '\n'
            << "Length scale [m]: " << derived_.l0 << '\n'
            << "Velocity scale [m/s]: " << derived_.u0 << '\n'
            << "Pressure scale [Pa]: " << derived_.p0 << '\n'
            << "Time scale [s]: " << derived_.t0 << '\n';
    }

private:
    static mp_phase_params to_c(const MaterialParameters& p)
    {
        mp_phase_params c;
        mp_default_phase_params(&c);
        c.Tm0 = p.Tm0; c.Te0 = p.Te0; c.nu = p.nu; c.L = p.L; c.cp = p.cp; c.Xe = p.Xe;
        c.rho = p.rho; c.rhor = p.rhor; c.mus = p.mus; c.mul = p.mul; c.k0 = p.k0; c.g = p.g;
        c.alpha0 = p.alpha0;
        c.pressure = 0.0;
        c.hypereutectic = 0;   // the reference's diagram: C > Xe is rejected
        return c;
    }

    MaterialParameters parameters_{};
    DerivedParameters derived_{};
    mp_phase_params params_{};
};

}  // namespace mantle::phase

namespace mantle::couple {

struct MeshIndex {
    int i = 0;
    int j = 0;
};

struct Point {
    double p[2]{};
};

// The transport state handed to Initialize / Update: the mesh and the cell
// averages of H and C (row-major j*nx+i), in place of the reference's
// InitialState with its DMDA and Vecs.
struct InitialState {
    mp_mesh_desc mesh{};
    const double* H = nullptr;
    const double* C = nullptr;
};

// The members of Configuration that Limit reads
// (src/couple/state_reconstruction.cpp:136-192): cell Gauss points per axis
// and the normal-line diffusion sampling.
struct Configuration {
    struct {
        struct {
            int cellPointsPerAxis = 5;
        } quadrature;
    } input;
    struct {
        int numberOfSamples = 4;
        double extentFraction = 0.9;
    } sampling;
};

// StateReconstruction (src/couple/state_reconstruction.h:11-22) of one rank's
// uniform mesh on the GPU.  The reconstruction itself is rebuilt by the device
// wherever it is needed; Update replaces the averages it is built from.
// Failed calls leave the previous state in place.  Not copyable; Destroy (or
// the destructor) releases the device memory.
class StateReconstruction {
public:
    StateReconstruction() = default;
    StateReconstruction(const StateReconstruction&) = delete;
    StateReconstruction& operator=(const StateReconstruction&) = delete;
    ~StateReconstruction() { Destroy(); }

    // weno/phase: the production options and the reference's default
    // material when null; precision MP_PREC_FP64 or MP_PREC_MIXED.
    mp_status Initialize(const InitialState& s, const mp_weno_options* weno = nullptr,
                         const mp_phase_params* phase = nullptr, int32_t precision = MP_PREC_FP64)
    {
        if (t_) return set_state_error("Reconstruction already initialized");
        mp_transport* t = nullptr;
        mp_status st = mp_transport_create(&s.mesh, weno, phase, precision, nullptr, &t);
        if (st) return st;
        t_ = t;
        mesh_ = s.mesh;
        st = Update(s, true);
        if (st) Destroy();
        return st;
    }

    // New cell averages of H, and of C unless updateComposition is false.
    mp_status Update(const InitialState& s, bool updateComposition = true)
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        if (s.mesh.nx != mesh_.nx || s.mesh.ny != mesh_.ny)
            return mp_report_error(MP_ERR_ARG, "InitialState mesh differs from the initialized mesh");
        const std::size_t n = std::size_t(mesh_.nx) * mesh_.ny;
        std::vector<double> c;
        const double* C = s.C;
        if (!updateComposition) {
            c.resize(n);
            mp_status st = mp_transport_download_state(t_, nullptr, c.data());
            if (st) return st;
            C = c.data();
        }
        mp_status st = mp_transport_upload_state(t_, s.H, C);
        if (st) return st;
        return refresh_averages();
    }

    mp_status Limit(const Configuration& c)
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        return mp_transport_set_limiter(t_, 1, c.input.quadrature.cellPointsPerAxis,
                                        c.sampling.numberOfSamples, c.sampling.extentFraction);
    }

    mp_status Destroy()
    {
        if (t_) mp_transport_destroy(t_);
        t_ = nullptr;
        return MP_OK;
    }

    mp_status Evaluate(MeshIndex cell, const Point& p, double& H, double& C) const
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        const int32_t k = cell_index(cell);
        double h = 0.0, c = 0.0;
        const mp_status st = mp_transport_evaluate(t_, 1, &k, &p.p[0], &p.p[1], &h, &c);
        if (st == MP_OK) { H = h; C = c; }
        return st;
    }

    mp_status Enthalpy(MeshIndex cell, const Point& p, double& H) const
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        const int32_t k = cell_index(cell);
        double h = 0.0;
        const mp_status st = mp_transport_evaluate(t_, 1, &k, &p.p[0], &p.p[1], &h, nullptr);
        if (st == MP_OK) H = h;
        return st;
    }

    // Many points in one launch (cells[k] = j*nx+i).
    mp_status Evaluate(std::size_t n, const int32_t* cells, const double* x, const double* y,
                       double* H, double* C) const
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        return mp_transport_evaluate(t_, int64_t(n), cells, x, y, H, C);
    }

    double AverageH(MeshIndex cell) const { return averagesH_.at(std::size_t(cell_index(cell))); }
    double AverageC(MeshIndex cell) const { return averagesC_.at(std::size_t(cell_index(cell))); }

    // ---- the transport step around the reconstruction -------------------
    mp_status SetFaceVelocity(const double* vnx, const double* vny)
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        return mp_transport_set_face_velocity(t_, vnx, vny);
    }
    mp_status StableStep(double cfl, double& dt) const
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        return mp_transport_stable_dt(t_, cfl, &dt);
    }
    // nsteps of the scheme at fixed dt, each ending with the per-cell phase;
    // the averages follow the new state.
    mp_status Step(int32_t nsteps, int32_t scheme, double dt, mp_step_report* report = nullptr)
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        mp_status st = mp_transport_step(t_, nsteps, scheme, dt, report);
        if (st) return st;
        return refresh_averages();
    }
    mp_status Phase(double* T, double* phi, double* Cl, double* Cs, uint8_t* region) const
    {
        if (!t_) return set_state_error("Reconstruction is not initialized");
        return mp_transport_phase(t_, T, phi, Cl, Cs, region);
    }
    const std::vector<double>& AveragesH() const { return averagesH_; }
    const std::vector<double>& AveragesC() const { return averagesC_; }
    mp_transport* handle() const { return t_; }

private:
    int32_t cell_index(MeshIndex c) const
    {
        if (c.i < 0 || c.j < 0 || c.i >= mesh_.nx || c.j >= mesh_.ny) return -1;   // rejected
        return c.j * mesh_.nx + c.i;
    }
    mp_status refresh_averages()
    {
        const std::size_t n = std::size_t(mesh_.nx) * mesh_.ny;
        averagesH_.resize(n);
        averagesC_.resize(n);
        return mp_transport_download_state(t_, averagesH_.data(), averagesC_.data());
    }
    static mp_status set_state_error(const char* msg) { return mp_report_error(MP_ERR_STATE, msg); }

    mp_transport* t_ = nullptr;
    mp_mesh_desc mesh_{};
    std::vector<double> averagesH_, averagesC_;
};

}  // namespace mantle::couple

#endif
