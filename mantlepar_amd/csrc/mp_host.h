// Host-side helpers shared by the single-mesh, batched and serving front ends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/mantlepar_amd.h"
#include "mp_common.h"

namespace mp {

// Thread-local last error (mp_last_error()).
void set_error(const char* fmt, ...);
mp_status fail(mp_status code, const char* fmt, ...);
mp_status status_from_bits(uint32_t bits);
const char* last_error();
const char* describe_bits(uint32_t bits);

#define MP_HIP(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::mp::fail(MP_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define MP_CALL(expr)                  \
    do {                               \
        mp_status s_ = (expr);         \
        if (s_ != MP_OK) return s_;    \
    } while (0)

// C ABI bodies: host exceptions become status codes.
#define MP_GUARD(...)                                                      \
    try {                                                                  \
        __VA_ARGS__                                                        \
    } catch (const std::bad_alloc&) {                                      \
        return ::mp::fail(MP_ERR_MEMORY, "host allocation failed");        \
    } catch (const std::exception& e) {                                    \
        return ::mp::fail(MP_ERR_DEVICE, "%s", e.what());                  \
    }

// Gauss-Legendre rule on [-1,1], ascending (Newton on P_n in long double).
void gauss_legendre(int n, double* x, double* w);

// Derived phase scales (EutecticModel::derive, src/phase/eutectic_rescaled.cpp:42-86)
// and the uniform-pressure boundaries used by the kernels.
struct PhaseHost {
    PhaseDev dev;
    double TDe0, nu, dT;
    double rho, g, l0;      // staticPressure's factors (l0 = sqrt(mus*k0/mul))
};
mp_status make_phase(const mp_phase_params& p, PhaseHost& out);

// Lithostatic {Tep, Tmp} at nondimensional ordinate y below the surface
// surface_y: P = staticPressure(surface_y - y) (src/phase/eutectic_rescaled.cpp:
// 151-161, src/couple/initialization.cpp:159-161), then the boundaries' pressure
// shift (:143-149, :166-169), each rounded once from long double.
mp_status lithostatic_pair(const PhaseHost& h, double surface_y, double y, TepPair& out);
// {Tep, Tmp} of each cell row at its centre (MapCellPoint of (0, 0) over the
// uniform vertices) for the pressure model "lithostatic"
// (src/couple/initialization.cpp:155-162); rejects a surface below the mesh top.
mp_status lithostatic_rows(const mp_mesh_desc& mesh, const PhaseHost& h, double surface_y,
                           std::vector<TepPair>& rows);
// Vertex ordinates of a uniform axis a0 + k*(a1 - a0)/n with the ends snapped
// (src/core/mesh.cpp:27-33, 86-114); v holds n + 1 values.
void uniform_vertices(double a0, double a1, int n, double* v);

// Work items of one mesh's sweep: `rows` per item over rows [0, split) and
// tail_rows per item above, nlong + tail items in all (nitems).  A mesh that
// gives every wave slot of the chip three or more items keeps its last rows
// for 4-row tail items, about two per slot, which the single-mesh launch
// dispatches last; otherwise split = ny and every item is `rows` tall.  The
// single-mesh handle and the batch engine share it, so a mesh's step totals,
// sums over its items in item order, are the same in both.
struct ItemLayout {
    int strips, rows, split, tail_rows, nlong, nitems;
};
ItemLayout item_layout(int nx, int ny, int rows, int32_t precision);
void apply_layout(const ItemLayout& L, StagePtrs& sp);
// MapEdgePoint's combination of two end values at reference t in [-1, 1]
// (src/core/integral.cpp:91-97).
inline double map_edge(double t, double a, double b)
{
    return (0.5 * (1.0 - t)) * a + (0.5 * (1.0 + t)) * b;
}

mp_status validate_mesh(const mp_mesh_desc& m);
mp_status validate_weno(const mp_weno_options& o);
// Closed-form tables of the production candidates for one uniform mesh.
// Bound limiter point tables (src/couple/state_reconstruction.cpp:136-192 with
// the sampling of src/mlweno/diffusiveflux.cpp:256-338) for a uniform mesh.
mp_status set_limiter_tables(MeshDev& m, int cell_points, int samples, double extent,
                             double c_upper);
// First-derivative weights at x of the Lagrange interpolant through nodes
// 0..npts-1 (src/mlweno/diffusiveflux.cpp:203-254), rounded to double.
void lagrange_weights(int npts, long double x, double* w);
// Thermal diffusion coefficients of the legacy preheat form
// (src/couple/preheat.cpp:176-190, src/mlweno/diffusiveflux.cpp:256-338).
mp_status set_diffusion_tables(MeshDev& m, double kappa, int samples, double extent);
mp_status make_mesh_dev(const mp_mesh_desc& mesh, const mp_weno_options& weno,
                        const PhaseHost& phase, MeshDev& out);

// Device allocation helper that frees on scope exit unless released.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) { reset(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DeviceBuffer() { reset(); }
    mp_status alloc(size_t n);
    void reset();
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

// Check host arrays for finite values (the reference rejects non-finite
// averages and velocities: src/mlweno/reconstruction.cpp:299-320,
// src/mlweno/advectiveflux.cpp:177-179).
bool all_finite(const double* v, size_t n);

}  // namespace mp
