// Host side of the coupled transport law: the mp_coupled_* C ABI and the
// SSPRK2 step loop with admissibility/CFL retries of src/couple/evolve.cpp:428-511.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mp_coupled.h"
#include "mp_host.h"

using namespace mp;

struct mp_coupled {
    mp_mesh_desc mesh{};
    mp_phase_params phase_params{};
    mp_coupled_params params{};
    PhaseHost phase{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    MeshDev md{};
    CoupledDev cd{};
    size_t N = 0;
    int64_t nxf = 0, nyf = 0;
    int npts = 0;                       // cell points
    // bnd: [5][N] per-cell boundary flux H, C, heat source, then rate and divergence
    DeviceBuffer d_md, state, old, rbuf, first, bnd, blend, cell_out, face, flowx, flowy, cmix;
    DeviceBuffer bcH, bcC, inflow, tbc, tval, error, red, kinds, part;
    DeviceBuffer tep;                   // lithostatic {Tep, Tmp} tables
    bool have_flow = false;
    double c_max = 0;                   // admissible C bound
    // Read-back of one step attempt, in pinned host memory: the totals in red
    // (offsets below) and the error words of the RHS evaluations (slots).
    double* host_red = nullptr;         // [kRedSize]
    uint32_t* host_err = nullptr;       // [kErrSlots]
    static constexpr int kRedRhs0 = kCoupledRedRhs0, kRedAdm0 = kCoupledRedAdm0, kRedAdm1 = kCoupledRedAdm1,
                         kRedRhs1 = kCoupledRedRhs1, kRedAdm2 = kCoupledRedAdm2, kRedRhs2 = kCoupledRedRhs2,
                         kRedSize = kCoupledRedSize;
    static constexpr int kErrSlots = kCoupledErrSlots;

    ~mp_coupled()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
        if (host_red) (void)hipHostFree(host_red);
        if (host_err) (void)hipHostFree(host_err);
    }

    mp_status init(const mp_mesh_desc& m, const mp_phase_params& p, const mp_coupled_params& cp,
                   void* s);
    mp_status sync_md();
    mp_status set_flow(const double* es, const double* ed, const double* ep, const double* cs,
                       const double* cdarcy, const double* cp);
    mp_status set_boundary(int field, const uint8_t* const side[4]);
    mp_status set_inflow(const double* h, const double* c);
    mp_status set_thermal(const uint8_t* types, const double* values);
    mp_status set_lithostatic(int enable, double surface_y);
    mp_status enqueue_rhs(const double* H, const double* C, double* rH, double* rC, int slot,
                          int off);
    mp_status enqueue_admissible(const double* S, int off);
    mp_status read_back();
    mp_status rhs_result(int slot, int off, mp_coupled_rates& tot) const;
    void admissible_result(int off, bool& ok, double& intH, double& intC) const;
    mp_status rhs_on(const double* H, const double* C, double* rH, double* rC,
                     mp_coupled_rates& tot);
    mp_status rhs(double* rH, double* rC, double* rate, double* div, double* T, double* phi,
                  mp_coupled_rates* tot);
    mp_status run(const mp_coupled_time& tc, double* time, int nsteps, mp_coupled_step* steps,
                  int32_t* accepted);
    CoupledBatchMesh batch_record() const;
};

mp_status mp_coupled::sync_md()
{
    MP_HIP(hipMemcpyAsync(d_md.ptr, &md, sizeof(MeshDev), hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status mp_coupled::init(const mp_mesh_desc& m, const mp_phase_params& p,
                           const mp_coupled_params& cp, void* s)
{
    if (!(std::isfinite(cp.theta) && std::isfinite(cp.cooling)))
        return fail(MP_ERR_ARG, "theta and cooling must be finite");
    if (!(std::isfinite(cp.thermal_diffusivity) && cp.thermal_diffusivity >= 0))
        return fail(MP_ERR_ARG, "diffusivity must be nonnegative");
    if (cp.lf_mode != 0 && cp.lf_mode != 1) return fail(MP_ERR_ARG, "unknown Lax-Friedrichs mode");
    if (cp.lf_mode == 1 && !(std::isfinite(cp.global_speed) && cp.global_speed >= 0))
        return fail(MP_ERR_ARG, "global LF speed must be finite and nonnegative");
    MP_CALL(make_phase(p, phase));
    mp_weno_options w;
    mp_default_weno_options(&w);
    w.lf_mode = cp.lf_mode;
    w.global_speed = cp.global_speed;
    MP_CALL(make_mesh_dev(m, w, phase, md));
    c_max = p.hypereutectic ? 1.0 : p.Xe;
    MP_CALL(set_limiter_tables(md, cp.cell_points_per_axis, cp.diffusion_samples,
                               cp.extent_fraction, c_max));
    mesh = m;
    phase_params = p;
    params = cp;
    N = size_t(m.nx) * m.ny;
    nxf = int64_t(m.ny) * (m.nx + 1);
    nyf = int64_t(m.ny + 1) * m.nx;
    npts = cp.cell_points_per_axis * cp.cell_points_per_axis;
    if (s) {
        stream = static_cast<hipStream_t>(s);
    } else {
        MP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        own_stream = true;
    }
    MP_CALL(d_md.alloc(sizeof(MeshDev)));
    MP_CALL(state.alloc(2 * N * sizeof(double)));
    MP_CALL(old.alloc(2 * N * sizeof(double)));
    MP_CALL(rbuf.alloc(2 * N * sizeof(double)));
    MP_CALL(first.alloc(2 * N * sizeof(double)));
    MP_CALL(bnd.alloc(5 * N * sizeof(double)));
    MP_CALL(blend.alloc(20 * N * sizeof(double)));
    MP_CALL(cell_out.alloc(4 * N * sizeof(double)));
    MP_CALL(face.alloc(4 * size_t(nxf + nyf) * sizeof(double)));
    MP_CALL(flowx.alloc(3 * size_t(m.q) * nxf * sizeof(double)));
    MP_CALL(flowy.alloc(3 * size_t(m.q) * nyf * sizeof(double)));
    MP_CALL(cmix.alloc(size_t(npts) * N * sizeof(double)));
    MP_CALL(error.alloc(kErrSlots * sizeof(uint32_t)));
    MP_CALL(red.alloc(kRedSize * sizeof(double)));
    MP_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_red), kRedSize * sizeof(double), hipHostMallocDefault));
    MP_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_err), kErrSlots * sizeof(uint32_t), hipHostMallocDefault));
    // reduction kinds and array maps, fixed: the RHS totals (sums of the three
    // boundary/source fields, max rate, max |divergence|) and the state checks
    // (min, max and sum of H and C)
    static const int kKinds[17] = {0, 0, 0, 3, 1,               // RHS kinds, arrays 0..4
                                   2, 2, 3, 3, 0, 0,            // state kinds
                                   0, 1, 0, 1, 0, 1};           // state arrays: H, C, H, C, H, C
    MP_CALL(kinds.alloc(sizeof kKinds));
    MP_HIP(hipMemcpyAsync(kinds.ptr, kKinds, sizeof kKinds, hipMemcpyHostToDevice, stream));
    MP_CALL(part.alloc(2 * 6 * size_t(kReducePartials) * sizeof(double)));
    MP_HIP(hipMemsetAsync(state.ptr, 0, state.bytes, stream));
    MP_HIP(hipMemsetAsync(flowx.ptr, 0, flowx.bytes, stream));
    MP_HIP(hipMemsetAsync(flowy.ptr, 0, flowy.bytes, stream));
    MP_HIP(hipMemsetAsync(cmix.ptr, 0, cmix.bytes, stream));

    CoupledDev& d = cd;
    std::memset(&d, 0, sizeof d);
    d.m = d_md.as<MeshDev>();
    d.blend = blend.as<double>();
    d.cell_out = cell_out.as<double>();
    d.face = face.as<double>();
    d.error = error.as<uint32_t>();
    const size_t q = m.q;
    d.xvs = flowx.as<double>();
    d.xqd = d.xvs + q * nxf;
    d.xphi = d.xqd + q * nxf;
    d.yvs = flowy.as<double>();
    d.yqd = d.yvs + q * nyf;
    d.yphi = d.yqd + q * nyf;
    d.cmix = cmix.as<double>();
    d.cooling = cp.cooling;
    d.kappa = cp.thermal_diffusivity;
    d.global_speed = cp.global_speed;
    d.lf_mode = cp.lf_mode;
    // normal-line sampling on uniform rectangles (src/mlweno/diffusiveflux.cpp:256-338)
    const int n = cp.diffusion_samples;
    d.n = n;
    const long double f = cp.extent_fraction;
    d.dsx = (double)(f * md.dx * 2.0L / (n - 1));
    d.dsy = (double)(f * md.dy * 2.0L / (n - 1));
    d.dsbx = (double)(f * md.dx / n);
    d.dsby = (double)(f * md.dy / n);
    lagrange_weights(n, (n - 1) / 2.0L, d.cm);
    lagrange_weights(n + 1, (long double)n, d.ce);
    long double last = 0, am = 0, ae = 0;
    for (int k = 0; k + 1 < n; ++k) {
        last -= d.cm[k];
        am += std::fabs(d.cm[k]);
    }
    am += fabsl(last);
    for (int k = 0; k < n; ++k) ae += std::fabs(d.ce[k]);
    d.cm_abs = (double)am;
    d.ce_abs = (double)ae;
    return sync_md();
}

// Flow samples by reference edge id (horizontal edges first) and canonical q;
// only the normal components of a uniform rectangle's faces are kept, with
// the Darcy flux phi^(1+theta)*u_d formed as PhaseTransportVelocities does
// (src/couple/evolve.cpp:40-46).
mp_status mp_coupled::set_flow(const double* es, const double* ed, const double* ep,
                               const double* cs, const double* cdarcy, const double* cp)
{
    if (!es || !ed || !ep || !cs || !cdarcy || !cp) return fail(MP_ERR_ARG, "null flow sample array");
    const int nx = mesh.nx, ny = mesh.ny, q = mesh.q;
    const size_t nh = size_t(nx) * (ny + 1);
    std::vector<double> hx(3 * size_t(q) * nxf), hy(3 * size_t(q) * nyf);
    auto factor = [&](double phi) { return phi == 0 ? 0.0 : std::pow(phi, 1 + params.theta); };
    for (size_t e = 0; e < nh + size_t(nxf); ++e) {
        const bool vert = e >= nh;
        const int comp = vert ? 0 : 1;
        for (int k = 0; k < q; ++k) {
            const size_t a = e * q + k;
            const double phi = ep[a];
            if (!(std::isfinite(phi) && phi >= 0 && phi <= 1))
                return fail(MP_ERR_ARG, "flow porosity must lie in [0, 1]");
            const double vs = es[2 * a + comp], dn = ed[2 * a + comp];
            if (!std::isfinite(vs) || !std::isfinite(dn)) return fail(MP_ERR_FP, "flow samples must be finite");
            const double qd = factor(phi) * dn;
            if (vert) {
                const size_t fl = e - nh;   // j*(nx+1) + i
                hx[size_t(k) * nxf + fl] = vs;
                hx[(q + k) * size_t(nxf) + fl] = qd;
                hx[(2 * q + k) * size_t(nxf) + fl] = phi;
            } else {
                hy[size_t(k) * nyf + e] = vs;
                hy[(q + k) * size_t(nyf) + e] = qd;
                hy[(2 * q + k) * size_t(nyf) + e] = phi;
            }
        }
    }
    std::vector<double> mix(size_t(npts) * N);
    for (size_t c = 0; c < N; ++c)
        for (int k = 0; k < npts; ++k) {
            const size_t a = c * npts + k;
            const double phi = cp[a];
            if (!(std::isfinite(phi) && phi >= 0 && phi <= 1))
                return fail(MP_ERR_ARG, "flow porosity must lie in [0, 1]");
            const double m = cs[2 * a + 1] + factor(phi) * cdarcy[2 * a + 1];
            if (!std::isfinite(m)) return fail(MP_ERR_FP, "flow samples must be finite");
            mix[size_t(k) * N + c] = m;
        }
    MP_HIP(hipMemcpyAsync(flowx.ptr, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(flowy.ptr, hy.data(), hy.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(cmix.ptr, mix.data(), mix.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    have_flow = true;
    return MP_OK;
}

mp_status mp_coupled::set_boundary(int field, const uint8_t* const side[4])
{
    if (field != 0 && field != 1) return fail(MP_ERR_ARG, "field must be 0 (H) or 1 (C)");
    const size_t n[4] = {size_t(mesh.nx), size_t(mesh.ny), size_t(mesh.nx), size_t(mesh.ny)};
    DeviceBuffer& buf = field ? bcC : bcH;
    const uint8_t** dst = field ? cd.bcC : cd.bcH;
    std::vector<uint8_t> all;
    for (int s = 0; s < 4; ++s) {
        for (size_t k = 0; k < n[s]; ++k) {
            const uint8_t v = side[s] ? side[s][k] : 0;
            if (v > 2) return fail(MP_ERR_ARG, "unknown boundary policy %d", int(v));
            all.push_back(v);
        }
    }
    MP_CALL(buf.alloc(all.size()));
    MP_HIP(hipMemcpyAsync(buf.ptr, all.data(), all.size(), hipMemcpyHostToDevice, stream));
    size_t off = 0;
    for (int s = 0; s < 4; ++s) {
        dst[s] = static_cast<const uint8_t*>(buf.ptr) + off;
        off += n[s];
    }
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status mp_coupled::set_inflow(const double* h, const double* c)
{
    const size_t total = size_t(2) * (mesh.nx + mesh.ny) * mesh.q;
    const size_t off[4] = {0, size_t(mesh.nx) * mesh.q, size_t(mesh.nx + mesh.ny) * mesh.q,
                           size_t(2 * mesh.nx + mesh.ny) * mesh.q};
    for (const double* src : {h, c})
        if (src && !all_finite(src, total))
            return fail(MP_ERR_FP, "Inflow boundary callback must return a finite state");
    MP_CALL(inflow.alloc(2 * total * sizeof(double)));
    double* base = inflow.as<double>();
    for (int f = 0; f < 2; ++f) {
        const double* src = f ? c : h;
        const double** dst = f ? cd.inC : cd.inH;
        if (src) MP_HIP(hipMemcpyAsync(base + f * total, src, total * sizeof(double), hipMemcpyHostToDevice, stream));
        for (int s = 0; s < 4; ++s) dst[s] = src ? base + f * total + off[s] : nullptr;
    }
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status mp_coupled::set_thermal(const uint8_t* types, const double* values)
{
    const size_t total = size_t(2) * (mesh.nx + mesh.ny) * mesh.q;
    const size_t off[4] = {0, size_t(mesh.nx) * mesh.q, size_t(mesh.nx + mesh.ny) * mesh.q,
                           size_t(2 * mesh.nx + mesh.ny) * mesh.q};
    if (!types) {
        for (int s = 0; s < 4; ++s) { md.tbc[s] = nullptr; md.tval[s] = nullptr; }
        return sync_md();
    }
    std::vector<double> v(total, 0.0);
    for (size_t k = 0; k < total; ++k) {
        if (types[k] > 2) return fail(MP_ERR_ARG, "Unknown thermal boundary policy");
        if (types[k] != 0) {
            if (!values) return fail(MP_ERR_ARG, "Missing thermal boundary profile.");
            if (!std::isfinite(values[k])) return fail(MP_ERR_ARG, "Thermal boundary data must be finite");
            v[k] = values[k];
        }
    }
    MP_CALL(tbc.alloc(total));
    MP_CALL(tval.alloc(total * sizeof(double)));
    MP_HIP(hipMemcpyAsync(tbc.ptr, types, total, hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(tval.ptr, v.data(), total * sizeof(double), hipMemcpyHostToDevice, stream));
    for (int s = 0; s < 4; ++s) {
        md.tbc[s] = static_cast<const uint8_t*>(tbc.ptr) + off[s];
        md.tval[s] = tval.as<double>() + off[s];
    }
    return sync_md();
}

// Lithostatic pressure at every phase evaluation point
// (src/couple/initialization.cpp:155-162; c.pressure(p, t) at the point p of
// src/couple/evolve.cpp:189, 273, 318).  Point ordinates follow the reference's
// construction on the uniform vertices: MapCellPoint for cell points,
// MapEdgePoint for face points (src/core/integral.cpp:91-106), and
// ShiftPoint along the unit normal by offset*spacing for the normal-line
// samples (src/mlweno/diffusiveflux.cpp:76-86, 256-338) -- the samples of a
// vertical face keep their point's ordinate.
mp_status mp_coupled::set_lithostatic(int enable, double surface_y)
{
    if (!enable) {
        tep.reset();
        cd.tep_cell = cd.tep_v = cd.tep_h = cd.tep_hs = nullptr;
        return MP_OK;
    }
    if (!std::isfinite(surface_y)) return fail(MP_ERR_ARG, "surface_y must be finite");
    if (surface_y < mesh.y1) return fail(MP_ERR_ARG, "surface_y must be at or above the mesh top");
    const int ny = mesh.ny, nq = mesh.q, n = params.cell_points_per_axis;
    const int ns = params.diffusion_samples;
    std::vector<double> y(size_t(ny) + 1);
    uniform_vertices(mesh.y0, mesh.y1, ny, y.data());
    double xc[kMaxQ], wc[kMaxQ], xq[kMaxQ], wq[kMaxQ];
    gauss_legendre(n, xc, wc);
    gauss_legendre(nq, xq, wq);
    const size_t ncell = size_t(ny) * n * n, nv = size_t(ny) * nq, nh = size_t(ny + 1) * nq;
    std::vector<TepPair> tab(ncell + nv + nh + nh * ns);
    TepPair *tc = tab.data(), *tv = tc + ncell, *th = tv + nv, *ths = th + nh;
    for (int j = 0; j < ny; ++j)
        for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
                const double lower = map_edge(xc[a], y[j], y[j]);
                const double upper = map_edge(xc[a], y[j + 1], y[j + 1]);
                MP_CALL(lithostatic_pair(phase, surface_y, map_edge(xc[b], lower, upper),
                                         tc[(size_t(j) * n + b) * n + a]));
            }
    for (int j = 0; j < ny; ++j)
        for (int q = 0; q < nq; ++q)
            MP_CALL(lithostatic_pair(phase, surface_y, map_edge(xq[q], y[j], y[j + 1]),
                                     tv[size_t(j) * nq + q]));
    const long double f = params.extent_fraction;
    for (int j = 0; j <= ny; ++j)
        for (int q = 0; q < nq; ++q) {
            const double py = map_edge(xq[q], y[j], y[j]);
            const size_t fq = size_t(j) * nq + q;
            MP_CALL(lithostatic_pair(phase, surface_y, py, th[fq]));
            // available normal distance: to the far face of the cell(s) beside the face
            const bool bottom = j == 0, top = j == ny;
            double avail = bottom ? y[1] - py : py - y[j - 1];
            if (!bottom && !top) avail = std::min(avail, y[j + 1] - py);
            const double spacing =
                (double)(f * avail * ((bottom || top) ? 1.0L / ns : 2.0L / (ns - 1)));
            const long double normal = bottom ? -1.0L : 1.0L;      // outward, or lower -> upper
            for (int k = 0; k < ns; ++k) {
                const long double offset = (bottom || top) ? (long double)k - ns
                                                           : (long double)k - (ns - 1) / 2.0L;
                const double ys = (double)((long double)py + offset * spacing * normal);
                MP_CALL(lithostatic_pair(phase, surface_y, ys, ths[fq * ns + k]));
            }
        }
    MP_CALL(tep.alloc(tab.size() * sizeof(TepPair)));
    MP_HIP(hipMemcpyAsync(tep.ptr, tab.data(), tep.bytes, hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    TepPair* base = tep.as<TepPair>();
    cd.tep_cell = base;
    cd.tep_v = base + ncell;
    cd.tep_h = base + ncell + nv;
    cd.tep_hs = base + ncell + nv + nh;
    return MP_OK;
}

// One RHS of the state (H, C) into (rH, rC), its error bits into error word
// `slot` and its totals (5 doubles) into red + off; nothing is waited for.
mp_status mp_coupled::enqueue_rhs(const double* H, const double* C, double* rH, double* rC,
                                  int slot, int off)
{
    if (!have_flow) return fail(MP_ERR_STATE, "set the flow samples first");
    MP_HIP(hipMemsetAsync(error.as<uint32_t>() + slot, 0, sizeof(uint32_t), stream));
    CoupledDev d = cd;
    d.error = error.as<uint32_t>() + slot;
    double* b = bnd.as<double>();
    MP_HIP(launch_coupled_rhs(d, mesh.nx, mesh.ny, H, C, rH, rC, b + 3 * N, b + 4 * N, b, stream));
    MP_HIP(launch_reduce(b, int64_t(N), 5, kinds.as<int>(), nullptr, red.as<double>() + off,
                         part.as<double>(), stream));
    return MP_OK;
}

// The state checks of Admissible (src/couple/evolve.cpp:386-393) and the
// integrals sum |K| u: min H, min C | max H, max C | sum H, sum C (H and C
// are contiguous) in one pass, into red + off.
mp_status mp_coupled::enqueue_admissible(const double* S, int off)
{
    MP_HIP(launch_reduce(S, int64_t(N), 6, kinds.as<int>() + 5, kinds.as<int>() + 11,
                         red.as<double>() + off, part.as<double>(), stream));
    return MP_OK;
}

// Copies every total and error word to the host and waits for the stream.
mp_status mp_coupled::read_back()
{
    MP_HIP(hipMemcpyAsync(host_red, red.ptr, kRedSize * sizeof(double), hipMemcpyDeviceToHost, stream));
    MP_HIP(hipMemcpyAsync(host_err, error.ptr, kErrSlots * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

// Totals of an RHS after read_back(); fails with its device error bits.
mp_status mp_coupled::rhs_result(int slot, int off, mp_coupled_rates& tot) const
{
    const uint32_t bits = host_err[slot];
    const double* h = host_red + off;
    if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
    if (!std::isfinite(h[3])) return fail(MP_ERR_FP, "Nonfinite coupled stage rate");
    tot.boundary_H = h[0];
    tot.boundary_C = h[1];
    tot.source_H = h[2];
    tot.maximum_rate = h[3];
    tot.divergence = h[4];
    return MP_OK;
}

// Admissible and the integrals after read_back().
void mp_coupled::admissible_result(int off, bool& ok, double& intH, double& intC) const
{
    const double* h = host_red + off;
    ok = std::isfinite(h[0]) && std::isfinite(h[2]) && std::isfinite(h[1]) && std::isfinite(h[3]) &&
         h[0] >= 0 && h[1] >= 0 && h[3] <= c_max;
    intH = h[4] * md.area;
    intC = h[5] * md.area;
}

// One RHS of the state (H, C) into (rH, rC) and the totals; fails with the
// device's error bits.
mp_status mp_coupled::rhs_on(const double* H, const double* C, double* rH, double* rC,
                             mp_coupled_rates& tot)
{
    MP_CALL(enqueue_rhs(H, C, rH, rC, 0, kRedRhs0));
    MP_CALL(read_back());
    return rhs_result(0, kRedRhs0, tot);
}

mp_status mp_coupled::rhs(double* rH, double* rC, double* rate, double* div, double* T, double* phi,
                          mp_coupled_rates* tot)
{
    mp_coupled_rates t{};
    double* S = state.as<double>();
    double* R = rbuf.as<double>();
    MP_CALL(rhs_on(S, S + N, R, R + N, t));
    const size_t nb = N * sizeof(double);
    double* outs[6] = {rH, rC, rate, div, T, phi};
    const double* srcs[6] = {R, R + N, bnd.as<double>() + 3 * N, bnd.as<double>() + 4 * N,
                             cell_out.as<double>() + N, cell_out.as<double>() + 2 * N};
    for (int k = 0; k < 6; ++k)
        if (outs[k]) MP_HIP(hipMemcpyAsync(outs[k], srcs[k], nb, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    if (tot) *tot = t;
    return MP_OK;
}

// The step loop of src/couple/evolve.cpp:428-511.  One attempt at a step size
// is enqueued whole -- stage 1, its state check, the second RHS, stage 2, its
// state check and the RHS of the new state -- and read back with one wait; the
// host then takes the reference's decisions in the reference's order and
// ignores whatever the attempt computed past the first failed check (a
// rejected attempt's later results, error bits included, are discarded, and a
// retry restarts from the saved state and first RHS).  The state checks of an
// accepted step serve as the next step's starting integrals, which the
// reference recomputes from the same state.
mp_status mp_coupled::run(const mp_coupled_time& tc, double* time, int nsteps,
                          mp_coupled_step* steps, int32_t* accepted_out)
{
    if (!(tc.cfl > 0 && tc.cfl <= 1)) return fail(MP_ERR_ARG, "Coupled CFL must be in (0,1]");
    if (!time) return fail(MP_ERR_ARG, "null time");
    const size_t nb = 2 * N * sizeof(double);
    double* S = state.as<double>();
    double* O = old.as<double>();
    double* R = rbuf.as<double>();
    double* F = first.as<double>();
    mp_coupled_rates current{}, second{};
    MP_CALL(enqueue_rhs(S, S + N, R, R + N, 0, kRedRhs0));
    MP_CALL(enqueue_admissible(S, kRedAdm0));
    MP_HIP(hipMemcpyAsync(O, S, nb, hipMemcpyDeviceToDevice, stream));
    MP_CALL(read_back());
    MP_CALL(rhs_result(0, kRedRhs0, current));
    bool ok;
    double oldH, oldC;
    admissible_result(kRedAdm0, ok, oldH, oldC);
    int accepted = 0;
    mp_status st = MP_OK;
    for (int s = 0; s < nsteps; ++s) {
        const double remaining = tc.end_time - *time;
        if (remaining <= 32 * 2.220446049250313e-16 * std::fmax(1.0, std::fabs(tc.end_time))) {
            *time = tc.end_time;
            break;
        }
        const double start = *time;
        double dt = std::fmin(remaining, tc.fixed ? tc.initial_step
                                                  : (accepted == 0 ? tc.initial_step : tc.maximum_step));
        const double limit = current.maximum_rate > 0 ? tc.cfl / current.maximum_rate : tc.maximum_step;
        if (tc.fixed && !(dt <= limit * (1 + 1e-12)))
            return fail(MP_ERR_ARG, "Fixed time step exceeds the phase-coupled CFL bound; use control: cfl");
        if (!tc.fixed) dt = std::fmin(dt, limit);
        MP_HIP(hipMemcpyAsync(F, R, nb, hipMemcpyDeviceToDevice, stream));
        const mp_coupled_rates firstRates = current;
        int retries = 0;
        double courant = 0, newH = 0, newC = 0;
        for (;;) {
            if (!(dt >= std::fmin(remaining, tc.minimum_step) && retries < 40)) {
                st = fail(MP_ERR_PHASE, "Coupled step cannot satisfy admissibility/CFL above minimum step");
                break;
            }
            MP_HIP(launch_coupled_stage(int64_t(N), 1, dt, O, O + N, F, F + N, S, S + N, stream));
            MP_CALL(enqueue_admissible(S, kRedAdm1));
            MP_CALL(enqueue_rhs(S, S + N, R, R + N, 1, kRedRhs1));
            MP_HIP(launch_coupled_stage(int64_t(N), 2, dt, O, O + N, R, R + N, S, S + N, stream));
            MP_CALL(enqueue_admissible(S, kRedAdm2));
            MP_CALL(enqueue_rhs(S, S + N, R, R + N, 2, kRedRhs2));
            MP_CALL(read_back());
            *time = start + dt;
            double ih, ic;
            admissible_result(kRedAdm1, ok, ih, ic);
            if (ok) {
                st = rhs_result(1, kRedRhs1, second);
                if (st) break;
                courant = dt * std::fmax(current.maximum_rate, second.maximum_rate);
                ok = courant <= tc.cfl * (1 + 1e-12);
            }
            if (ok) admissible_result(kRedAdm2, ok, newH, newC);
            if (ok) break;
            if (tc.fixed) {
                st = fail(MP_ERR_ARG, "Fixed coupled step failed stage admissibility/CFL");
                break;
            }
            dt *= .5;
            ++retries;
        }
        double balH = 0, balC = 0;
        if (!st) {
            balH = newH - oldH + .5 * dt * (current.boundary_H + second.boundary_H - current.source_H - second.source_H);
            balC = newC - oldC + .5 * dt * (current.boundary_C + second.boundary_C);
            if (!(std::fabs(balH) < 1e-10 * std::fmax(1.0, std::fabs(oldH)) && std::fabs(balC) < 1e-10))
                st = fail(MP_ERR_PHASE, "Coupled finite-volume conservation balance failed");
        }
        if (!st) st = rhs_result(2, kRedRhs2, current);
        if (st) {
            // roll back to the latest accepted state (src/couple/evolve.cpp:493-499)
            *time = start;
            MP_HIP(hipMemcpyAsync(S, O, nb, hipMemcpyDeviceToDevice, stream));
            MP_HIP(hipStreamSynchronize(stream));
            break;
        }
        ++accepted;
        oldH = newH;
        oldC = newC;
        MP_HIP(hipMemcpyAsync(O, S, nb, hipMemcpyDeviceToDevice, stream));
        if (steps) {
            mp_coupled_step& r = steps[s];
            r.dt = dt;
            r.courant = courant;
            r.balance_H = balH;
            r.balance_C = balC;
            r.boundary_H = .5 * (firstRates.boundary_H + second.boundary_H);
            r.boundary_C = .5 * (firstRates.boundary_C + second.boundary_C);
            r.source_H = .5 * (firstRates.source_H + second.source_H);
            r.maximum_rate = current.maximum_rate;
            r.retries = retries;
            r.pad = 0;
        }
    }
    MP_HIP(hipStreamSynchronize(stream));
    if (accepted_out) *accepted_out = accepted;
    return st;
}

CoupledBatchMesh mp_coupled::batch_record() const
{
    CoupledBatchMesh b;
    std::memset(&b, 0, sizeof b);
    b.d = cd;
    b.S = state.as<double>();
    b.O = old.as<double>();
    b.R = rbuf.as<double>();
    b.F = first.as<double>();
    b.bnd = bnd.as<double>();
    b.red = red.as<double>();
    b.part = part.as<double>();
    b.kinds = kinds.as<int>();
    b.N = int64_t(N);
    b.nf = nxf + nyf;
    b.nb = coupled_reduce_blocks(int64_t(N));
    return b;
}

namespace {

// Pinned host memory of a batched run, freed on every exit.
struct PinnedBlock {
    void* p = nullptr;
    ~PinnedBlock() { if (p) (void)hipHostFree(p); }
};

// One handle's place in the step loop of mp_coupled::run (same variables,
// same decisions), advanced one step attempt per round.
struct BatchMeshState {
    enum { kRunning, kDone, kFailed } phase = kRunning;
    bool new_step = true;          // the next attempt starts step s
    int s = 0, accepted = 0, retries = 0;
    double start = 0, dt = 0, remaining = 0, courant = 0;
    double oldH = 0, oldC = 0, newH = 0, newC = 0;
    mp_coupled_rates current{}, second{}, first{};
    mp_status st = MP_OK;
    std::string msg;
};

}  // namespace

// Several handles advanced together.  Each round enqueues, for every handle
// still running, the copies its loop needs and one step attempt (stage 1, its
// state check, the second RHS, stage 2, its state check, the RHS of the new
// state) as one launch per kernel over all those handles, then reads every
// handle's totals and error words back with one wait; each handle then takes
// the decisions of mp_coupled::run in the same order.  The kernels run the
// single-mesh bodies and reductions on each handle's own buffers, so every
// handle ends as its own run would: states, times and step records.
static mp_status coupled_run_batch(mp_coupled* const* hs, int n, const mp_coupled_time* tcs, double* times,
                                   int nsteps, mp_coupled_step* steps, int32_t* accepted_out,
                                   mp_status* status_out)
{
    if (n <= 0 || !hs || !tcs || !times) return fail(MP_ERR_ARG, "batched coupled run needs handles, time controls and times");
    if (nsteps < 0) return fail(MP_ERR_ARG, "nsteps must be nonnegative");
    for (int k = 0; k < n; ++k) {
        if (!hs[k]) return fail(MP_ERR_ARG, "null coupled handle %d", k);
        for (int j = 0; j < k; ++j)
            if (hs[j] == hs[k]) return fail(MP_ERR_ARG, "coupled handle %d appears twice", k);
    }
    hipStream_t stream = hs[0]->stream;
    for (int k = 0; k < n; ++k) MP_HIP(hipStreamSynchronize(hs[k]->stream));
    std::vector<BatchMeshState> ms(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        const mp_coupled_time& tc = tcs[k];
        if (!(tc.cfl > 0 && tc.cfl <= 1)) ms[k].st = fail(MP_ERR_ARG, "Coupled CFL must be in (0,1]");
        else if (!hs[k]->have_flow) ms[k].st = fail(MP_ERR_STATE, "set the flow samples first");
        if (ms[k].st) {
            ms[k].phase = BatchMeshState::kFailed;
            ms[k].msg = last_error();
        }
    }

    // Device records, the per-round tables and the read-back, sized for n.
    // ints: act[n], 5 bases [n+1] (cell, face, gather, reduction, stage),
    // copy segments [2][2n] and their base [2n+1]; doubles: dt[n].
    const size_t n1 = size_t(n) + 1;
    const size_t nints = size_t(n) + 5 * n1 + 4 * size_t(n) + 2 * size_t(n) + 1;
    DeviceBuffer d_rec, d_int, d_dbl, d_rb_red, d_rb_err;
    MP_CALL(d_rec.alloc(sizeof(CoupledBatchMesh) * size_t(n)));
    MP_CALL(d_int.alloc(sizeof(int) * nints));
    MP_CALL(d_dbl.alloc(sizeof(double) * size_t(n)));
    MP_CALL(d_rb_red.alloc(sizeof(double) * kCoupledRedSize * size_t(n)));
    MP_CALL(d_rb_err.alloc(sizeof(uint32_t) * kCoupledErrSlots * size_t(n)));
    PinnedBlock h_rec, h_int, h_dbl, h_rb_red, h_rb_err;
    MP_HIP(hipHostMalloc(&h_rec.p, sizeof(CoupledBatchMesh) * size_t(n), hipHostMallocDefault));
    MP_HIP(hipHostMalloc(&h_int.p, sizeof(int) * nints, hipHostMallocDefault));
    MP_HIP(hipHostMalloc(&h_dbl.p, sizeof(double) * size_t(n), hipHostMallocDefault));
    MP_HIP(hipHostMalloc(&h_rb_red.p, sizeof(double) * kCoupledRedSize * size_t(n), hipHostMallocDefault));
    MP_HIP(hipHostMalloc(&h_rb_err.p, sizeof(uint32_t) * kCoupledErrSlots * size_t(n), hipHostMallocDefault));
    auto* rec = static_cast<CoupledBatchMesh*>(h_rec.p);
    for (int k = 0; k < n; ++k) rec[k] = hs[k]->batch_record();
    MP_HIP(hipMemcpyAsync(d_rec.ptr, rec, sizeof(CoupledBatchMesh) * size_t(n), hipMemcpyHostToDevice, stream));
    const CoupledBatchMesh* M = d_rec.as<CoupledBatchMesh>();
    int* hi = static_cast<int*>(h_int.p);
    int* di = d_int.as<int>();
    double* hd = static_cast<double*>(h_dbl.p);
    int* h_act = hi;
    int* h_base[5];
    for (int t = 0; t < 5; ++t) h_base[t] = hi + n + t * n1;
    int* h_seg = hi + n + 5 * n1;
    int* h_segbase = h_seg + 4 * size_t(n);
    auto dev = [&](const int* h) { return di + (h - hi); };
    std::vector<int> seg_pending;  // (handle, kind) pairs owed before the next attempt

    // One round: the owed copies, then (if act non-empty) one attempt, or
    // in the first round the starting RHS and state check, then the read-back.
    auto round = [&](const std::vector<int>& act, bool first_round) -> mp_status {
        const int na = int(act.size());
        int tot[5] = {0, 0, 0, 0, 0};
        for (int a = 0; a < na; ++a) {
            const CoupledBatchMesh& r = rec[act[a]];
            h_act[a] = act[a];
            const int blocks[5] = {coupled_cell_blocks(r.N), coupled_face_blocks(r.nf), coupled_gather_blocks(r.N),
                                   r.nb, coupled_stage_blocks(r.N)};
            for (int t = 0; t < 5; ++t) {
                h_base[t][a] = tot[t];
                tot[t] += blocks[t];
            }
            hd[a] = ms[act[a]].dt;
        }
        for (int t = 0; t < 5; ++t) h_base[t][na] = tot[t];
        const int nseg = int(seg_pending.size() / 2);
        int copy_blocks = 0;
        for (int q = 0; q < nseg; ++q) {
            h_seg[2 * q] = seg_pending[2 * q];
            h_seg[2 * q + 1] = seg_pending[2 * q + 1];
            h_segbase[q] = copy_blocks;
            copy_blocks += coupled_copy_blocks(rec[seg_pending[2 * q]].N);
        }
        h_segbase[nseg] = copy_blocks;
        seg_pending.clear();
        MP_HIP(hipMemcpyAsync(di, hi, sizeof(int) * nints, hipMemcpyHostToDevice, stream));
        MP_HIP(hipMemcpyAsync(d_dbl.ptr, hd, sizeof(double) * size_t(n), hipMemcpyHostToDevice, stream));
        MP_HIP(launch_coupled_batch_copy(M, dev(h_seg), dev(h_segbase), nseg, copy_blocks, stream));
        if (na == 0) {
            MP_HIP(hipStreamSynchronize(stream));
            return MP_OK;
        }
        const CoupledBatchSpan cell{dev(h_act), dev(h_base[0]), na}, face{dev(h_act), dev(h_base[1]), na},
            gath{dev(h_act), dev(h_base[2]), na}, red{dev(h_act), dev(h_base[3]), na},
            stg{dev(h_act), dev(h_base[4]), na};
        const double* ddt = d_dbl.as<double>();
        if (first_round) {
            MP_HIP(launch_coupled_batch_clear(M, dev(h_act), na, 0, 1, stream));
            MP_HIP(launch_coupled_batch_rhs(M, cell, tot[0], face, tot[1], gath, tot[2], red, tot[3], 0,
                                            kCoupledRedRhs0, stream));
            MP_HIP(launch_coupled_batch_admissible(M, red, tot[3], kCoupledRedAdm0, stream));
        } else {
            MP_HIP(launch_coupled_batch_clear(M, dev(h_act), na, 1, 2, stream));
            MP_HIP(launch_coupled_batch_stage(M, stg, tot[4], 1, ddt, stream));
            MP_HIP(launch_coupled_batch_admissible(M, red, tot[3], kCoupledRedAdm1, stream));
            MP_HIP(launch_coupled_batch_rhs(M, cell, tot[0], face, tot[1], gath, tot[2], red, tot[3], 1,
                                            kCoupledRedRhs1, stream));
            MP_HIP(launch_coupled_batch_stage(M, stg, tot[4], 2, ddt, stream));
            MP_HIP(launch_coupled_batch_admissible(M, red, tot[3], kCoupledRedAdm2, stream));
            MP_HIP(launch_coupled_batch_rhs(M, cell, tot[0], face, tot[1], gath, tot[2], red, tot[3], 2,
                                            kCoupledRedRhs2, stream));
        }
        MP_HIP(launch_coupled_batch_readback(M, dev(h_act), na, d_rb_red.as<double>(), d_rb_err.as<uint32_t>(),
                                             stream));
        MP_HIP(hipMemcpyAsync(h_rb_red.p, d_rb_red.ptr, sizeof(double) * kCoupledRedSize * size_t(na),
                              hipMemcpyDeviceToHost, stream));
        MP_HIP(hipMemcpyAsync(h_rb_err.p, d_rb_err.ptr, sizeof(uint32_t) * kCoupledErrSlots * size_t(na),
                              hipMemcpyDeviceToHost, stream));
        MP_HIP(hipStreamSynchronize(stream));
        for (int a = 0; a < na; ++a) {
            mp_coupled& h = *hs[act[a]];
            std::memcpy(h.host_red, static_cast<double*>(h_rb_red.p) + size_t(a) * kCoupledRedSize,
                        sizeof(double) * kCoupledRedSize);
            std::memcpy(h.host_err, static_cast<uint32_t*>(h_rb_err.p) + size_t(a) * kCoupledErrSlots,
                        sizeof(uint32_t) * kCoupledErrSlots);
        }
        return MP_OK;
    };
    auto failed = [&](int k, mp_status st) {
        ms[k].st = st;
        ms[k].msg = last_error();
        ms[k].phase = BatchMeshState::kFailed;
    };

    // the starting RHS and state check (mp_coupled::run before its loop)
    std::vector<int> act;
    for (int k = 0; k < n; ++k)
        if (ms[k].phase == BatchMeshState::kRunning) {
            act.push_back(k);
            seg_pending.push_back(k);
            seg_pending.push_back(1);      // O <- S
        }
    MP_CALL(round(act, true));
    for (int k : act) {
        BatchMeshState& m = ms[k];
        mp_status st = hs[k]->rhs_result(0, kCoupledRedRhs0, m.current);
        if (st) {
            failed(k, st);
            continue;
        }
        bool ok;
        hs[k]->admissible_result(kCoupledRedAdm0, ok, m.oldH, m.oldC);
    }
    for (;;) {
        act.clear();
        for (int k = 0; k < n; ++k) {
            BatchMeshState& m = ms[k];
            if (m.phase != BatchMeshState::kRunning) continue;
            const mp_coupled_time& tc = tcs[k];
            if (m.new_step) {
                if (m.s >= nsteps) {
                    m.phase = BatchMeshState::kDone;
                    continue;
                }
                m.remaining = tc.end_time - times[k];
                if (m.remaining <= 32 * 2.220446049250313e-16 * std::fmax(1.0, std::fabs(tc.end_time))) {
                    times[k] = tc.end_time;
                    m.phase = BatchMeshState::kDone;
                    continue;
                }
                m.start = times[k];
                m.dt = std::fmin(m.remaining, tc.fixed ? tc.initial_step
                                                        : (m.accepted == 0 ? tc.initial_step : tc.maximum_step));
                const double limit = m.current.maximum_rate > 0 ? tc.cfl / m.current.maximum_rate : tc.maximum_step;
                if (tc.fixed && !(m.dt <= limit * (1 + 1e-12))) {
                    failed(k, fail(MP_ERR_ARG, "Fixed time step exceeds the phase-coupled CFL bound; use control: cfl"));
                    continue;
                }
                if (!tc.fixed) m.dt = std::fmin(m.dt, limit);
                seg_pending.push_back(k);
                seg_pending.push_back(0);  // F <- R
                m.first = m.current;
                m.retries = 0;
                m.new_step = false;
            }
            if (!(m.dt >= std::fmin(m.remaining, tc.minimum_step) && m.retries < 40)) {
                failed(k, fail(MP_ERR_PHASE, "Coupled step cannot satisfy admissibility/CFL above minimum step"));
                times[k] = m.start;        // roll back to the latest accepted state
                seg_pending.push_back(k);
                seg_pending.push_back(2);  // S <- O
                continue;
            }
            act.push_back(k);
        }
        if (act.empty()) break;
        MP_CALL(round(act, false));
        for (int k : act) {
            BatchMeshState& m = ms[k];
            const mp_coupled_time& tc = tcs[k];
            mp_coupled& h = *hs[k];
            times[k] = m.start + m.dt;
            bool ok;
            double ih, ic;
            mp_status st = MP_OK;
            h.admissible_result(kCoupledRedAdm1, ok, ih, ic);
            if (ok) {
                st = h.rhs_result(1, kCoupledRedRhs1, m.second);
                if (!st) {
                    m.courant = m.dt * std::fmax(m.current.maximum_rate, m.second.maximum_rate);
                    ok = m.courant <= tc.cfl * (1 + 1e-12);
                }
            }
            if (!st && ok) h.admissible_result(kCoupledRedAdm2, ok, m.newH, m.newC);
            if (!st && !ok) {
                if (!tc.fixed) {           // retry at half the step next round
                    m.dt *= .5;
                    ++m.retries;
                    continue;
                }
                st = fail(MP_ERR_ARG, "Fixed coupled step failed stage admissibility/CFL");
            }
            double balH = 0, balC = 0;
            if (!st) {
                balH = m.newH - m.oldH + .5 * m.dt * (m.current.boundary_H + m.second.boundary_H -
                                                      m.current.source_H - m.second.source_H);
                balC = m.newC - m.oldC + .5 * m.dt * (m.current.boundary_C + m.second.boundary_C);
                if (!(std::fabs(balH) < 1e-10 * std::fmax(1.0, std::fabs(m.oldH)) && std::fabs(balC) < 1e-10))
                    st = fail(MP_ERR_PHASE, "Coupled finite-volume conservation balance failed");
            }
            if (!st) st = h.rhs_result(2, kCoupledRedRhs2, m.current);
            if (st) {
                failed(k, st);
                times[k] = m.start;        // roll back to the latest accepted state
                seg_pending.push_back(k);
                seg_pending.push_back(2);  // S <- O
                continue;
            }
            ++m.accepted;
            m.oldH = m.newH;
            m.oldC = m.newC;
            seg_pending.push_back(k);
            seg_pending.push_back(1);      // O <- S
            if (steps) {
                mp_coupled_step& r = steps[size_t(k) * nsteps + m.s];
                r.dt = m.dt;
                r.courant = m.courant;
                r.balance_H = balH;
                r.balance_C = balC;
                r.boundary_H = .5 * (m.first.boundary_H + m.second.boundary_H);
                r.boundary_C = .5 * (m.first.boundary_C + m.second.boundary_C);
                r.source_H = .5 * (m.first.source_H + m.second.source_H);
                r.maximum_rate = m.current.maximum_rate;
                r.retries = m.retries;
                r.pad = 0;
            }
            ++m.s;
            m.new_step = true;
        }
    }
    act.clear();
    MP_CALL(round(act, false));            // the copies still owed (accepted states, roll-backs)
    mp_status first_st = MP_OK;
    int first_k = -1;
    for (int k = 0; k < n; ++k) {
        if (accepted_out) accepted_out[k] = ms[k].accepted;
        if (status_out) status_out[k] = ms[k].st;
        if (ms[k].st && first_k < 0) {
            first_st = ms[k].st;
            first_k = k;
        }
    }
    if (first_k >= 0) return fail(first_st, "coupled handle %d: %s", first_k, ms[first_k].msg.c_str());
    return MP_OK;
}

// ===================================================================== C ABI

extern "C" {

void mp_default_coupled_params(mp_coupled_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->theta = 1.0;
    p->cooling = 0.0;
    p->thermal_diffusivity = 0.0;
    p->cell_points_per_axis = 5;
    p->diffusion_samples = 4;
    p->extent_fraction = 0.9;
    p->lf_mode = 0;
    p->global_speed = 0.0;
}

mp_status mp_coupled_create(const mp_mesh_desc* mesh, const mp_phase_params* phase,
                            const mp_coupled_params* params, void* stream, mp_coupled** out)
{
    MP_GUARD({
        if (!mesh || !phase || !params || !out) return fail(MP_ERR_ARG, "null argument");
        *out = nullptr;
        auto* c = new mp_coupled();
        const mp_status s = c->init(*mesh, *phase, *params, stream);
        if (s) {
            delete c;
            return s;
        }
        *out = c;
        return MP_OK;
    })
}

mp_status mp_coupled_destroy(mp_coupled* c)
{
    delete c;
    return MP_OK;
}

mp_status mp_coupled_set_flow(mp_coupled* c, const double* edge_stokes, const double* edge_darcy,
                              const double* edge_phi, const double* cell_stokes,
                              const double* cell_darcy, const double* cell_phi)
{
    if (!c) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return c->set_flow(edge_stokes, edge_darcy, edge_phi, cell_stokes, cell_darcy, cell_phi); })
}

mp_status mp_coupled_set_boundary(mp_coupled* c, int32_t field, const uint8_t* bottom,
                                  const uint8_t* right, const uint8_t* top, const uint8_t* left)
{
    if (!c) return fail(MP_ERR_ARG, "null handle");
    const uint8_t* sides[4] = {bottom, right, top, left};
    MP_GUARD({ return c->set_boundary(field, sides); })
}

mp_status mp_coupled_set_inflow(mp_coupled* c, const double* inflow_H, const double* inflow_C)
{
    if (!c) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return c->set_inflow(inflow_H, inflow_C); })
}

mp_status mp_coupled_set_thermal_boundary(mp_coupled* c, const uint8_t* types, const double* values)
{
    if (!c) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return c->set_thermal(types, values); })
}

mp_status mp_coupled_set_lithostatic(mp_coupled* c, int32_t enable, double surface_y)
{
    if (!c) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return c->set_lithostatic(enable, surface_y); })
}

mp_status mp_coupled_upload_state(mp_coupled* c, const double* H, const double* C)
{
    if (!c || !H || !C) return fail(MP_ERR_ARG, "null argument");
    MP_GUARD({
        const size_t nb = c->N * sizeof(double);
        MP_HIP(hipMemcpyAsync(c->state.ptr, H, nb, hipMemcpyHostToDevice, c->stream));
        MP_HIP(hipMemcpyAsync(c->state.as<double>() + c->N, C, nb, hipMemcpyHostToDevice, c->stream));
        MP_HIP(hipStreamSynchronize(c->stream));
        return MP_OK;
    })
}

mp_status mp_coupled_download_state(mp_coupled* c, double* H, double* C)
{
    if (!c) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({
        const size_t nb = c->N * sizeof(double);
        if (H) MP_HIP(hipMemcpyAsync(H, c->state.ptr, nb, hipMemcpyDeviceToHost, c->stream));
        if (C) MP_HIP(hipMemcpyAsync(C, c->state.as<double>() + c->N, nb, hipMemcpyDeviceToHost, c->stream));
        MP_HIP(hipStreamSynchronize(c->stream));
        return MP_OK;
    })
}

mp_status mp_coupled_rhs(mp_coupled* c, double* rhsH, double* rhsC, double* rate, double* div,
                         double* T, double* phi, mp_coupled_rates* totals)
{
    if (!c) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return c->rhs(rhsH, rhsC, rate, div, T, phi, totals); })
}

mp_status mp_coupled_run(mp_coupled* c, const mp_coupled_time* time_controls, double* time,
                         int32_t nsteps, mp_coupled_step* steps, int32_t* accepted)
{
    if (!c || !time_controls) return fail(MP_ERR_ARG, "null argument");
    if (nsteps < 0) return fail(MP_ERR_ARG, "nsteps must be >= 0");
    MP_GUARD({ return c->run(*time_controls, time, nsteps, steps, accepted); })
}

mp_status mp_coupled_run_batch(mp_coupled* const* handles, int32_t n, const mp_coupled_time* time_controls,
                               double* times, int32_t nsteps, mp_coupled_step* steps, int32_t* accepted,
                               mp_status* status)
{
    MP_GUARD({ return coupled_run_batch(handles, n, time_controls, times, nsteps, steps, accepted, status); })
}

}  // extern "C"
