// Single-mesh transport handle and the C ABI of include/mantlepar_amd.h.
//
// The handle owns all device buffers of one rank's share of the mesh: three
// full H/C states (the state at the start of a call is never overwritten, so a
// failed call leaves it intact), two stage buffers, face velocities, boundary
// data, two phase-output sets and the reduction scratch.  A call to
// mp_transport_step launches the stage kernels asynchronously, synchronises
// once, and commits the new state only if the device error word is clean,
// which mirrors the reference's collective Agree after every step
// (src/couple/evolve.cpp:23-28, 493-499).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <vector>

#include "mp_host.h"
#include "mp_launch.h"
#include "mp_transport.h"

using namespace mp;

namespace mp {
const char* last_error();
}

// Stage tables: u_out = A*u_n + B*(u_k + dt*L(u_k)); W weights the stage
// boundary fluxes in the step's conservation balance.
static const double kA[3][3] = {{0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
static const double kB[3][3] = {{1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};
static const double kW[3][3] = {{1, 0, 0}, {0.5, 0.5, 0}, {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}};

mp_status mp_transport::init(const mp_mesh_desc& m, const mp_weno_options& w,
                             const mp_phase_params& p, int32_t prec, void* s)
{
    if (prec != MP_PREC_FP64 && prec != MP_PREC_MIXED) return fail(MP_ERR_ARG, "unknown precision");
    MP_CALL(make_phase(p, phase));
    MP_CALL(make_mesh_dev(m, w, phase, md));
    mesh = m;
    weno = w;
    phase_params = p;
    precision = prec;
    N = size_t(m.nx) * size_t(m.ny);
    MP_HIP(hipGetDevice(&device));
    if (s) {
        stream = static_cast<hipStream_t>(s);
        own_stream = false;
    } else {
        MP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        own_stream = true;
    }
    {   // side stream: each step's phase and totals beside the next step
        MP_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        for (int k = 0; k < 2; ++k) {
            MP_HIP(hipEventCreateWithFlags(&stage3[k], hipEventDisableTiming));
            MP_HIP(hipEventCreateWithFlags(&side_done[k], hipEventDisableTiming));
        }
        MP_HIP(hipEventCreateWithFlags(&side_join, hipEventDisableTiming));
    }
    for (int k = 0; k < 5; ++k) {
        MP_CALL(H[k].alloc(N * sizeof(double)));
        MP_CALL(C[k].alloc(N * sizeof(double)));
    }
    for (int k = 0; k < 2; ++k) {
        MP_CALL(phase_buf[k].alloc(N * (4 * sizeof(double) + 1)));
    }
    MP_CALL(vnx.alloc(size_t(m.q) * m.ny * (m.nx + 1) * sizeof(double)));
    MP_CALL(vny.alloc(size_t(m.q) * (m.ny + 1) * m.nx * sizeof(double)));
    if (prec == MP_PREC_MIXED) {   // face moment records: the mixed sweep's flux input
        const size_t nxf = size_t(m.ny) * (m.nx + 1), nyf = size_t(m.ny + 1) * m.nx;
        MP_CALL(fx.alloc(nxf * sizeof(FaceMom)));
        MP_CALL(fy.alloc(nyf * sizeof(FaceMom)));
        MP_CALL(rx.alloc(nxf * sizeof(FaceRec)));
        MP_CALL(ry.alloc(nyf * sizeof(FaceRec)));
        md.fx = fx.as<FaceMom>();
        md.fy = fy.as<FaceMom>();
        md.rx = rx.as<FaceRec>();
        md.ry = ry.as<FaceRec>();
    } else {                       // one-sided FP64 records: the FP64 sweep's flux input
        const size_t nxf = size_t(m.ny) * (m.nx + 1), nyf = size_t(m.ny + 1) * m.nx;
        MP_CALL(rx64.alloc(nxf * sizeof(FaceRec64)));
        MP_CALL(ry64.alloc(nyf * sizeof(FaceRec64)));
        md.rx64 = rx64.as<FaceRec64>();
        md.ry64 = ry64.as<FaceRec64>();
    }
    MP_CALL(error.alloc(sizeof(uint64_t) * 2));
    MP_CALL(d_md.alloc(sizeof(MeshDev)));
    MP_CALL(d_tab.alloc(sizeof(StagePtrs) * 9));
    MP_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_tab), sizeof(StagePtrs) * 9, hipHostMallocDefault));
    MP_HIP(hipEventCreateWithFlags(&tab_copied, hipEventDisableTiming));
    MP_HIP(hipEventRecord(tab_copied, stream));
    MP_HIP(hipMemsetAsync(H[0].ptr, 0, N * sizeof(double), stream));
    MP_HIP(hipMemsetAsync(C[0].ptr, 0, N * sizeof(double), stream));
    md.vnx = vnx.as<double>();
    md.vny = vny.as<double>();
    MP_HIP(hipMemsetAsync(vnx.ptr, 0, vnx.bytes, stream));
    MP_HIP(hipMemsetAsync(vny.ptr, 0, vny.bytes, stream));
    MP_CALL(sync_md());
    set_tile(default_rows());
    return MP_OK;
}

// Rows per work item when the caller sets none, by how many items of a height
// each wave slot of the chip gets.  FP64 sweep (2 waves per SIMD): 16 rows
// when every slot gets two 16-row items, else 8.  Mixed sweep (3 waves per
// SIMD): 16 rows at four items per slot, 12 at two, else 8.  These were the
// fastest measured on MI355X (mixed: 2048x4096 12 rows, 4096x4096 and larger
// 16; FP64: 2048x4096 16; items sized to fill the chip in a single round ran
// slower), and a small mesh's stage time is the time of one item, so short
// items serve small meshes best.
int mp_transport::default_rows() const
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const long long strips = (mesh.nx + kStripCells - 1) / kStripCells;
    auto items = [&](int rows) { return strips * ((mesh.ny + rows - 1) / rows); };
    if (precision == MP_PREC_MIXED) {
        const long long slots = (long long)cus * 4 * 3;
        if (items(16) >= 4 * slots) return 16;
        return items(12) >= 2 * slots ? 12 : 8;
    }
    const long long slots = (long long)cus * 4 * 2;
    return items(16) >= 2 * slots ? 16 : 8;
}

mp_transport::~mp_transport()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    for (int k = 0; k < 2; ++k) {
        if (stage3[k]) (void)hipEventDestroy(stage3[k]);
        if (side_done[k]) (void)hipEventDestroy(side_done[k]);
    }
    if (side_join) (void)hipEventDestroy(side_join);
    if (tab_copied) (void)hipEventDestroy(tab_copied);
    if (h_tab) (void)hipHostFree(h_tab);
    if (side) (void)hipStreamDestroy(side);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

mp_status mp_transport::sync_md()
{
    MP_HIP(hipMemcpyAsync(d_md.ptr, &md, sizeof(MeshDev), hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

void mp_transport::set_tile(int r)
{
    rows = r < 4 ? 4 : r;
    layout = item_layout(mesh.nx, mesh.ny, rows, precision);
    strips = layout.strips;
    nitems = layout.nitems;
}

// Work memory of a step over nitems work items and nsteps steps: two sets of
// three stage partials, the step reduction's block results, the totals (of the
// starting state, then of each step) and the error word (the last two
// 256-byte aligned).
size_t mp_transport::work_bytes(int nitems, int nsteps)
{
    const size_t part = (size_t(6) * nitems * kPartials + size_t(kStepSlices) * 6) * sizeof(double);
    const size_t tot = size_t(nsteps + 1) * 4 * sizeof(double);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    return up(part) + up(tot) + 256;
}

mp_status mp_transport::ensure_scratch(int nsteps)
{
    // stage partials of two consecutive steps, then the step reduction's
    // block results
    const size_t pbytes = (size_t(6) * nitems * kPartials + size_t(kStepSlices) * 6) * sizeof(double);
    if (partials.bytes < pbytes) MP_CALL(partials.alloc(pbytes));
    // per-step totals: at least 256 steps, grown geometrically, so that calls
    // of varying length do not reallocate (a reallocation synchronises)
    const size_t tbytes = size_t(nsteps + 1) * 4 * sizeof(double);
    if (totals.bytes < tbytes)
        MP_CALL(totals.alloc(std::max({tbytes, 2 * totals.bytes, size_t(257) * 4 * sizeof(double)})));
    const size_t fbytes = size_t(3) * nitems * sizeof(uint32_t);
    if (flags.bytes < fbytes) {
        // a new flag array holds no epoch yet
        MP_CALL(flags.alloc(fbytes));
        MP_HIP(hipMemsetAsync(flags.ptr, 0, fbytes, stream));
    }
    return MP_OK;
}

mp_transport::StepWork mp_transport::own_work()
{
    return StepWork{partials.as<double>(), totals.as<double>(), reinterpret_cast<uint32_t*>(error.ptr),
                    flags.as<uint32_t>(), flags.bytes / sizeof(uint32_t), &epoch};
}

// One launch per step on the mixed path.  MP_LAUNCH_AUTO takes it for meshes
// that give each wave slot of the chip fewer than eight items: there a stage
// launch's ramp and drain (about 90 us) are a large part of it, and the
// per-step launch, which pays them once per step, measured 5% faster at
// 2048 x 4096 (0.666 against 3 x 0.234 ms); at 8192 x 8192 the per-stage
// launches were 3% faster (3 x 1.32 against 4.09 ms), the waits for
// neighbouring items costing what the shared ramp saves.
// MP_LAUNCH_AUTO takes the per-step launch at every size: with its rows
// unrolled it measured faster than three per-stage launches from 2048x4096
// (0.584 against 0.652 ms per step) to 8192^2 (4.37 against 4.53 ms).
bool mp_transport::step_fused(int scheme) const
{
    if (precision != MP_PREC_MIXED || scheme < 2 || md.limit != 0) return false;
    return launch_mode != MP_LAUNCH_PER_STAGE;
}

mp_status mp_transport::upload(const double* h, const double* c)
{
    if (!h || !c) return fail(MP_ERR_ARG, "null state array");
    if (!all_finite(h, N) || !all_finite(c, N)) return fail(MP_ERR_ARG, "Cell averages must be finite");
    MP_HIP(hipMemcpyAsync(H[cur].ptr, h, N * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(C[cur].ptr, c, N * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    phase_valid = false;
    return MP_OK;
}

mp_status mp_transport::download(double* h, double* c)
{
    if (h) MP_HIP(hipMemcpyAsync(h, H[cur].ptr, N * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (c) MP_HIP(hipMemcpyAsync(c, C[cur].ptr, N * sizeof(double), hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status mp_transport::set_velocity(const double* x, const double* y)
{
    if (!x || !y) return fail(MP_ERR_ARG, "null velocity array");
    const size_t nxv = size_t(mesh.q) * mesh.ny * (mesh.nx + 1);
    const size_t nyv = size_t(mesh.q) * (mesh.ny + 1) * mesh.nx;
    if (!all_finite(x, nxv) || !all_finite(y, nyv)) return fail(MP_ERR_ARG, "Edge velocity must be finite");
    if (weno.lf_mode == 1) {
        double vmax = 0;
        for (size_t k = 0; k < nxv; ++k) vmax = std::fmax(vmax, std::fabs(x[k]));
        for (size_t k = 0; k < nyv; ++k) vmax = std::fmax(vmax, std::fabs(y[k]));
        if (weno.global_speed + 1e-14 < vmax)
            return fail(MP_ERR_ARG, "Global LF speed is smaller than the flow normal speed");
    }
    MP_HIP(hipMemcpyAsync(vnx.ptr, x, nxv * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(vny.ptr, y, nyv * sizeof(double), hipMemcpyHostToDevice, stream));
    if (precision == MP_PREC_MIXED)
        MP_HIP(launch_face_prep(d_md.as<MeshDev>(), mesh.nx, mesh.ny, fx.as<FaceMom>(), fy.as<FaceMom>(),
                                rx.as<FaceRec>(), ry.as<FaceRec>(), stream));
    else
        MP_HIP(launch_face_rec64(d_md.as<MeshDev>(), mesh.nx, mesh.ny, rx64.as<FaceRec64>(),
                                 ry64.as<FaceRec64>(), stream));
    MP_CALL(compute_rate());
    have_velocity = true;
    return MP_OK;
}

mp_status mp_transport::compute_rate()
{
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(error.as<uint64_t>() + 1);
    MP_HIP(hipMemsetAsync(bits, 0, sizeof(unsigned long long), stream));
    MP_HIP(launch_rate(d_md.as<MeshDev>(), int(N), bits, stream));
    unsigned long long host = 0;
    MP_HIP(hipMemcpyAsync(&host, bits, sizeof host, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    std::memcpy(&max_rate, &host, sizeof max_rate);
    return MP_OK;
}

mp_status mp_transport::set_boundary(const uint8_t* const side[4])
{
    const size_t len[4] = {size_t(mesh.nx), size_t(mesh.ny), size_t(mesh.nx), size_t(mesh.ny)};
    for (int s = 0; s < 4; ++s) {
        if (side[s])
            for (size_t k = 0; k < len[s]; ++k)
                if (side[s][k] > MP_BC_INFLOW_OUTFLOW) return fail(MP_ERR_ARG, "Unknown advection boundary policy");
    }
    for (int s = 0; s < 4; ++s) {
        if (!side[s]) {
            bc[s].reset();
            md.bc[s] = nullptr;
            continue;
        }
        MP_CALL(bc[s].alloc(len[s]));
        MP_HIP(hipMemcpyAsync(bc[s].ptr, side[s], len[s], hipMemcpyHostToDevice, stream));
        md.bc[s] = bc[s].as<uint8_t>();
    }
    return sync_md();
}

mp_status mp_transport::set_inflow(const double* h, const double* c)
{
    const size_t total = size_t(2) * (mesh.nx + mesh.ny) * mesh.q;
    const size_t off[4] = {0, size_t(mesh.nx) * mesh.q, size_t(mesh.nx + mesh.ny) * mesh.q,
                           size_t(2 * mesh.nx + mesh.ny) * mesh.q};
    for (int f = 0; f < 2; ++f) {
        const double* src = f ? c : h;
        DeviceBuffer& buf = f ? inflowC : inflowH;
        const double** dst = f ? md.inflowC : md.inflowH;
        if (!src) {
            buf.reset();
            for (int s = 0; s < 4; ++s) dst[s] = nullptr;
            continue;
        }
        if (!all_finite(src, total)) return fail(MP_ERR_FP, "Inflow boundary callback must return a finite state");
        MP_CALL(buf.alloc(total * sizeof(double)));
        MP_HIP(hipMemcpyAsync(buf.ptr, src, total * sizeof(double), hipMemcpyHostToDevice, stream));
        for (int s = 0; s < 4; ++s) dst[s] = buf.as<double>() + off[s];
    }
    return sync_md();
}

StagePtrs mp_transport::stage_ptrs(const StepWork& w, int src, int kin, int out, int scheme, int st,
                                   double dt, bool phase_stage, int phase_set, int part_stage) const
{
    StagePtrs sp;
    std::memset(&sp, 0, sizeof sp);
    sp.Hn = H[src].as<double>();
    sp.Cn = C[src].as<double>();
    sp.Hk = H[kin].as<double>();
    sp.Ck = C[kin].as<double>();
    sp.Ho = H[out].as<double>();
    sp.Co = C[out].as<double>();
    sp.a = kA[scheme - 1][st];
    sp.b = kB[scheme - 1][st];
    sp.dt = dt;
    if (phase_stage) {
        double* base = phase_buf[phase_set].as<double>();
        sp.T = base;
        sp.phi = base + N;
        sp.Cl = base + 2 * N;
        sp.Cs = base + 3 * N;
        sp.region = reinterpret_cast<uint8_t*>(base + 4 * N);
    }
    sp.partials = w.partials + size_t(part_stage) * nitems * kPartials;
    sp.error = w.error;
    apply_layout(layout, sp);
    return sp;
}

// Enqueue nsteps; returns the index of the final state buffer.
//
// Each step's phase kernel and step totals run on the side stream, beside
// the next step's stages: the phase reads the step's output
// state, which the next step only reads, and writes the phase outputs and the
// sum slots of the step's partials.  Partials alternate between two sets by
// step parity, and step s + 2, which writes step s's output buffer and
// partial set again, first waits for step s's phase and totals.  The call
// ends with the main stream waiting for the side stream.  Profiling (events)
// keeps everything on the main stream so each launch is timed alone.
mp_status mp_transport::enqueue_steps(const StepWork& w, int nsteps, int scheme, double dt,
                                      int& final_state, std::vector<hipEvent_t>* events)
{
    int others[2], n = 0;
    for (int k = 0; k < 3; ++k)
        if (k != cur) others[n++] = k;
    int src = cur;
    final_state = cur;
    const int phase_set = 1 - phase_cur;
    const bool fused = step_fused(scheme);
    // One launch per mixed step: each launch also runs the previous step's
    // phase items after its own sweep items, in the launch's drain, and the
    // previous step's totals follow it; the last step's phase and totals
    // close the call.  Per-stage launches (both precisions) run each step's
    // phase and totals on the side stream, beside the next step.  Profiling (events)
    // keeps every kernel in its own launch on the main stream.
    const bool fold = fused && precision == MP_PREC_MIXED && !events;
    const bool overlap = side && !events && !fused;
    // However the call ends, error returns included, the main stream waits for
    // everything enqueued on the side stream.
    struct Join {
        hipStream_t main, side;
        hipEvent_t ev;
        bool on;
        ~Join()
        {
            if (on && hipEventRecord(ev, side) == hipSuccess) (void)hipStreamWaitEvent(main, ev, 0);
        }
    } join{stream, side, side_join, overlap && nsteps > 0};
    const size_t set_doubles = size_t(nitems) * kPartials;
    double* scratch = w.partials + 6 * set_doubles;
    // the starting state's totals, against which the first step is balanced
    if (nsteps > 0) MP_HIP(launch_state_sums(H[cur].as<double>(), C[cur].as<double>(), int64_t(N), md.area,
                                             w.totals, scratch, stream));
    if (fused) {
        // Stage records of the call's three kinds of step (step 0, odd steps,
        // even steps >= 2), uploaded once; the previous call's upload from the
        // same pinned buffer has completed first.
        MP_HIP(hipEventSynchronize(tab_copied));
        const int src_of[3] = {cur, others[0], others[1]};
        const int dst_of[3] = {others[0], others[1], others[0]};
        const int par_of[3] = {0, overlap || fold ? 1 : 0, 0};
        for (int set = 0; set < 3; ++set) {
            int kin = src_of[set];
            for (int st = 0; st < scheme; ++st) {
                const bool last = st == scheme - 1;
                const int out = last ? dst_of[set] : 3 + st;
                h_tab[set * 3 + st] = stage_ptrs(w, src_of[set], kin, out, scheme, st, dt, last, phase_set,
                                                 par_of[set] * 3 + st);
                kin = out;
            }
        }
        MP_HIP(hipMemcpyAsync(d_tab.ptr, h_tab, sizeof(StagePtrs) * 9, hipMemcpyHostToDevice, stream));
        MP_HIP(hipEventRecord(tab_copied, stream));
    }
    const size_t set_doubles_ = set_doubles;
    auto reduce_on = [&](int step, int par, hipStream_t st) {
        // totals of step `step` in slot step + 1, checked against slot step
        // (the preheat's balance, src/couple/preheat.cpp:268-273)
        return launch_reduce_step(w.partials + size_t(par) * 3 * set_doubles_, nitems, scheme, kW[scheme - 1],
                                  md.area, w.totals + size_t(step + 1) * 4, scratch, st,
                                  w.totals + size_t(step) * 4, dt, balance_tol, w.error);
    };
    for (int s = 0; s < nsteps; ++s) {
        const int dst = others[s % 2];
        const int par = overlap || fold ? s % 2 : 0;
        if (overlap && s >= 2) MP_HIP(hipStreamWaitEvent(stream, side_done[par], 0));
        if (fused) {
            // one launch for the step's stages, then its phase as below
            const int set = s == 0 ? 0 : (s % 2 ? 1 : 2);
            if (*w.epoch >= 0xfffffff0u) {   // flags hold only epochs: restart them
                MP_HIP(launch_clear(w.flags, int(w.flag_words), stream));
                *w.epoch = 0;
            }
            const uint32_t ep = ++*w.epoch;
            if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3)], stream));
            const bool stall = launch_mode == MP_LAUNCH_STALL_CHECK;
            // the previous step's last stage record: its phase items ride along
            const int prev = s == 1 ? 0 : (s % 2 ? 2 : 1);
            const StagePtrs* ptab = fold && s > 0 ? d_tab.as<StagePtrs>() + prev * 3 + scheme - 1 : nullptr;
            MP_HIP(launch_step32(d_md.as<MeshDev>(), d_tab.as<StagePtrs>() + set * 3, nitems, scheme,
                                 w.flags, ep, stall ? kEpochNever : ep, stall ? 64 : kSpinLimit, stream,
                                 has_diffusion(), md.lf_mode != 0, ptab, ptab ? nitems : 0));
            if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3) + 1], stream));
            const StagePtrs sp = h_tab[set * 3 + scheme - 1];
            if (fold) {
                if (s > 0) MP_HIP(reduce_on(s - 1, 1 - par, stream));
                if (s == nsteps - 1) {
                    MP_HIP(launch_phase_items(d_md.as<MeshDev>(), sp, nitems, stream));
                    MP_HIP(reduce_on(s, par, stream));
                }
                src = dst;
                final_state = dst;
                continue;
            } else {
                if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3 + 1)], stream));
                MP_HIP(launch_phase_items(d_md.as<MeshDev>(), sp, nitems, stream));
                if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3 + 1) + 1], stream));
            }
        }
        int kin = src;
        for (int st = 0; st < scheme && !fused; ++st) {
            const bool last = st == scheme - 1;
            const int out = last ? dst : 3 + st;
            const StagePtrs sp = stage_ptrs(w, src, kin, out, scheme, st, dt, last, phase_set, par * 3 + st);
            // both sweeps leave the phase to its own streaming kernel: fused
            // into the FP64 sweep's last stage it cost that stage more than
            // the streaming kernel takes, and it could not run beside the
            // next step
            if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3 + st)], stream));
            MP_HIP(launch_stage(precision, mesh.q, d_md.as<MeshDev>(), sp, nitems, stream, limit_mode32(md),
                                has_diffusion(), md.lf_mode != 0));
            if (last && overlap) {
                MP_HIP(hipEventRecord(stage3[par], stream));
                MP_HIP(hipStreamWaitEvent(side, stage3[par], 0));
                MP_HIP(launch_phase_items(d_md.as<MeshDev>(), sp, nitems, side));
            } else if (last) {
                MP_HIP(launch_phase_items(d_md.as<MeshDev>(), sp, nitems, stream));
            }
            if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3 + st) + 1], stream));
            kin = out;
        }
        MP_HIP(reduce_on(s, par, overlap ? side : stream));
        if (overlap) MP_HIP(hipEventRecord(side_done[par], side));
        src = dst;
        final_state = dst;
    }
    return MP_OK;
}

bool mp_transport::has_diffusion() const
{
    return mp::has_diffusion(md);
}

mp_status mp_transport::check_step_args(int nsteps, int scheme, double dt)
{
    if (!have_velocity) return fail(MP_ERR_STATE, "set face velocities before stepping");
    if (scheme < 1 || scheme > 3) return fail(MP_ERR_ARG, "scheme must be 1 (FE), 2 (SSPRK2) or 3 (SSPRK3)");
    if (nsteps < 0) return fail(MP_ERR_ARG, "nsteps must be nonnegative");
    if (!std::isfinite(dt) || !(dt > 0)) return fail(MP_ERR_ARG, "Time step does not advance finite time");
    return MP_OK;
}

mp_status mp_transport::step(int nsteps, int scheme, double dt, mp_step_report* report)
{
    MP_CALL(check_step_args(nsteps, scheme, dt));
    MP_CALL(ensure_scratch(nsteps));
    const StepWork w = own_work();
    MP_HIP(launch_clear(w.error, 1, stream));
    int final_state = cur;
    MP_CALL(enqueue_steps(w, nsteps, scheme, dt, final_state, nullptr));
    return finish_step(w, nsteps, scheme, dt, final_state, report);
}

// Wait for the enqueued steps, commit the new state only if the error word is
// clean, and fill the report from the per-step totals.  The totals are read
// through a fixed host buffer, 64 steps at a time, so this allocates nothing.
mp_status mp_transport::finish_step(const StepWork& w, int nsteps, int scheme, double dt,
                                    int final_state, mp_step_report* report)
{
    (void)scheme;
    uint32_t bits = 0;
    MP_HIP(hipMemcpyAsync(&bits, w.error, sizeof bits, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x); state unchanged",
                          describe_bits(bits), bits);
    if (nsteps > 0) {
        cur = final_state;
        phase_cur = 1 - phase_cur;
        phase_valid = true;
    }
    if (report) {
        std::memset(report, 0, sizeof *report);
        report->dt = dt;
        report->max_rate = max_rate;
        report->courant = dt * max_rate;
        report->steps_done = nsteps;
        long double bh = 0, bc = 0;
        constexpr int kChunk = 64;
        double tot[kChunk * 4];
        for (int s0 = 0; s0 < nsteps; s0 += kChunk) {
            const int n = nsteps - s0 < kChunk ? nsteps - s0 : kChunk;
            MP_HIP(hipMemcpy(tot, w.totals + size_t(s0 + 1) * 4, size_t(n) * 4 * sizeof(double),
                             hipMemcpyDeviceToHost));
            for (int s = 0; s < n; ++s) {
                bh += (long double)dt * tot[size_t(s) * 4 + 2];
                bc += (long double)dt * tot[size_t(s) * 4 + 3];
            }
            if (s0 + n == nsteps) {
                report->total_H = tot[size_t(n - 1) * 4 + 0];
                report->total_C = tot[size_t(n - 1) * 4 + 1];
            }
        }
        report->boundary_H = double(bh);
        report->boundary_C = double(bc);
        if (nsteps == 0) MP_CALL(state_totals(report->total_H, report->total_C));
    }
    return MP_OK;
}

// ------------------------------------------------- allocation-free stepping
//
// mp_transport_step_noalloc keeps a step's work memory (stage partials, the
// step reduction's block results, per-step totals, the error word) in one
// file-scope device array instead of buffers the handle allocates.  Each GPU
// holds its own copy of the array at its own address until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device, never through the array's name.  Every call on one
// GPU shares that GPU's array, so calls that overlap on one GPU race.
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
// completion flags of the one-launch step: 3 per work item, up to 262144 items
constexpr size_t kScratchFlagWords = size_t(3) << 18;
std::mutex scratch_mutex;
void* scratch_addr[kMaxDevices] = {};
uint32_t scratch_epoch[kMaxDevices] = {};   // epochs used on each device's flags

// The scratch address on `device` (which must be current), looked up once.
mp_status scratch_on(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(MP_ERR_DEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MP_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        scratch_addr[device] = p;
    }
    *out = scratch_addr[device];
    return MP_OK;
}
}  // namespace

mp_status mp_transport::step_noalloc(int nsteps, int scheme, double dt, mp_step_report* report)
{
    MP_CALL(check_step_args(nsteps, scheme, dt));
    const size_t need = work_bytes(nitems, nsteps) + kScratchFlagWords * sizeof(uint32_t);
    if (need > kNoAllocScratchBytes)
        return fail(MP_ERR_MEMORY,
                    "mp_transport_step_noalloc: %d work items x %d steps need %zu bytes of scratch, "
                    "more than its %zu; use mp_transport_step or fewer steps per call",
                    nitems, nsteps, need, kNoAllocScratchBytes);
    int dev = -1;
    MP_HIP(hipGetDevice(&dev));
    if (dev != device)
        return fail(MP_ERR_STATE, "mp_transport_step_noalloc: the handle's device %d is not current (%d)",
                    device, dev);
    if (size_t(3) * nitems > kScratchFlagWords)
        return fail(MP_ERR_MEMORY, "mp_transport_step_noalloc: %d work items exceed the scratch's %zu",
                    nitems, kScratchFlagWords / 3);
    void* base = nullptr;
    MP_CALL(scratch_on(dev, &base));
    // The error word, the completion flags (a fixed region that never holds
    // anything but epochs), the totals, then the partials (each 256-byte aligned).
    char* p = static_cast<char*>(base);
    const size_t tot = (size_t(nsteps + 1) * 4 * sizeof(double) + 255) & ~size_t(255);
    char* q = p + 256 + kScratchFlagWords * sizeof(uint32_t);
    const StepWork w{reinterpret_cast<double*>(q + tot), reinterpret_cast<double*>(q),
                     reinterpret_cast<uint32_t*>(p), reinterpret_cast<uint32_t*>(p + 256),
                     kScratchFlagWords, &scratch_epoch[dev]};
    MP_HIP(launch_clear(w.error, 1, stream));
    // before the first epoch on this device the flags are cleared, whatever
    // the array held at load time
    if (scratch_epoch[dev] == 0) MP_HIP(launch_clear(w.flags, int(kScratchFlagWords), stream));
    int final_state = cur;
    MP_CALL(enqueue_steps(w, nsteps, scheme, dt, final_state, nullptr));
    return finish_step(w, nsteps, scheme, dt, final_state, report);
}

mp_status mp_transport::state_totals(double& th, double& tc)
{
    std::vector<double> h(N), c(N);
    MP_CALL(download(h.data(), c.data()));
    double s[2] = {0, 0}, e[2] = {0, 0};
    for (int f = 0; f < 2; ++f) {
        const std::vector<double>& u = f ? c : h;
        for (size_t k = 0; k < N; ++k) {
            const double y = s[f] + u[k];
            e[f] += std::fabs(s[f]) >= std::fabs(u[k]) ? (s[f] - y) + u[k] : (u[k] - y) + s[f];
            s[f] = y;
        }
    }
    th = (s[0] + e[0]) * md.area;
    tc = (s[1] + e[1]) * md.area;
    return MP_OK;
}

mp_status mp_transport::profile(int nsteps, int scheme, double dt, double* ms_total,
                                double* ms_stage, double* ms_phase_stage)
{
    MP_CALL(check_step_args(nsteps, scheme, dt));
    if (nsteps < 1) return fail(MP_ERR_ARG, "profile needs at least one step");
    MP_CALL(ensure_scratch(nsteps));
    std::vector<hipEvent_t> ev(size_t(nsteps) * 6);
    for (auto& e : ev) MP_HIP(hipEventCreate(&e));
    hipEvent_t t0, t1;
    MP_HIP(hipEventCreate(&t0));
    MP_HIP(hipEventCreate(&t1));
    MP_HIP(hipMemsetAsync(error.ptr, 0, sizeof(uint32_t), stream));
    int final_state = cur;
    MP_HIP(hipEventRecord(t0, stream));
    mp_status st = enqueue_steps(own_work(), nsteps, scheme, dt, final_state, &ev);
    MP_HIP(hipEventRecord(t1, stream));
    MP_HIP(hipStreamSynchronize(stream));
    double plain = 0, fused = 0;
    int np = 0, nf = 0;
    const bool one_launch = step_fused(scheme);
    for (int s = 0; s < nsteps && st == MP_OK; ++s)
        for (int k = 0; k < (one_launch ? 2 : scheme); ++k) {
            // one launch per step: pair 0 times the step's launch, pair 1 its phase
            float ms = 0;
            MP_HIP(hipEventElapsedTime(&ms, ev[2 * (s * 3 + k)], ev[2 * (s * 3 + k) + 1]));
            const bool second = one_launch ? k == 1 : k == scheme - 1;
            if (second) { fused += ms; ++nf; } else { plain += ms; ++np; }
        }
    float total = 0;
    MP_HIP(hipEventElapsedTime(&total, t0, t1));
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    MP_CALL(st);
    if (ms_total) *ms_total = total;
    if (ms_stage) *ms_stage = np ? plain / np : fused / nf;
    if (ms_phase_stage) *ms_phase_stage = fused / nf;
    return MP_OK;
}

mp_status mp_transport::phase_out(double* T, double* phi, double* Cl, double* Cs, uint8_t* region)
{
    if (!phase_valid) {
        double* base = phase_buf[phase_cur].as<double>();
        MP_HIP(hipMemsetAsync(error.ptr, 0, sizeof(uint32_t), stream));
        MP_HIP(launch_state_phase(d_md.as<MeshDev>(), int(N), H[cur].as<double>(), C[cur].as<double>(),
                                  base, base + N, base + 2 * N, base + 3 * N,
                                  reinterpret_cast<uint8_t*>(base + 4 * N),
                                  reinterpret_cast<uint32_t*>(error.ptr), stream));
        uint32_t bits = 0;
        MP_HIP(hipMemcpyAsync(&bits, error.ptr, sizeof bits, hipMemcpyDeviceToHost, stream));
        MP_HIP(hipStreamSynchronize(stream));
        if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
        phase_valid = true;
    }
    const double* base = phase_buf[phase_cur].as<double>();
    double* outs[4] = {T, phi, Cl, Cs};
    for (int k = 0; k < 4; ++k)
        if (outs[k]) MP_HIP(hipMemcpyAsync(outs[k], base + k * N, N * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (region) MP_HIP(hipMemcpyAsync(region, base + 4 * N, N, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status mp_transport::rhs(double* rh, double* rc, double* boundary)
{
    if (!have_velocity) return fail(MP_ERR_STATE, "set face velocities before evaluating the RHS");
    DeviceBuffer buf;
    MP_CALL(buf.alloc(N * 4 * sizeof(double)));
    double* d = buf.as<double>();
    MP_HIP(hipMemsetAsync(error.ptr, 0, sizeof(uint32_t), stream));
    MP_HIP(launch_rhs_diag(precision, d_md.as<MeshDev>(), int(N), H[cur].as<double>(), C[cur].as<double>(),
                           d, d + N, d + 2 * N, reinterpret_cast<uint32_t*>(error.ptr), stream));
    std::vector<double> host(N * 4);
    uint32_t bits = 0;
    MP_HIP(hipMemcpyAsync(host.data(), d, N * 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
    MP_HIP(hipMemcpyAsync(&bits, error.ptr, sizeof bits, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
    if (rh) std::memcpy(rh, host.data(), N * sizeof(double));
    if (rc) std::memcpy(rc, host.data() + N, N * sizeof(double));
    if (boundary) {
        long double b[2] = {0, 0};
        for (size_t k = 0; k < N; ++k) {
            b[0] += host[2 * N + 2 * k];
            b[1] += host[2 * N + 2 * k + 1];
        }
        boundary[0] = double(b[0]);
        boundary[1] = double(b[1]);
    }
    return MP_OK;
}

mp_status mp_transport::reconstruct(int field, double* sigma, double* omega, double* traces)
{
    if (field != 0 && field != 1) return fail(MP_ERR_ARG, "field must be 0 (H) or 1 (C)");
    const size_t nt = N * 4 * size_t(mesh.q);
    DeviceBuffer buf;
    MP_CALL(buf.alloc((N * 10 + nt) * sizeof(double)));
    double* d = buf.as<double>();
    MP_HIP(hipMemsetAsync(error.ptr, 0, sizeof(uint32_t), stream));
    MP_HIP(launch_recon_diag(precision, d_md.as<MeshDev>(), int(N),
                             field ? C[cur].as<double>() : H[cur].as<double>(), field, d,
                             d + 5 * N, d + 10 * N, reinterpret_cast<uint32_t*>(error.ptr), stream));
    std::vector<double> host(N * 10 + nt);
    uint32_t bits = 0;
    MP_HIP(hipMemcpyAsync(host.data(), d, host.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    MP_HIP(hipMemcpyAsync(&bits, error.ptr, sizeof bits, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
    if (sigma) std::memcpy(sigma, host.data(), 5 * N * sizeof(double));
    if (omega) std::memcpy(omega, host.data() + 5 * N, 5 * N * sizeof(double));
    if (traces) std::memcpy(traces, host.data() + 10 * N, nt * sizeof(double));
    return MP_OK;
}

mp_status mp_transport::evaluate(int64_t n, const int32_t* cells, const double* x, const double* y,
                                 double* outH, double* outC)
{
    if (n < 0 || (n > 0 && (!cells || !x || !y))) return fail(MP_ERR_ARG, "null argument");
    if (n == 0) return MP_OK;
    const int64_t ncells = int64_t(mesh.nx) * mesh.ny;
    // cell centres and h as the reference forms them: the mean of the cell's
    // (snapped) vertices, h = sqrt(area)
    std::vector<double> vx(size_t(mesh.nx) + 1), vy(size_t(mesh.ny) + 1);
    uniform_vertices(mesh.x0, mesh.x1, mesh.nx, vx.data());
    uniform_vertices(mesh.y0, mesh.y1, mesh.ny, vy.data());
    const double h = std::sqrt(md.area);
    std::vector<double> X(static_cast<size_t>(n)), Y(static_cast<size_t>(n));
    for (int64_t k = 0; k < n; ++k) {
        if (cells[k] < 0 || cells[k] >= ncells) return fail(MP_ERR_ARG, "Cell outside reconstruction halo");
        if (!std::isfinite(x[k]) || !std::isfinite(y[k])) return fail(MP_ERR_ARG, "Evaluation point must be finite");
        const int i = cells[k] % mesh.nx, j = cells[k] / mesh.nx;
        X[size_t(k)] = (x[k] - 0.5 * (vx[size_t(i)] + vx[size_t(i) + 1])) / h;
        Y[size_t(k)] = (y[k] - 0.5 * (vy[size_t(j)] + vy[size_t(j) + 1])) / h;
    }
    DeviceBuffer buf;
    const size_t nd = size_t(n) * sizeof(double);
    MP_CALL(buf.alloc(4 * nd + size_t(n) * sizeof(int32_t)));
    double* d = buf.as<double>();
    int32_t* dc = reinterpret_cast<int32_t*>(d + 4 * n);
    MP_HIP(hipMemcpyAsync(d, X.data(), nd, hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(d + n, Y.data(), nd, hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(dc, cells, size_t(n) * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemsetAsync(error.ptr, 0, sizeof(uint32_t), stream));
    MP_HIP(launch_point_eval(precision, d_md.as<MeshDev>(), H[cur].as<double>(), C[cur].as<double>(), n,
                             dc, d, d + n, d + 2 * n, d + 3 * n, reinterpret_cast<uint32_t*>(error.ptr),
                             stream));
    uint32_t bits = 0;
    std::vector<double> host(2 * size_t(n));
    MP_HIP(hipMemcpyAsync(host.data(), d + 2 * n, 2 * nd, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipMemcpyAsync(&bits, error.ptr, sizeof bits, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
    if (outH) std::memcpy(outH, host.data(), nd);
    if (outC) std::memcpy(outC, host.data() + n, nd);
    return MP_OK;
}

mp_status mp_transport::set_limiter(int enable, int cell_points, int samples, double extent)
{
    if (!enable) {
        md.limit = 0;
        return sync_md();
    }
    MP_CALL(set_limiter_tables(md, cell_points, samples, extent,
                               phase_params.hypereutectic ? 1.0 : phase_params.Xe));
    return sync_md();
}

mp_status mp_transport::limits(double* thetaH, double* thetaC)
{
    if (!md.limit) return fail(MP_ERR_STATE, "the bound limiter is off");
    DeviceBuffer buf;
    MP_CALL(buf.alloc(2 * N * sizeof(double)));
    double* d = buf.as<double>();
    MP_HIP(hipMemsetAsync(error.ptr, 0, sizeof(uint32_t), stream));
    MP_HIP(launch_limit_diag(d_md.as<MeshDev>(), int(N), H[cur].as<double>(), C[cur].as<double>(),
                             d, d + N, reinterpret_cast<uint32_t*>(error.ptr), stream));
    uint32_t bits = 0;
    MP_HIP(hipMemcpyAsync(&bits, error.ptr, sizeof bits, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
    if (thetaH) MP_HIP(hipMemcpyAsync(thetaH, d, N * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (thetaC) MP_HIP(hipMemcpyAsync(thetaC, d + N, N * sizeof(double), hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status mp_transport::set_diffusion(double kappa, int samples, double extent)
{
    MeshDev m = md;
    MP_CALL(set_diffusion_tables(m, kappa, samples, extent));
    md = m;
    MP_CALL(sync_md());
    if (have_velocity) MP_CALL(compute_rate());
    return MP_OK;
}

mp_status mp_transport::set_thermal_boundary(const uint8_t* types, const double* values)
{
    const size_t faces = size_t(2) * (mesh.nx + mesh.ny), total = faces * mesh.q;
    const size_t off[4] = {0, size_t(mesh.nx) * mesh.q, size_t(mesh.nx + mesh.ny) * mesh.q,
                           size_t(2 * mesh.nx + mesh.ny) * mesh.q};
    if (!types) {
        tbc.reset();
        tval.reset();
        for (int s = 0; s < 4; ++s) { md.tbc[s] = nullptr; md.tval[s] = nullptr; }
    } else {
        for (size_t k = 0; k < total; ++k)
            if (types[k] > 2) return fail(MP_ERR_ARG, "Unknown thermal boundary policy");
        std::vector<double> v(total, 0.0);
        if (values) {
            for (size_t k = 0; k < total; ++k)
                if (types[k] != 0 && !std::isfinite(values[k]))
                    return fail(MP_ERR_ARG, "Thermal boundary data must be finite");
            std::memcpy(v.data(), values, total * sizeof(double));
        } else {
            for (size_t k = 0; k < total; ++k)
                if (types[k] != 0) return fail(MP_ERR_ARG, "Missing thermal boundary profile.");
        }
        MP_CALL(tbc.alloc(total));
        MP_CALL(tval.alloc(total * sizeof(double)));
        MP_HIP(hipMemcpyAsync(tbc.ptr, types, total, hipMemcpyHostToDevice, stream));
        MP_HIP(hipMemcpyAsync(tval.ptr, v.data(), total * sizeof(double), hipMemcpyHostToDevice, stream));
        for (int s = 0; s < 4; ++s) {
            md.tbc[s] = static_cast<const uint8_t*>(tbc.ptr) + off[s];
            md.tval[s] = tval.as<double>() + off[s];
        }
    }
    MP_CALL(sync_md());
    if (have_velocity) MP_CALL(compute_rate());
    return MP_OK;
}

mp_status mp_transport::set_lithostatic(int enable, double surface_y)
{
    if (!enable) {
        tep.reset();
        md.tep_row = nullptr;
    } else {
        // the phase at a cell centre sees P = staticPressure(surface_y - y_c)
        std::vector<TepPair> rows;
        MP_CALL(lithostatic_rows(mesh, phase, surface_y, rows));
        MP_CALL(tep.alloc(rows.size() * sizeof(TepPair)));
        MP_HIP(hipMemcpyAsync(tep.ptr, rows.data(), tep.bytes, hipMemcpyHostToDevice, stream));
        md.tep_row = tep.as<TepPair>();
    }
    phase_valid = false;
    return sync_md();
}

// ===================================================================== C ABI

extern "C" {

void mp_default_weno_options(mp_weno_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->epsilon = 1e-2;
    o->s = 1;
    o->large_weight = 1.0;
    for (int k = 0; k < 4; ++k) o->small_weight[k] = 1.0;
    o->use_constant = 0;
    o->constant_linear_weight = 1e-5;
    o->lf_mode = 0;
    o->global_speed = 0.0;
}

void mp_default_phase_params(mp_phase_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->Tm0 = 2000.0; p->Te0 = 1480.0; p->nu = 6.5e6; p->L = 5.0e5; p->cp = 1200.0; p->Xe = 0.25;
    p->rho = 3000.0; p->rhor = 500.0; p->mus = 1.0e19; p->mul = 1.0; p->k0 = 1.0e-8; p->g = 10.0;
    p->alpha0 = 3.0e-5;
    p->pressure = 0.0;
    p->hypereutectic = 0;
}

const char* mp_last_error(void) { return mp::last_error(); }
mp_status mp_report_error(mp_status code, const char* message)
{
    return fail(code, "%s", message ? message : "");
}
const char* mp_version(void) { return "mantlepar_amd 0.1 (gfx950)"; }

mp_status mp_transport_create(const mp_mesh_desc* mesh, const mp_weno_options* weno,
                              const mp_phase_params* phase, int32_t precision, void* stream,
                              mp_transport** out)
{
    MP_GUARD({
        if (!mesh || !out) return fail(MP_ERR_ARG, "null argument");
        mp_weno_options w;
        mp_phase_params p;
        mp_default_weno_options(&w);
        mp_default_phase_params(&p);
        if (weno) w = *weno;
        if (phase) p = *phase;
        mp_transport* t = new mp_transport();
        mp_status s = t->init(*mesh, w, p, precision, stream);
        if (s) { delete t; return s; }
        *out = t;
        return MP_OK;
    })
}

mp_status mp_transport_destroy(mp_transport* t)
{
    delete t;
    return MP_OK;
}

mp_status mp_transport_upload_state(mp_transport* t, const double* H, const double* C)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->upload(H, C); })
}

mp_status mp_transport_download_state(mp_transport* t, double* H, double* C)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->download(H, C); })
}

mp_status mp_transport_set_face_velocity(mp_transport* t, const double* vnx, const double* vny)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->set_velocity(vnx, vny); })
}

mp_status mp_transport_set_boundary(mp_transport* t, const uint8_t* bottom, const uint8_t* right,
                                    const uint8_t* top, const uint8_t* left)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    const uint8_t* sides[4] = {bottom, right, top, left};
    MP_GUARD({ return t->set_boundary(sides); })
}

mp_status mp_transport_set_inflow(mp_transport* t, const double* inflow_H, const double* inflow_C)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->set_inflow(inflow_H, inflow_C); })
}

mp_status mp_transport_stable_dt(mp_transport* t, double cfl, double* dt)
{
    if (!t || !dt) return fail(MP_ERR_ARG, "null argument");
    if (!t->have_velocity) return fail(MP_ERR_STATE, "set face velocities first");
    if (!std::isfinite(cfl) || !(cfl > 0)) return fail(MP_ERR_ARG, "CFL must be finite and positive");
    if (!(t->max_rate > 0)) return fail(MP_ERR_STATE, "zero advective rate: the step is unbounded");
    *dt = cfl / t->max_rate;
    return MP_OK;
}

mp_status mp_transport_step(mp_transport* t, int32_t nsteps, int32_t scheme, double dt,
                            mp_step_report* report)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->step(nsteps, scheme, dt, report); })
}

mp_status mp_transport_step_noalloc(mp_transport* t, int32_t nsteps, int32_t scheme, double dt,
                                    mp_step_report* report)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->step_noalloc(nsteps, scheme, dt, report); })
}

mp_status mp_transport_phase(mp_transport* t, double* T, double* phi, double* Cl, double* Cs,
                             uint8_t* region)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->phase_out(T, phi, Cl, Cs, region); })
}

mp_status mp_transport_rhs(mp_transport* t, double* rhsH, double* rhsC, double* boundary)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->rhs(rhsH, rhsC, boundary); })
}

mp_status mp_transport_reconstruct(mp_transport* t, int32_t field, double* sigma, double* omega,
                                   double* traces)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->reconstruct(field, sigma, omega, traces); })
}

mp_status mp_transport_evaluate(mp_transport* t, int64_t n, const int32_t* cells, const double* x,
                                const double* y, double* H, double* C)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->evaluate(n, cells, x, y, H, C); })
}

mp_status mp_transport_device_state(mp_transport* t, double** H, double** C)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    if (H) *H = t->H[t->cur].as<double>();
    if (C) *C = t->C[t->cur].as<double>();
    return MP_OK;
}

mp_status mp_transport_synchronize(mp_transport* t)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_HIP(hipStreamSynchronize(t->stream));
    return MP_OK;
}

mp_status mp_transport_profile(mp_transport* t, int32_t nsteps, int32_t scheme, double dt,
                               double* ms_total, double* ms_stage, double* ms_phase_stage)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->profile(nsteps, scheme, dt, ms_total, ms_stage, ms_phase_stage); })
}

mp_status mp_transport_set_limiter(mp_transport* t, int32_t enable, int32_t cell_points_per_axis,
                                   int32_t diffusion_samples, double extent_fraction)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->set_limiter(enable, cell_points_per_axis, diffusion_samples, extent_fraction); })
}

mp_status mp_transport_set_diffusion(mp_transport* t, double kappa, int32_t diffusion_samples,
                                     double extent_fraction)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->set_diffusion(kappa, diffusion_samples, extent_fraction); })
}

mp_status mp_transport_set_thermal_boundary(mp_transport* t, const uint8_t* types,
                                            const double* values)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->set_thermal_boundary(types, values); })
}

mp_status mp_transport_set_lithostatic(mp_transport* t, int32_t enable, double surface_y)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->set_lithostatic(enable, surface_y); })
}

mp_status mp_transport_limits(mp_transport* t, double* thetaH, double* thetaC)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return t->limits(thetaH, thetaC); })
}

mp_status mp_transport_set_launch(mp_transport* t, int32_t mode)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    if (mode != MP_LAUNCH_PER_STAGE && mode != MP_LAUNCH_PER_STEP && mode != MP_LAUNCH_STALL_CHECK &&
        mode != MP_LAUNCH_AUTO)
        return fail(MP_ERR_ARG, "launch mode must be MP_LAUNCH_PER_STAGE, _PER_STEP, _STALL_CHECK or _AUTO");
    t->launch_mode = mode;
    return MP_OK;
}

mp_status mp_transport_layout(mp_transport* t, int32_t* rows_per_item, int32_t* items,
                              int32_t* tail_split, int32_t* one_launch_steps)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    if (rows_per_item) *rows_per_item = t->rows;
    if (items) *items = t->nitems;
    if (tail_split) *tail_split = t->layout.split;
    if (one_launch_steps) *one_launch_steps = t->step_fused(MP_SCHEME_SSPRK3) ? 1 : 0;
    return MP_OK;
}

mp_status mp_transport_set_balance_tolerance(mp_transport* t, double tol)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    if (!std::isfinite(tol) || tol < 0.0) return fail(MP_ERR_ARG, "balance tolerance must be finite and >= 0");
    t->balance_tol = tol;
    return MP_OK;
}

mp_status mp_transport_set_tile(mp_transport* t, int32_t rows_per_item)
{
    if (!t) return fail(MP_ERR_ARG, "null handle");
    if (rows_per_item < 4 || rows_per_item > 1 << 20) return fail(MP_ERR_ARG, "rows per item must be >= 4");
    t->set_tile(rows_per_item);
    return MP_OK;
}

mp_status mp_phase_evaluate(const mp_phase_params* params, int64_t n, const double* H,
                            const double* C, const double* P, double* T, double* phi, double* Cl,
                            double* Cs, uint8_t* region, double* phi1, double* phi2, double* dTdC,
                            double* dTdH)
{
    MP_GUARD({
        if (!params || n < 0 || (n > 0 && (!H || !C))) return fail(MP_ERR_ARG, "null argument");
        PhaseHost ph;
        MP_CALL(make_phase(*params, ph));
        if (n == 0) return MP_OK;
        hipStream_t s;
        MP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } guard{s};
        const size_t nb = size_t(n) * sizeof(double);
        DeviceBuffer in, out, err;
        MP_CALL(in.alloc(nb * 3));
        MP_CALL(out.alloc(nb * 8 + size_t(n)));
        MP_CALL(err.alloc(sizeof(uint32_t)));
        double* dH = in.as<double>();
        double* dC = dH + n;
        double* dP = P ? dC + n : nullptr;
        MP_HIP(hipMemcpyAsync(dH, H, nb, hipMemcpyHostToDevice, s));
        MP_HIP(hipMemcpyAsync(dC, C, nb, hipMemcpyHostToDevice, s));
        if (P) MP_HIP(hipMemcpyAsync(dP, P, nb, hipMemcpyHostToDevice, s));
        MP_HIP(hipMemsetAsync(err.ptr, 0, sizeof(uint32_t), s));
        double* o = out.as<double>();
        uint8_t* dreg = reinterpret_cast<uint8_t*>(o + 8 * n);
        MP_HIP(launch_phase(ph.dev, ph.TDe0, ph.nu, ph.dT, n, dH, dC, dP, o, o + n, o + 2 * n,
                            o + 3 * n, dreg, o + 4 * n, o + 5 * n, o + 6 * n, o + 7 * n,
                            err.as<uint32_t>(), s));
        uint32_t bits = 0;
        MP_HIP(hipMemcpyAsync(&bits, err.ptr, sizeof bits, hipMemcpyDeviceToHost, s));
        MP_HIP(hipStreamSynchronize(s));
        if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
        double* outs[8] = {T, phi, Cl, Cs, phi1, phi2, dTdC, dTdH};
        for (int k = 0; k < 8; ++k)
            if (outs[k]) MP_HIP(hipMemcpyAsync(outs[k], o + k * n, nb, hipMemcpyDeviceToHost, s));
        if (region) MP_HIP(hipMemcpyAsync(region, dreg, size_t(n), hipMemcpyDeviceToHost, s));
        MP_HIP(hipStreamSynchronize(s));
        return MP_OK;
    })
}

// EutecticModel's derive(), src/phase/eutectic_rescaled.cpp:42-86: the same
// checks in the same order (invalid input -> MP_ERR_ARG, like
// std::invalid_argument; an unrepresentable scale -> MP_ERR_FP, like
// std::overflow_error), long-double arithmetic rounded once per scale.
mp_status mp_phase_derive(const mp_phase_params* params, mp_phase_derived* out)
{
    if (!params || !out) return fail(MP_ERR_ARG, "null argument");
    const mp_phase_params& p = *params;
    auto finite = [](double v, const char* n) -> mp_status {
        return std::isfinite(v) ? mp_status(MP_OK) : fail(MP_ERR_ARG, "%s must be finite", n);
    };
    auto positive = [&](double v, const char* n) -> mp_status {
        MP_CALL(finite(v, n));
        return v > 0.0 ? mp_status(MP_OK) : fail(MP_ERR_ARG, "%s must be positive", n);
    };
    MP_CALL(positive(p.Tm0, "Tm0"));
    MP_CALL(positive(p.Te0, "Te0"));
    if (p.Tm0 <= p.Te0) return fail(MP_ERR_ARG, "Tm0 must be greater than Te0");
    MP_CALL(finite(p.nu, "nu"));
    if (p.nu == 0.0) return fail(MP_ERR_ARG, "nu must be nonzero");
    MP_CALL(positive(p.L, "L"));
    MP_CALL(positive(p.cp, "cp"));
    MP_CALL(finite(p.Xe, "Xe"));
    if (p.Xe <= 0.0 || p.Xe >= 1.0) return fail(MP_ERR_ARG, "Xe must lie strictly between 0 and 1");
    MP_CALL(positive(p.rho, "rho"));
    MP_CALL(positive(p.rhor, "rhor"));
    MP_CALL(positive(p.mus, "mus"));
    MP_CALL(positive(p.mul, "mul"));
    MP_CALL(positive(p.k0, "k0"));
    MP_CALL(positive(p.g, "g"));
    MP_CALL(finite(p.alpha0, "alpha0"));
    const long double dmax = DBL_MAX;
    auto checked = [&](long double v, const char* n, double& r) -> mp_status {
        if (!std::isfinite(v) || fabsl(v) > dmax) return fail(MP_ERR_FP, "%s is not representable", n);
        r = double(v);
        return MP_OK;
    };
    auto scale = [&](long double v, const char* n, double& r) -> mp_status {
        MP_CALL(checked(v, n, r));
        return r > 0.0 ? mp_status(MP_OK) : fail(MP_ERR_FP, "%s is not a positive scale", n);
    };
    mp_phase_derived d;
    const long double deltaT = (long double)p.Tm0 - p.Te0;
    const long double h0 = (long double)p.cp * deltaT;
    MP_CALL(scale(deltaT, "dT", d.dT));
    MP_CALL(scale(h0, "h0", d.h0));
    MP_CALL(checked(1.0L / p.nu, "gamma", d.gamma));
    if (d.gamma == 0.0) return fail(MP_ERR_FP, "gamma underflows to zero");
    MP_CALL(scale(p.L / h0, "LD", d.LD));
    MP_CALL(scale(p.Tm0 / deltaT, "TDm0", d.TDm0));
    MP_CALL(scale(p.Te0 / deltaT, "TDe0", d.TDe0));
    MP_CALL(scale(1.0L / p.k0, "invk0", d.invk0));
    const long double l0 = sqrtl((long double)p.mus * p.k0 / p.mul);
    const long double u0 = (long double)p.k0 * p.rhor * p.g / p.mul;
    MP_CALL(scale(l0, "l0", d.l0));
    MP_CALL(scale(u0, "u0", d.u0));
    MP_CALL(scale((long double)p.rhor * p.g * l0, "p0", d.p0));
    MP_CALL(scale(l0 / u0, "t0", d.t0));
    *out = d;
    return MP_OK;
}

// EutecticModel::staticPressure, src/phase/eutectic_rescaled.cpp:151-161.
mp_status mp_phase_static_pressure(const mp_phase_params* params, double zD, double length_scale,
                                   int32_t use_length, double* P)
{
    if (!params || !P) return fail(MP_ERR_ARG, "null argument");
    mp_phase_derived d;
    MP_CALL(mp_phase_derive(params, &d));
    const double len = use_length ? length_scale : d.l0;
    if (!std::isfinite(zD)) return fail(MP_ERR_ARG, "zD must be finite");
    if (!std::isfinite(len)) return fail(MP_ERR_ARG, "length_scale must be finite");
    if (!(len > 0.0)) return fail(MP_ERR_ARG, "length_scale must be positive");
    const long double v = (long double)params->rho * params->g * zD * len;
    if (!std::isfinite(v) || fabsl(v) > (long double)DBL_MAX)
        return fail(MP_ERR_FP, "static pressure is not representable");
    *P = double(v);
    return MP_OK;
}

// EutecticModel::pressureCorrectedTemperature, src/phase/eutectic_rescaled.cpp:143-149.
mp_status mp_phase_pressure_corrected_temperature(const mp_phase_params* params, double TD, double P,
                                                  double* out)
{
    if (!params || !out) return fail(MP_ERR_ARG, "null argument");
    mp_phase_derived d;
    MP_CALL(mp_phase_derive(params, &d));
    if (!std::isfinite(TD)) return fail(MP_ERR_ARG, "TD must be finite");
    if (!std::isfinite(P)) return fail(MP_ERR_ARG, "P must be finite");
    const long double v = (long double)TD + (long double)P / params->nu / d.dT;
    if (!std::isfinite(v) || fabsl(v) > (long double)DBL_MAX)
        return fail(MP_ERR_FP, "pressure-corrected temperature is not representable");
    *out = double(v);
    return MP_OK;
}

mp_status mp_phase_boundaries(const mp_phase_params* params, double C, double P, double* Tep,
                              double* Tmp, double* H_eutectic_end, double* H_liquidus)
{
    // EutecticModel::boundaries, src/phase/eutectic_rescaled.cpp:163-185 (host, long double)
    if (!params) return fail(MP_ERR_ARG, "null argument");
    mp_phase_params p = *params;
    p.pressure = P;
    PhaseHost ph;
    MP_CALL(make_phase(p, ph));
    if (!std::isfinite(C) || C < 0.0 || C > p.Xe) return fail(MP_ERR_COMPOSITION, "CD must lie in [0, Xe]");
    const long double ratio = (long double)C / p.Xe;
    const long double heut = (long double)ph.dev.Tep + ph.dev.LD * ratio;
    long double hliq = (long double)ph.dev.Tep + (1.0L - ratio) + ph.dev.LD;
    if (C == 0.0) hliq = (long double)ph.dev.Tmp + ph.dev.LD;
    if (!std::isfinite(heut) || !std::isfinite(hliq)) return fail(MP_ERR_FP, "boundary is not representable");
    const double he = double(heut), hl = double(hliq);
    if (ph.dev.Tmp <= ph.dev.Tep || hl <= ph.dev.Tep || (C == 0.0 && hl <= ph.dev.Tmp))
        return fail(MP_ERR_FP, "Phase temperature/enthalpy intervals are not resolvable");
    if (Tep) *Tep = ph.dev.Tep;
    if (Tmp) *Tmp = ph.dev.Tmp;
    if (H_eutectic_end) *H_eutectic_end = he;
    if (H_liquidus) *H_liquidus = hl;
    return MP_OK;
}

}  // extern "C"
