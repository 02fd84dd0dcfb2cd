// Multi-mesh engine: arena, slots, work lists and stage sequencing.
#include "mp_engine.h"

#include <cmath>
#include <algorithm>
#include <climits>
#include <cstring>

#include "mp_launch.h"

namespace mp {

// Stage tables as in the single-mesh handle.
static const double kA[3][3] = {{0, 0, 0}, {0, 0.5, 0}, {0, 0.75, 1.0 / 3.0}};
static const double kB[3][3] = {{1, 0, 0}, {1, 0.5, 0}, {1, 0.25, 2.0 / 3.0}};
static const double kW[3][3] = {{1, 0, 0}, {0.5, 0.5, 0}, {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}};

// ------------------------------------------------------------------ Arena

Arena::~Arena()
{
    for (auto& c : chunks_) (void)hipFree(c.base);
}

void Arena::configure(size_t budget_bytes, size_t chunk_bytes)
{
    budget_ = budget_bytes;
    chunk_ = chunk_bytes;
}

mp_status Arena::alloc(size_t bytes, void** out)
{
    const size_t need = round(bytes);
    if (in_use_ + need > budget_) return fail(MP_ERR_MEMORY, "HBM budget exhausted (%zu + %zu > %zu)", in_use_, need, budget_);
    for (size_t ci = 0; ci < chunks_.size(); ++ci) {
        Chunk& c = chunks_[ci];
        for (auto it = c.free.begin(); it != c.free.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, sz = it->second;
            c.free.erase(it);
            if (sz > need) c.free[off + need] = sz - need;
            void* p = c.base + off;
            live_[p] = {ci, need};
            in_use_ += need;
            *out = p;
            return MP_OK;
        }
    }
    Chunk c;
    const size_t chunk = chunk_ < budget_ ? chunk_ : budget_;
    c.size = need > chunk ? need : chunk;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&c.base), c.size);
    if (e != hipSuccess) return fail(MP_ERR_MEMORY, "hipMalloc(%zu) failed: %s", c.size, hipGetErrorString(e));
    if (c.size > need) c.free[need] = c.size - need;
    chunks_.push_back(c);
    void* p = chunks_.back().base;
    live_[p] = {chunks_.size() - 1, need};
    in_use_ += need;
    *out = p;
    return MP_OK;
}

void Arena::release(void* p)
{
    auto it = live_.find(p);
    if (it == live_.end()) return;
    Chunk& c = chunks_[it->second.first];
    size_t off = size_t(static_cast<char*>(p) - c.base), sz = it->second.second;
    in_use_ -= sz;
    live_.erase(it);
    // coalesce with neighbours
    auto next = c.free.lower_bound(off);
    if (next != c.free.end() && next->first == off + sz) {
        sz += next->second;
        next = c.free.erase(next);
    }
    if (next != c.free.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += sz;
            return;
        }
    }
    c.free[off] = sz;
}

// ------------------------------------------------------------- PinnedRing

PinnedRing::~PinnedRing()
{
    for (Buf& b : bufs_) {
        if (b.pending) (void)hipEventSynchronize(b.read);
        if (b.read) (void)hipEventDestroy(b.read);
        if (b.p) (void)hipHostFree(b.p);
    }
}

mp_status PinnedRing::next(size_t bytes, char** out)
{
    Buf& b = bufs_[cur_];
    if (b.pending) {
        MP_HIP(hipEventSynchronize(b.read));
        b.pending = false;
    }
    if (!b.read) MP_HIP(hipEventCreateWithFlags(&b.read, hipEventDisableTiming));
    if (b.cap < bytes) {
        if (b.p) MP_HIP(hipHostFree(b.p));
        b.p = nullptr;
        b.cap = 0;
        const size_t cap = std::max(bytes + bytes / 2, size_t(4096));
        MP_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.p), cap, hipHostMallocDefault));
        b.cap = cap;
    }
    *out = b.p;
    return MP_OK;
}

mp_status PinnedRing::issued(hipStream_t stream)
{
    Buf& b = bufs_[cur_];
    MP_HIP(hipEventRecord(b.read, stream));
    b.pending = true;
    cur_ = (cur_ + 1) % kBufs;
    return MP_OK;
}

// ----------------------------------------------------------------- Engine

Engine::~Engine()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    for (int k = 0; k < 2; ++k) {
        if (stage3[k]) (void)hipEventDestroy(stage3[k]);
        if (side_done[k]) (void)hipEventDestroy(side_done[k]);
    }
    if (side_join) (void)hipEventDestroy(side_join);
    if (side) (void)hipStreamDestroy(side);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    for (int k = 0; k < 2; ++k) {
        if (h_errors[k]) (void)hipHostFree(h_errors[k]);
        if (h_totals[k]) (void)hipHostFree(h_totals[k]);
    }
}

mp_status Engine::init(int cap, int32_t prec, const mp_weno_options& w, const mp_phase_params& p,
                       void* s, size_t budget_bytes)
{
    if (cap < 1) return fail(MP_ERR_ARG, "capacity must be positive");
    if (prec != MP_PREC_FP64 && prec != MP_PREC_MIXED) return fail(MP_ERR_ARG, "unknown precision");
    MP_CALL(validate_weno(w));
    MP_CALL(make_phase(p, phase));
    precision = prec;
    weno = w;
    phase_params = p;
    capacity = cap;
    if (s) {
        stream = static_cast<hipStream_t>(s);
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
    arena.configure(budget_bytes, size_t(1) << 30);
    slots.assign(size_t(cap), Slot());
    MP_CALL(d_meshes.alloc(sizeof(MeshDev) * cap));
    MP_CALL(d_stages.alloc(sizeof(StagePtrs) * 9 * cap));
    MP_CALL(d_nitems.alloc(sizeof(int) * cap));
    MP_CALL(d_active.alloc(sizeof(int) * cap));
    MP_CALL(d_errors.alloc(sizeof(uint32_t) * cap));
    MP_CALL(d_maxbits.alloc(sizeof(unsigned long long) * cap));
    MP_CALL(d_flag_off.alloc(sizeof(int) * cap));
    MP_CALL(d_init_totals.alloc(sizeof(double) * 4 * size_t(cap)));
    MP_CALL(d_init_part.alloc(sizeof(double) * 4 * size_t(kStateSlicesMax) * size_t(cap)));
    MP_CALL(d_epoch.alloc(256));
    MP_HIP(hipMemsetAsync(d_epoch.ptr, 0, 256, stream));
    MP_CALL(reserve_steps(1));
    h_nitems.assign(size_t(cap), 0);
    return MP_OK;
}

static size_t faces_x(const mp_mesh_desc& d) { return size_t(d.ny) * (d.nx + 1); }
static size_t faces_y(const mp_mesh_desc& d) { return size_t(d.ny + 1) * d.nx; }

// An upper bound on the work items of any layout with at least `rows` rows
// per item: tail items are 4 rows tall and number at most two bands' worth per
// wave slot, fewer than a 4-row layout of the whole mesh has.
static int items_for(const mp_mesh_desc& d, int rows)
{
    const int strips = (d.nx + kStripCells - 1) / kStripCells;
    return strips * ((d.ny + rows - 1) / rows) + strips * ((d.ny + 3) / 4);
}

size_t Engine::slot_bytes(const mp_mesh_desc& d, int32_t precision, int rows_min)
{
    const size_t N = size_t(d.nx) * d.ny;
    size_t b = 0;
    b += Arena::round(10 * N * sizeof(double));
    b += Arena::round(size_t(d.q) * faces_x(d) * sizeof(double));
    b += Arena::round(size_t(d.q) * faces_y(d) * sizeof(double));
    if (precision == MP_PREC_MIXED) {   // face moment records, full and one-sided
        b += Arena::round(faces_x(d) * sizeof(FaceMom)) + Arena::round(faces_y(d) * sizeof(FaceMom));
        b += Arena::round(faces_x(d) * sizeof(FaceRec)) + Arena::round(faces_y(d) * sizeof(FaceRec));
    } else {                            // one-sided FP64 records
        b += Arena::round(faces_x(d) * sizeof(FaceRec64)) + Arena::round(faces_y(d) * sizeof(FaceRec64));
    }
    b += 2 * Arena::round(N * (4 * sizeof(double) + 1));
    b += Arena::round(size_t(6) * items_for(d, rows_min) * kPartials * sizeof(double));
    b += Arena::round(size_t(2) * (d.nx + d.ny));
    b += Arena::round(size_t(4) * (d.nx + d.ny) * d.q * sizeof(double));
    b += Arena::round(size_t(d.ny) * sizeof(TepPair));
    b += Arena::round(size_t(2) * (d.nx + d.ny) * d.q);                    // thermal policies
    b += Arena::round(size_t(2) * (d.nx + d.ny) * d.q * sizeof(double));   // thermal data
    return b;
}

mp_status Engine::add_mesh(const mp_mesh_desc& d, int* out)
{
    int k = -1;
    for (int i = 0; i < capacity; ++i)
        if (!slots[i].live) { k = i; break; }
    if (k < 0) return fail(MP_ERR_MEMORY, "no free mesh slot (capacity %d)", capacity);
    Slot sl;
    MP_CALL(make_mesh_dev(d, weno, phase, sl.md));
    sl.desc = d;
    sl.N = size_t(d.nx) * d.ny;
    MP_CALL(arena.alloc(slot_bytes(d, precision), &sl.block));
    char* p = static_cast<char*>(sl.block);
    auto take = [&](size_t bytes) { char* r = p; p += Arena::round(bytes); return r; };
    const size_t N = sl.N;
    double* states = reinterpret_cast<double*>(take(10 * N * sizeof(double)));
    for (int i = 0; i < 5; ++i) {
        sl.H[i] = states + size_t(2 * i) * N;
        sl.C[i] = states + size_t(2 * i + 1) * N;
    }
    sl.md.vnx = reinterpret_cast<double*>(take(size_t(d.q) * faces_x(d) * sizeof(double)));
    sl.md.vny = reinterpret_cast<double*>(take(size_t(d.q) * faces_y(d) * sizeof(double)));
    if (precision == MP_PREC_MIXED) {
        sl.md.fx = reinterpret_cast<FaceMom*>(take(faces_x(d) * sizeof(FaceMom)));
        sl.md.fy = reinterpret_cast<FaceMom*>(take(faces_y(d) * sizeof(FaceMom)));
        sl.md.rx = reinterpret_cast<FaceRec*>(take(faces_x(d) * sizeof(FaceRec)));
        sl.md.ry = reinterpret_cast<FaceRec*>(take(faces_y(d) * sizeof(FaceRec)));
    } else {
        sl.md.rx64 = reinterpret_cast<FaceRec64*>(take(faces_x(d) * sizeof(FaceRec64)));
        sl.md.ry64 = reinterpret_cast<FaceRec64*>(take(faces_y(d) * sizeof(FaceRec64)));
    }
    sl.phase[0] = reinterpret_cast<double*>(take(N * (4 * sizeof(double) + 1)));
    sl.phase[1] = reinterpret_cast<double*>(take(N * (4 * sizeof(double) + 1)));
    sl.partials = reinterpret_cast<double*>(take(size_t(6) * items_for(d, 8) * kPartials * sizeof(double)));
    sl.bc = reinterpret_cast<uint8_t*>(take(size_t(2) * (d.nx + d.ny)));
    sl.inflow = reinterpret_cast<double*>(take(size_t(4) * (d.nx + d.ny) * d.q * sizeof(double)));
    sl.tep = reinterpret_cast<TepPair*>(take(size_t(d.ny) * sizeof(TepPair)));
    sl.tbc = reinterpret_cast<uint8_t*>(take(size_t(2) * (d.nx + d.ny) * d.q));
    sl.tval = reinterpret_cast<double*>(take(size_t(2) * (d.nx + d.ny) * d.q * sizeof(double)));
    sl.live = true;
    slots[k] = sl;
    set_rows(k, 64);
    mp_status s = MP_OK;
    // zero state until the first upload (as a fresh single-mesh handle); the
    // record is copied from pageable memory, which the copy reads before it
    // returns
    hipError_t e = hipMemsetAsync(sl.H[0], 0, 10 * N * sizeof(double), stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_meshes.as<MeshDev>() + k, &slots[k].md, sizeof(MeshDev),
                                  hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) s = fail(MP_ERR_DEVICE, "mesh upload: %s", hipGetErrorString(e));
    if (s) {
        remove_mesh(k);
        return s;
    }
    *out = k;
    return MP_OK;
}

void Engine::remove_mesh(int k)
{
    if (k < 0 || k >= capacity || !slots[k].live) return;
    arena.release(slots[k].block);
    slots[k] = Slot();
}

void Engine::set_rows(int k, int rows)
{
    Slot& s = slots[k];
    s.rows = rows < 8 ? 8 : (rows > s.desc.ny ? s.desc.ny : rows);
    if (s.rows < 8) s.rows = 8;
    s.layout = item_layout(s.desc.nx, s.desc.ny, s.rows, precision);
    s.strips = s.layout.strips;
    s.nitems = s.layout.nitems;
}

mp_status Engine::upload_async(int k, const double* H, const double* C, bool checked)
{
    Slot& s = slots[k];
    if (!H || !C) return fail(MP_ERR_ARG, "null state array");
    if (!checked && (!all_finite(H, s.N) || !all_finite(C, s.N)))
        return fail(MP_ERR_ARG, "Cell averages must be finite");
    MP_HIP(hipMemcpyAsync(s.H[s.cur], H, s.N * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(s.C[s.cur], C, s.N * sizeof(double), hipMemcpyHostToDevice, stream));
    s.phase_valid = false;
    return MP_OK;
}

mp_status Engine::upload(int k, const double* H, const double* C)
{
    MP_CALL(upload_async(k, H, C));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status Engine::download_async(int k, double* H, double* C)
{
    Slot& s = slots[k];
    if (H) MP_HIP(hipMemcpyAsync(H, s.H[s.cur], s.N * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (C) MP_HIP(hipMemcpyAsync(C, s.C[s.cur], s.N * sizeof(double), hipMemcpyDeviceToHost, stream));
    return MP_OK;
}

mp_status Engine::download(int k, double* H, double* C)
{
    MP_CALL(download_async(k, H, C));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status Engine::set_velocity(int k, const double* vx, const double* vy, bool wait, bool checked)
{
    Slot& s = slots[k];
    const mp_mesh_desc& d = s.desc;
    const size_t nxv = size_t(d.q) * faces_x(d), nyv = size_t(d.q) * faces_y(d);
    if (!vx || !vy) return fail(MP_ERR_ARG, "null velocity array");
    if (!checked && (!all_finite(vx, nxv) || !all_finite(vy, nyv)))
        return fail(MP_ERR_ARG, "Edge velocity must be finite");
    if (weno.lf_mode == 1) {
        double vmax = 0;
        for (size_t i = 0; i < nxv; ++i) vmax = std::fmax(vmax, std::fabs(vx[i]));
        for (size_t i = 0; i < nyv; ++i) vmax = std::fmax(vmax, std::fabs(vy[i]));
        if (weno.global_speed + 1e-14 < vmax)
            return fail(MP_ERR_ARG, "Global LF speed is smaller than the flow normal speed");
    }
    MP_HIP(hipMemcpyAsync(const_cast<double*>(s.md.vnx), vx, nxv * sizeof(double), hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(const_cast<double*>(s.md.vny), vy, nyv * sizeof(double), hipMemcpyHostToDevice, stream));
    MeshDev* dm = d_meshes.as<MeshDev>() + k;
    if (precision == MP_PREC_MIXED)
        MP_HIP(launch_face_prep(dm, d.nx, d.ny, const_cast<FaceMom*>(s.md.fx), const_cast<FaceMom*>(s.md.fy),
                                const_cast<FaceRec*>(s.md.rx), const_cast<FaceRec*>(s.md.ry), stream));
    else
        MP_HIP(launch_face_rec64(dm, d.nx, d.ny, const_cast<FaceRec64*>(s.md.rx64),
                                 const_cast<FaceRec64*>(s.md.ry64), stream));
    unsigned long long* bits = d_maxbits.as<unsigned long long>() + k;
    MP_CALL(launch_clear(reinterpret_cast<uint32_t*>(bits), 2, stream) == hipSuccess
                ? MP_OK : fail(MP_ERR_DEVICE, "clear failed"));
    MP_HIP(launch_rate(dm, int(s.N), bits, stream));
    if (wait) {
        unsigned long long host = 0;
        MP_HIP(hipMemcpyAsync(&host, bits, sizeof host, hipMemcpyDeviceToHost, stream));
        MP_HIP(hipStreamSynchronize(stream));
        std::memcpy(&s.max_rate, &host, sizeof s.max_rate);
    }
    s.have_velocity = true;
    return MP_OK;
}

mp_status Engine::set_boundary(int k, const uint8_t* const side[4])
{
    Slot& s = slots[k];
    const mp_mesh_desc& d = s.desc;
    const size_t len[4] = {size_t(d.nx), size_t(d.ny), size_t(d.nx), size_t(d.ny)};
    const size_t off[4] = {0, len[0], len[0] + len[1], len[0] + len[1] + len[2]};
    for (int f = 0; f < 4; ++f)
        if (side[f])
            for (size_t i = 0; i < len[f]; ++i)
                if (side[f][i] > MP_BC_INFLOW_OUTFLOW) return fail(MP_ERR_ARG, "Unknown advection boundary policy");
    for (int f = 0; f < 4; ++f) {
        if (!side[f]) {
            s.md.bc[f] = nullptr;
            continue;
        }
        MP_HIP(hipMemcpyAsync(s.bc + off[f], side[f], len[f], hipMemcpyHostToDevice, stream));
        s.md.bc[f] = s.bc + off[f];
    }
    MP_HIP(hipMemcpyAsync(d_meshes.as<MeshDev>() + k, &s.md, sizeof(MeshDev), hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status Engine::set_inflow(int k, const double* h, const double* c)
{
    Slot& s = slots[k];
    const mp_mesh_desc& d = s.desc;
    const size_t total = size_t(2) * (d.nx + d.ny) * d.q;
    const size_t off[4] = {0, size_t(d.nx) * d.q, size_t(d.nx + d.ny) * d.q, size_t(2 * d.nx + d.ny) * d.q};
    for (int f = 0; f < 2; ++f) {
        const double* src = f ? c : h;
        if (src && !all_finite(src, total))
            return fail(MP_ERR_FP, "Inflow boundary callback must return a finite state");
    }
    for (int f = 0; f < 2; ++f) {
        const double* src = f ? c : h;
        const double** dst = f ? s.md.inflowC : s.md.inflowH;
        double* buf = s.inflow + size_t(f) * total;
        if (!src) {
            for (int e = 0; e < 4; ++e) dst[e] = nullptr;
            continue;
        }
        MP_HIP(hipMemcpyAsync(buf, src, total * sizeof(double), hipMemcpyHostToDevice, stream));
        for (int e = 0; e < 4; ++e) dst[e] = buf + off[e];
    }
    MP_HIP(hipMemcpyAsync(d_meshes.as<MeshDev>() + k, &s.md, sizeof(MeshDev), hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

// Pressure model "lithostatic" for one mesh (src/couple/initialization.cpp:155-162):
// the phase of each cell at the static pressure of its centre.
mp_status Engine::set_lithostatic(int k, int enable, double surface_y)
{
    Slot& s = slots[k];
    if (enable) {
        MP_CALL(lithostatic_rows(s.desc, phase, surface_y, s.tep_host));
        MP_HIP(hipMemcpyAsync(s.tep, s.tep_host.data(), s.tep_host.size() * sizeof(TepPair),
                              hipMemcpyHostToDevice, stream));
    }
    s.md.tep_row = enable ? s.tep : nullptr;
    MP_HIP(hipMemcpyAsync(d_meshes.as<MeshDev>() + k, &s.md, sizeof(MeshDev), hipMemcpyHostToDevice,
                          stream));
    s.phase_valid = false;
    return MP_OK;
}

mp_status Engine::set_limiter(int k, int enable, int cell_points, int samples, double extent)
{
    Slot& s = slots[k];
    if (!enable) {
        s.md.limit = 0;
    } else {
        MeshDev md = s.md;
        MP_CALL(set_limiter_tables(md, cell_points, samples, extent,
                                   phase_params.hypereutectic ? 1.0 : phase_params.Xe));
        s.md = md;
    }
    MP_HIP(hipMemcpyAsync(d_meshes.as<MeshDev>() + k, &s.md, sizeof(MeshDev), hipMemcpyHostToDevice, stream));
    MP_HIP(hipStreamSynchronize(stream));
    ++generation;       // the launches' limiter instance may change
    return MP_OK;
}

mp_status Engine::refresh_rate(int k)
{
    Slot& s = slots[k];
    if (!s.have_velocity) return MP_OK;
    unsigned long long* bits = d_maxbits.as<unsigned long long>() + k;
    MP_HIP(launch_clear(reinterpret_cast<uint32_t*>(bits), 2, stream));
    MP_HIP(launch_rate(d_meshes.as<MeshDev>() + k, int(s.N), bits, stream));
    unsigned long long host = 0;
    MP_HIP(hipMemcpyAsync(&host, bits, sizeof host, hipMemcpyDeviceToHost, stream));
    MP_HIP(hipStreamSynchronize(stream));
    std::memcpy(&s.max_rate, &host, sizeof s.max_rate);
    return MP_OK;
}

mp_status Engine::set_diffusion(int k, double kappa, int samples, double extent)
{
    Slot& s = slots[k];
    MeshDev md = s.md;
    MP_CALL(set_diffusion_tables(md, kappa, samples, extent));
    s.md = md;
    MP_HIP(hipMemcpyAsync(d_meshes.as<MeshDev>() + k, &s.md, sizeof(MeshDev), hipMemcpyHostToDevice, stream));
    ++generation;
    return refresh_rate(k);
}

// Same checks and messages as mp_transport_set_thermal_boundary.
mp_status Engine::set_thermal_boundary(int k, const uint8_t* types, const double* values)
{
    Slot& s = slots[k];
    const mp_mesh_desc& d = s.desc;
    const size_t total = size_t(2) * (d.nx + d.ny) * d.q;
    const size_t off[4] = {0, size_t(d.nx) * d.q, size_t(d.nx + d.ny) * d.q, size_t(2 * d.nx + d.ny) * d.q};
    if (!types) {
        for (int f = 0; f < 4; ++f) { s.md.tbc[f] = nullptr; s.md.tval[f] = nullptr; }
    } else {
        for (size_t i = 0; i < total; ++i)
            if (types[i] > 2) return fail(MP_ERR_ARG, "Unknown thermal boundary policy");
        std::vector<double> v(total, 0.0);
        if (values) {
            for (size_t i = 0; i < total; ++i)
                if (types[i] != 0 && !std::isfinite(values[i]))
                    return fail(MP_ERR_ARG, "Thermal boundary data must be finite");
            std::memcpy(v.data(), values, total * sizeof(double));
        } else {
            for (size_t i = 0; i < total; ++i)
                if (types[i] != 0) return fail(MP_ERR_ARG, "Missing thermal boundary profile.");
        }
        MP_HIP(hipMemcpyAsync(s.tbc, types, total, hipMemcpyHostToDevice, stream));
        MP_HIP(hipMemcpyAsync(s.tval, v.data(), total * sizeof(double), hipMemcpyHostToDevice, stream));
        MP_HIP(hipStreamSynchronize(stream));     // v is a local
        for (int f = 0; f < 4; ++f) {
            s.md.tbc[f] = s.tbc + off[f];
            s.md.tval[f] = s.tval + off[f];
        }
    }
    MP_HIP(hipMemcpyAsync(d_meshes.as<MeshDev>() + k, &s.md, sizeof(MeshDev), hipMemcpyHostToDevice, stream));
    ++generation;
    return refresh_rate(k);
}

bool Engine::any_diffusion() const
{
    for (int k : active)
        if (has_diffusion(slots[k].md)) return true;
    return false;
}

int Engine::limit_mode() const
{
    int mode = 0;
    for (int k : active) {
        const int m = limit_mode32(slots[k].md);
        if (m == 1) return 1;
        if (m == 2) mode = 2;
    }
    return mode;
}

mp_status Engine::phase_out(int k, double* T, double* phi, double* Cl, double* Cs, uint8_t* region,
                            bool async)
{
    Slot& s = slots[k];
    if (!s.phase_valid) {
        double* base = s.phase[s.phase_cur];
        MP_HIP(launch_clear(d_errors.as<uint32_t>() + k, 1, stream));
        MP_HIP(launch_state_phase(d_meshes.as<MeshDev>() + k, int(s.N), s.H[s.cur], s.C[s.cur], base,
                                  base + s.N, base + 2 * s.N, base + 3 * s.N,
                                  reinterpret_cast<uint8_t*>(base + 4 * s.N),
                                  d_errors.as<uint32_t>() + k, stream));
        uint32_t bits = 0;
        MP_HIP(hipMemcpyAsync(&bits, d_errors.as<uint32_t>() + k, sizeof bits, hipMemcpyDeviceToHost, stream));
        MP_HIP(hipStreamSynchronize(stream));
        if (bits) return fail(status_from_bits(bits), "%s (device error bits 0x%x)", describe_bits(bits), bits);
        s.phase_valid = true;
    }
    const double* base = s.phase[s.phase_cur];
    double* outs[4] = {T, phi, Cl, Cs};
    for (int i = 0; i < 4; ++i)
        if (outs[i]) MP_HIP(hipMemcpyAsync(outs[i], base + i * s.N, s.N * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (region) MP_HIP(hipMemcpyAsync(region, base + 4 * s.N, s.N, hipMemcpyDeviceToHost, stream));
    if (!async) MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status Engine::set_active(const std::vector<int>& list, const std::vector<int>& steps)
{
    if (!steps.empty() && steps.size() != list.size()) return fail(MP_ERR_ARG, "steps and slots differ in length");
    std::vector<int> order(list.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
    if (!steps.empty())
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return steps[a] > steps[b]; });
    active.clear();
    active_steps.clear();
    slot_steps.assign(size_t(capacity), 0);
    q_common = -1;
    for (int i : order) {
        const int k = list[i];
        const Slot& s = slots[k];
        if (k < 0 || k >= capacity || !s.live) return fail(MP_ERR_STATE, "inactive slot %d in the work list", k);
        active.push_back(k);
        active_steps.push_back(steps.empty() ? INT32_MAX : steps[i]);
        slot_steps[k] = active_steps.back();
        h_nitems[k] = s.nitems;
        q_common = q_common < 0 ? s.desc.q : (q_common == s.desc.q ? q_common : 0);
    }
    if (q_common < 0) q_common = 0;
    // Two item lists per distinct prefix of the ordered slots (a prefix is the
    // set still stepping at some step), both over the same logical order:
    // items by decreasing cost, mesh after mesh among equal costs.
    //  * queue order: the logical order itself, for the persistent grid whose
    //    waves take items from a counter;
    //  * XCD slabs, for one wave per item: the dispatcher places workgroup b on
    //    XCD b % 8, so the logical order is cut into eight runs of equal cost
    //    and run x fills the positions b = x, x + 8, ...  A mesh's adjacent row
    //    bands, which share halo rows, then sit in one XCD's L2, and the
    //    waves of a CU mostly serve one mesh, whose tables stay in the scalar
    //    cache.  Runs of unequal length are padded with (-1, 0) entries that
    //    the kernels skip.  An item costs about its rows plus the two rows it
    //    reads before its first output row.
    // The order is built once for all active slots (a stable bucket sort on
    // the integer cost); a prefix's order is the subsequence of its slots,
    // which is what sorting the prefix alone would give.
    std::vector<int> prefixes;
    for (size_t i = 0; i < active.size(); ++i)
        if (i + 1 == active.size() || active_steps[i + 1] != active_steps[i]) prefixes.push_back(int(i + 1));
    struct Entry {
        int2 item;
        int rank, cost;
    };
    std::vector<Entry> all;
    int max_cost = 0;
    for (size_t i = 0; i < active.size(); ++i) {
        const Slot& sl = slots[active[i]];
        for (int it = 0; it < sl.nitems; ++it) {
            const bool tail = it >= sl.layout.nlong;
            const int band = tail ? (it - sl.layout.nlong) / sl.layout.strips : it / sl.layout.strips;
            const int j0 = tail ? sl.layout.split + band * sl.layout.tail_rows : band * sl.layout.rows;
            const int j1 = std::min(j0 + (tail ? sl.layout.tail_rows : sl.layout.rows),
                                    tail ? sl.desc.ny : sl.layout.split);
            const int cost = (j1 - j0) + 2;
            all.push_back({make_int2(active[i], it), int(i), cost});
            max_cost = std::max(max_cost, cost);
        }
    }
    std::vector<Entry> sorted(all.size());
    {
        // longest items first: the items a launch dispatches last, whose
        // durations set its drain, are the shortest
        std::vector<size_t> start(size_t(max_cost) + 2, 0);
        for (const Entry& e : all) ++start[size_t(max_cost - e.cost) + 1];
        for (size_t c = 1; c < start.size(); ++c) start[c] += start[c - 1];
        for (const Entry& e : all) sorted[start[size_t(max_cost - e.cost)]++] = e;
    }
    // completion flags of the one-launch step: per slot, a run of its items
    std::vector<int> flag_off(size_t(capacity), 0);
    nflags = 0;
    for (int k : active) {
        flag_off[size_t(k)] = nflags;
        nflags += slots[k].nitems;
    }
    std::vector<int2> items;
    lists.clear();
    for (int p : prefixes) {
        std::vector<int2> logical;
        std::vector<double> cost;
        double total = 0.0;
        for (const Entry& e : sorted) {
            if (e.rank >= p) continue;
            logical.push_back(e.item);
            cost.push_back(double(e.cost));
            total += double(e.cost);
        }
        List L;
        L.queue_off = int(items.size());
        L.queue_count = int(logical.size());
        items.insert(items.end(), logical.begin(), logical.end());
        int start[kXcds + 1], count[kXcds];
        {
            size_t pos = 0;
            double acc = 0.0;
            for (int x = 0; x < kXcds; ++x) {
                start[x] = int(pos);
                const double goal = total * double(x + 1) / kXcds;
                while (pos < logical.size() && (x == kXcds - 1 || acc + 0.5 * cost[pos] <= goal))
                    acc += cost[pos++];
                count[x] = int(pos) - start[x];
            }
            start[kXcds] = int(pos);
        }
        int per = 0;
        for (int x = 0; x < kXcds; ++x) per = std::max(per, count[x]);
        L.slab_off = int(items.size());
        L.slab_count = per * kXcds;
        for (int k = 0; k < per; ++k)
            for (int x = 0; x < kXcds; ++x)
                items.push_back(k < count[x] ? logical[size_t(start[x] + k)] : make_int2(-1, 0));
        lists[p] = L;
    }
    nitems_total = int(items.size());
    ++generation;
    // Uploaded from pinned memory without waiting: the copies run after the
    // work already queued on the stream, which still reads the old lists.
    // The item array grows geometrically (a reallocation synchronises).
    if (d_items.bytes < items.size() * sizeof(int2))
        MP_CALL(d_items.alloc(std::max(items.size() * sizeof(int2) + 1024, 2 * d_items.bytes)));
    MP_CALL(ensure_flags());
    const size_t b_items = items.size() * sizeof(int2), b_slots = sizeof(int) * size_t(capacity);
    char* h = nullptr;
    MP_CALL(ring_lists.next(b_items + 3 * b_slots, &h));
    if (b_items) std::memcpy(h, items.data(), b_items);
    std::memcpy(h + b_items, h_nitems.data(), b_slots);
    if (!active.empty()) std::memcpy(h + b_items + b_slots, active.data(), sizeof(int) * active.size());
    std::memcpy(h + b_items + 2 * b_slots, flag_off.data(), b_slots);
    if (b_items) MP_HIP(hipMemcpyAsync(d_items.ptr, h, b_items, hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(d_nitems.ptr, h + b_items, b_slots, hipMemcpyHostToDevice, stream));
    if (!active.empty())
        MP_HIP(hipMemcpyAsync(d_active.ptr, h + b_items + b_slots, sizeof(int) * active.size(),
                              hipMemcpyHostToDevice, stream));
    MP_HIP(hipMemcpyAsync(d_flag_off.ptr, h + b_items + 2 * b_slots, b_slots, hipMemcpyHostToDevice, stream));
    MP_CALL(ring_lists.issued(stream));
    return MP_OK;
}

mp_status Engine::ensure_flags()
{
    const size_t need = sizeof(uint32_t) * 2 * size_t(std::max(nflags, 1));
    if (d_flags.bytes >= need) return MP_OK;
    MP_CALL(d_flags.alloc(std::max(need, 2 * d_flags.bytes)));
    MP_HIP(hipMemsetAsync(d_flags.ptr, 0, d_flags.bytes, stream));
    ++generation;
    return MP_OK;
}

int Engine::auto_step_items()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return 8 * cus * 4 * 3;
}

mp_status Engine::reserve_epochs(int nsteps)
{
    // flags hold counter values 1 .. 2^32 - 1; restart well before a wrap
    if (epochs_used + uint64_t(std::max(nsteps, 0)) >= (uint64_t(1) << 31)) {
        MP_HIP(hipMemsetAsync(d_flags.ptr, 0, d_flags.bytes, stream));
        MP_HIP(hipMemsetAsync(d_epoch.ptr, 0, sizeof(uint32_t), stream));
        epochs_used = 0;
    }
    epochs_used += uint64_t(std::max(nsteps, 0));
    return MP_OK;
}

int Engine::steps_of(int k) const
{
    return k >= 0 && k < int(slot_steps.size()) ? slot_steps[k] : 0;
}

// Three pointer sets per slot: step 0 (cur -> o1), odd steps (o1 -> o2),
// even steps >= 2 (o2 -> o1); each with the scheme's stages.  Odd steps use
// the second set of stage partials, so a step's totals can be reduced beside
// the next step.
mp_status Engine::write_tables(int scheme)
{
    const size_t bytes = sizeof(StagePtrs) * 9 * size_t(capacity);
    char* h = nullptr;
    MP_CALL(ring_tables.next(bytes, &h));
    StagePtrs* h_stages = reinterpret_cast<StagePtrs*>(h);
    for (int k : active) {
        Slot& s = slots[k];
        int o[2], n = 0;
        for (int i = 0; i < 3; ++i)
            if (i != s.cur) o[n++] = i;
        const int src_of[3] = {s.cur, o[0], o[1]};
        const int dst_of[3] = {o[0], o[1], o[0]};
        for (int set = 0; set < 3; ++set)
            for (int st = 0; st < 3; ++st) {
                StagePtrs sp;
                std::memset(&sp, 0, sizeof sp);
                const int src = src_of[set], dst = dst_of[set];
                const bool last = st == scheme - 1;
                const int kin = st == 0 ? src : 3 + (st - 1);
                const int out = last ? dst : 3 + st;
                sp.Hn = s.H[src]; sp.Cn = s.C[src];
                sp.Hk = s.H[kin]; sp.Ck = s.C[kin];
                sp.Ho = s.H[out]; sp.Co = s.C[out];
                sp.a = st < scheme ? kA[scheme - 1][st] : 0.0;
                sp.b = st < scheme ? kB[scheme - 1][st] : 0.0;
                sp.dt = s.dt;
                if (last) {
                    double* base = s.phase[1 - s.phase_cur];
                    sp.T = base; sp.phi = base + s.N; sp.Cl = base + 2 * s.N; sp.Cs = base + 3 * s.N;
                    sp.region = reinterpret_cast<uint8_t*>(base + 4 * s.N);
                }
                sp.partials = s.partials + size_t((set == 1 ? 3 : 0) + st) * s.nitems * kPartials;
                sp.error = d_errors.as<uint32_t>() + k;
                apply_layout(s.layout, sp);
                h_stages[size_t(set * 3 + st) * capacity + k] = sp;
            }
    }
    MP_HIP(hipMemcpyAsync(d_stages.ptr, h_stages, bytes, hipMemcpyHostToDevice, stream));
    MP_CALL(ring_tables.issued(stream));
    return MP_OK;
}

mp_status Engine::reserve_steps(int nsteps)
{
    if (nsteps <= steps_capacity) return MP_OK;
    MP_CALL(d_totals.alloc(sizeof(double) * 4 * size_t(capacity) * nsteps));
    MP_CALL(d_queues.alloc(sizeof(int) * 3 * size_t(nsteps)));
    steps_capacity = nsteps;
    ++generation;
    return MP_OK;
}

// Each step's phase kernel and per-mesh totals run on the side stream,
// beside the next step's stages: the phase reads the step's
// output states, which the next step only reads, and writes the phase outputs
// and the sum slots of the step's partials (the parity set of its step).
// Step s + 2 writes step s's output states and partial set again, so it first
// waits for step s's phase and totals; the main stream waits for the side
// stream at the end, error returns included.  The fork and join are events,
// so a captured graph holds them too.  Profiling (events) keeps everything on
// the main stream, so each launch is timed alone.
mp_status Engine::enqueue(int nsteps, int scheme, std::vector<hipEvent_t>* events)
{
    if (nsteps > steps_capacity) return fail(MP_ERR_STATE, "per-step totals not reserved");
    const bool overlap = side && !events;
    struct Join {
        hipStream_t main, side;
        hipEvent_t ev;
        bool on;
        ~Join()
        {
            if (on && hipEventRecord(ev, side) == hipSuccess) (void)hipStreamWaitEvent(main, ev, 0);
        }
    } join{stream, side, side_join, overlap && nsteps > 0};
    MP_HIP(launch_clear(d_errors.as<uint32_t>(), capacity, stream));
    if (blocks > 0) MP_HIP(launch_clear(d_queues.as<uint32_t>(), 3 * nsteps, stream));
    const StagePtrs* base = d_stages.as<StagePtrs>();
    // the first step's stage-0 records name each mesh's starting state
    if (nsteps > 0 && !active.empty())
        MP_HIP(launch_batch_state_sums(base, d_meshes.as<MeshDev>(), d_active.as<int>(), int(active.size()),
                                       d_init_totals.as<double>(), d_init_part.as<double>(), stream));
    for (int s = 0; s < nsteps; ++s) {
        const int set = s == 0 ? 0 : (s % 2 ? 1 : 2);
        int p = 0;      // slots still stepping: a prefix of the ordered list
        while (p < int(active.size()) && active_steps[p] > s) ++p;
        if (p == 0) {
            if (events)
                for (int st = 0; st < scheme; ++st) {
                    MP_HIP(hipEventRecord((*events)[2 * (s * 3 + st)], stream));
                    MP_HIP(hipEventRecord((*events)[2 * (s * 3 + st) + 1], stream));
                }
            continue;
        }
        const List& Lst = lists.at(p);
        const std::pair<int, int> L = blocks > 0 ? std::make_pair(Lst.queue_off, Lst.queue_count)
                                                 : std::make_pair(Lst.slab_off, Lst.slab_count);
        const int par = set == 1 ? 1 : 0;    // the partial set of this step
        if (overlap && s >= 2) MP_HIP(hipStreamWaitEvent(stream, side_done[par], 0));
        const bool one = one_launch_for(L.second, events != nullptr);
        const int lim = limit_mode();
        const bool dif = any_diffusion();
        if (one)
            MP_HIP(launch_batch_step32(d_meshes.as<MeshDev>(), base + size_t(set * 3) * capacity, capacity,
                                       d_items.as<int2>() + L.first, L.second, scheme, d_flags.as<uint32_t>(),
                                       d_flag_off.as<int>(), nflags, d_epoch.as<uint32_t>(), kSpinLimit,
                                       weno.lf_mode != 0, stream));
        for (int st = 0; st < scheme; ++st) {
            // both sweeps leave the phase to its own streaming kernel, as the
            // single-mesh handle does
            const bool last = st == scheme - 1;
            const StagePtrs* sps = base + size_t(set * 3 + st) * capacity;
            if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3 + st)], stream));
            if (!one)
                MP_HIP(launch_batch_stage(precision, q_common, d_meshes.as<MeshDev>(), sps,
                                          d_items.as<int2>() + L.first, L.second, blocks,
                                          d_queues.as<int>() + 3 * s + st, stream, weno.lf_mode != 0, lim,
                                          dif));
            if (last && overlap) {
                MP_HIP(hipEventRecord(stage3[par], stream));
                MP_HIP(hipStreamWaitEvent(side, stage3[par], 0));
                MP_HIP(launch_batch_phase_items(d_meshes.as<MeshDev>(), sps, d_items.as<int2>() + L.first,
                                                L.second, side));
            } else if (last) {
                MP_HIP(launch_batch_phase_items(d_meshes.as<MeshDev>(), sps, d_items.as<int2>() + L.first,
                                                L.second, stream));
            }
            if (events) MP_HIP(hipEventRecord((*events)[2 * (s * 3 + st) + 1], stream));
        }
        const double* prev = s == 0 ? d_init_totals.as<double>()
                                    : d_totals.as<double>() + size_t(s - 1) * 4 * capacity;
        MP_HIP(launch_batch_reduce(base + size_t(set * 3) * capacity, d_nitems.as<int>(),
                                   d_meshes.as<MeshDev>(), d_active.as<int>(), p, scheme,
                                   kW[scheme - 1], d_totals.as<double>() + size_t(s) * 4 * capacity,
                                   overlap ? side : stream, prev, balance_tol));
        if (overlap) MP_HIP(hipEventRecord(side_done[par], side));
    }
    return MP_OK;
}

mp_status Engine::fetch_status(int nsteps, std::vector<uint32_t>& errors, std::vector<double>* totals)
{
    errors.assign(size_t(capacity), 0u);
    MP_HIP(hipMemcpyAsync(errors.data(), d_errors.ptr, sizeof(uint32_t) * capacity, hipMemcpyDeviceToHost, stream));
    if (totals && nsteps > 0) {
        totals->assign(size_t(4) * capacity * nsteps, 0.0);
        MP_HIP(hipMemcpyAsync(totals->data(), d_totals.ptr, sizeof(double) * totals->size(),
                              hipMemcpyDeviceToHost, stream));
    }
    MP_HIP(hipStreamSynchronize(stream));
    return MP_OK;
}

mp_status Engine::copy_status(int set, int nsteps)
{
    if (set < 0 || set > 1) return fail(MP_ERR_ARG, "status set must be 0 or 1");
    if (nsteps > steps_capacity) return fail(MP_ERR_STATE, "per-step totals not reserved");
    if (!h_errors[set]) MP_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_errors[set]),
                                             sizeof(uint32_t) * capacity, hipHostMallocDefault));
    if (status_steps < steps_capacity) {
        // both sets are reallocated together; a caller reserves its steps
        // before it keeps copies in flight
        for (int k = 0; k < 2; ++k) {
            if (h_totals[k]) MP_HIP(hipHostFree(h_totals[k]));
            h_totals[k] = nullptr;
            MP_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_totals[k]),
                                 sizeof(double) * 4 * size_t(capacity) * steps_capacity, hipHostMallocDefault));
        }
        status_steps = steps_capacity;
    }
    MP_HIP(hipMemcpyAsync(h_errors[set], d_errors.ptr, sizeof(uint32_t) * capacity, hipMemcpyDeviceToHost,
                          stream));
    if (nsteps > 0)
        MP_HIP(hipMemcpyAsync(h_totals[set], d_totals.ptr, sizeof(double) * 4 * size_t(capacity) * nsteps,
                              hipMemcpyDeviceToHost, stream));
    return MP_OK;
}

void Engine::make_report(int k, int nsteps, const std::vector<double>& tot, uint32_t bits,
                         mp_step_report* r) const
{
    const Slot& s = slots[k];
    std::memset(r, 0, sizeof *r);
    r->dt = s.dt;
    r->max_rate = s.max_rate;
    r->courant = s.dt * s.max_rate;
    r->error_bits = bits;
    r->steps_done = bits ? 0 : nsteps;
    if (bits || nsteps <= 0) return;
    long double bh = 0, bc = 0;
    for (int st = 0; st < nsteps; ++st) {
        const double* t = tot.data() + (size_t(st) * capacity + k) * 4;
        bh += (long double)s.dt * t[2];
        bc += (long double)s.dt * t[3];
    }
    r->boundary_H = double(bh);
    r->boundary_C = double(bc);
    const double* t = tot.data() + (size_t(nsteps - 1) * capacity + k) * 4;
    r->total_H = t[0];
    r->total_C = t[1];
}

void Engine::commit(int nsteps, const std::vector<uint32_t>& errors)
{
    if (nsteps <= 0) return;
    for (size_t i = 0; i < active.size(); ++i) {
        const int k = active[i];
        const int done = std::min(nsteps, active_steps[i]);
        Slot& s = slots[k];
        if (errors[k] || done <= 0) continue;
        int o[2], n = 0;
        for (int c = 0; c < 3; ++c)
            if (c != s.cur) o[n++] = c;
        s.cur = (done - 1) % 2 ? o[1] : o[0];
        s.phase_cur = 1 - s.phase_cur;
        s.phase_valid = true;
    }
}

}  // namespace mp
