// Serving front end: continuous batching of transport runs on one device.
//
// Callers submit requests from any thread.  Stager threads copy each
// request's inputs, in arrival order, into pinned host blocks and reject
// non-finite values on the way.  One server thread owns the engine and keeps
// two epochs queued on the device: while they run it admits staged requests
// (arrival order, first fit, while slots and the HBM budget allow; uploads on
// a second stream, not waited for: a run joins the epochs once its uploads
// are seen complete), then waits for the older epoch, reads its status (copied
// on the stream into pinned memory), retires runs that reached their step
// count and queues the next epoch, which advances every resident run by up to
// E steps (one launch per stage over all of them).  A run's results are
// independent of its co-residents: each mesh has its own tables, dt,
// partials and error word, and the stage kernel's per-mesh arithmetic does
// not depend on the work list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "mp_engine.h"

using namespace mp;

namespace {

struct Request {
    int64_t id = 0;
    mp_request r{};
    size_t bytes = 0;
    int slot = -1;
    int rows_now = 12;        // rows per work item, fixed at admission
    int remaining = 0;        // steps not yet queued
    int inflight = 0;         // queued epochs that advance the run
    long double bH = 0, bC = 0;
    mp_request_result res{};
    char message[256] = {0};
    bool done = false;
    mp_status fail_st = MP_OK;   // a failure found while later epochs were queued
    // Inputs staged in pinned memory: H, C (N each), then the vertical and
    // horizontal face velocities.
    enum { kWaiting, kStaging, kStaged } stage = kWaiting;
    char* staged = nullptr;
    size_t staged_bytes = 0;
    mp_status stage_st = MP_OK;
    const char* stage_msg = nullptr;
};

size_t staged_size(const mp_mesh_desc& d)
{
    const size_t N = size_t(d.nx) * d.ny;
    const size_t fx = size_t(d.ny) * (d.nx + 1), fy = size_t(d.ny + 1) * d.nx;
    return sizeof(double) * (2 * N + size_t(d.q) * (fx + fy));
}

// Copies n values and reports whether all are finite (one pass over each).
bool copy_finite(double* dst, const double* src, size_t n)
{
    uint64_t bad = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t b;
        std::memcpy(&b, src + i, sizeof b);
        bad |= uint64_t((b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull);
        std::memcpy(dst + i, &b, sizeof b);
    }
    return bad == 0;
}

// Pinned host blocks for staged inputs, reused by size class (powers of
// two): pinned allocations are slow, copies from them are not.
class PinnedPool {
public:
    ~PinnedPool()
    {
        for (auto& kv : blocks_) (void)hipHostFree(kv.first);
    }
    char* get(size_t bytes, size_t* cap)
    {
        size_t c = 4096;
        while (c < bytes) c <<= 1;
        *cap = c;
        std::lock_guard<std::mutex> g(mu_);
        auto it = free_.find(c);
        if (it != free_.end()) {
            char* p = it->second;
            free_.erase(it);
            return p;
        }
        char* p = nullptr;
        if (hipHostMalloc(reinterpret_cast<void**>(&p), c, hipHostMallocDefault) != hipSuccess) return nullptr;
        blocks_[p] = c;
        return p;
    }
    void put(char* p, size_t cap)
    {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        free_.emplace(cap, p);
    }

private:
    std::mutex mu_;
    std::multimap<size_t, char*> free_;
    std::map<char*, size_t> blocks_;
};

}  // namespace

struct mp_server {
    Engine eng;
    int scheme = MP_SCHEME_SSPRK3;
    int max_epoch = 8;
    int device = 0;
    int rows = 12;                     // rows per work item of every admitted run (the batch autotuner's usual pick)
    std::vector<int> last_list, last_steps;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_stage;
    std::deque<Request*> pending;      // submitted, not admitted (arrival order)
    std::deque<Request*> to_stage;     // submitted, not yet taken by a stager
    std::map<int64_t, Request*> all;   // submitted and not yet collected by mp_server_wait
    // Latencies (ms) of the most recent completed runs, a ring of kLatencies.
    static constexpr size_t kLatencies = size_t(1) << 16;
    std::vector<double> latencies;
    size_t latency_next = 0;
    // (completion time, cumulative cell-steps) of the most recent epochs, a ring
    std::vector<std::pair<double, double>> progress;
    size_t progress_next = 0;
    double work_done = 0;
    std::vector<Request*> resident;
    int64_t next_id = 1, submitted = 0, completed = 0, failed = 0;
    int64_t epochs = 0, max_resident = 0;
    uint64_t hbm_in_use = 0, hbm_peak = 0;
    double cell_steps = 0, first_submit = -1, last_done = 0;
    bool stop = false;
    bool membership_changed = true;   // a slot was filled or freed since the last work list
    std::thread worker;
    // Stagers copy at most kStageAhead bytes ahead of admission.
    static constexpr int kStagers = 4;
    static constexpr size_t kStageAhead = size_t(512) << 20;
    size_t staged_ahead = 0;
    int staging = 0;                   // requests a stager is copying now
    std::vector<std::thread> stagers;
    PinnedPool pool;
    // Requests are admitted on `upload` while the queued epochs run on the
    // engine's stream; an admitted run joins an epoch only after its uploads
    // are seen complete.  Results of finished runs come back on `download`.
    hipStream_t upload = nullptr, download = nullptr;
    // Admissions whose uploads are still on their way, oldest first: their
    // runs hold slots and memory, and become resident once `done` has
    // completed (their advective rates are then in the pinned `bits`).  The
    // worker does not wait for the uploads, so it keeps the epochs queued.
    struct Admission {
        hipEvent_t done = nullptr;
        unsigned long long* bits = nullptr;   // [capacity], pinned
        std::vector<Request*> runs;
    };
    std::deque<Admission> admitting;
    std::vector<unsigned long long*> spare_bits;
    std::vector<hipEvent_t> spare_done;   // the worker's own; spare_events is shared with the finisher
    int admitting_runs = 0;
    // Queued epochs (at most two), oldest first; epoch i uses status set i % 2.
    struct Flight {
        int E = 0, set = 0;
        mp_status st = MP_OK;
        std::vector<Request*> run;
        std::vector<int> steps;        // steps of each run in this epoch
    };
    std::deque<Flight> flights;
    hipEvent_t flight_done[2] = {nullptr, nullptr};
    int next_set = 0;
    // Finished runs whose outputs are on their way into pinned blocks: the
    // finisher thread waits for the copies, moves the outputs into the
    // caller's arrays and reports the runs.
    struct Retired {
        hipEvent_t copied = nullptr;
        std::vector<Request*> runs;
        std::vector<char*> blocks;
        std::vector<size_t> caps;
    };
    std::deque<Retired> to_finish;
    std::vector<hipEvent_t> spare_events;
    std::condition_variable cv_finish;
    bool finisher_stop = false;
    std::thread finisher;

    ~mp_server()
    {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        cv_stage.notify_all();
        if (worker.joinable()) worker.join();
        {
            std::lock_guard<std::mutex> g(mu);
            finisher_stop = true;
        }
        cv_finish.notify_all();
        if (finisher.joinable()) finisher.join();
        for (std::thread& t : stagers)
            if (t.joinable()) t.join();
        for (hipEvent_t e : spare_events) (void)hipEventDestroy(e);
        for (unsigned long long* b : spare_bits) (void)hipHostFree(b);
        for (hipEvent_t e : spare_done) (void)hipEventDestroy(e);
        for (auto& kv : all) delete kv.second;
        for (hipEvent_t e : flight_done)
            if (e) (void)hipEventDestroy(e);
        if (upload) (void)hipStreamDestroy(upload);
        if (download) (void)hipStreamDestroy(download);
    }

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void loop();
    void stage_loop();
    void finish_loop();
    void admit(std::vector<Request*>& batch, int rows_now);
    void promote(bool wait);
    bool start_epoch();
    void finish_oldest();
    void retire();
    void finish(Request* q, mp_status st, const char* msg) { release(q); report(q, st, msg); }
    void release(Request* q);
    void report(Request* q, mp_status st, const char* msg);
    void unstage(Request* q);
};

// Returns the request's pinned block to the pool (server thread, or the
// destructor after the threads have ended).
void mp_server::unstage(Request* q)
{
    if (!q->staged) return;
    pool.put(q->staged, q->staged_bytes);
    q->staged = nullptr;
    std::lock_guard<std::mutex> g(mu);
    staged_ahead -= q->staged_bytes;
    cv_stage.notify_all();
}

// Frees the run's slot (server thread).
void mp_server::release(Request* q)
{
    if (q->slot >= 0) {
        eng.remove_mesh(q->slot);
        q->slot = -1;
        membership_changed = true;
    }
}

// Publishes the run's result (server or finisher thread).
void mp_server::report(Request* q, mp_status st, const char* msg)
{
    unstage(q);
    std::lock_guard<std::mutex> g(mu);
    q->res.status = st;
    if (msg) std::snprintf(q->message, sizeof q->message, "%s", msg);
    q->res.t_done = now();
    q->done = true;
    if (st == MP_OK) {
        ++completed;
        cell_steps += double(q->r.mesh.nx) * q->r.mesh.ny * q->res.steps_done;
    } else {
        ++failed;
    }
    last_done = q->res.t_done;
    if (st == MP_OK) {
        const double ms = 1e3 * (q->res.t_done - q->res.t_submit);
        if (latencies.size() < kLatencies) latencies.push_back(ms);
        else latencies[latency_next] = ms;
        latency_next = (latency_next + 1) % kLatencies;
    }
    q->r.H = q->r.C = q->r.vnx = q->r.vny = nullptr;
    cv_done.notify_all();
}

// Stager thread: takes requests in arrival order and copies their inputs
// into a pinned block, checking them as the engine's upload does (the
// reference rejects non-finite averages and velocities).  It stays at most
// kStageAhead bytes ahead of admission.
void mp_server::stage_loop()
{
    for (;;) {
        Request* q = nullptr;
        size_t need = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_stage.wait(lk, [&] { return stop || (!to_stage.empty() && staged_ahead < kStageAhead); });
            if (stop) return;
            q = to_stage.front();
            to_stage.pop_front();
            q->stage = Request::kStaging;
            ++staging;
            need = staged_size(q->r.mesh);
            staged_ahead += need;
        }
        size_t cap = 0;
        char* p = pool.get(need, &cap);
        mp_status st = MP_OK;
        const char* msg = nullptr;
        if (!p) {
            st = MP_ERR_MEMORY;
            msg = "pinned staging allocation failed";
        } else {
            const mp_mesh_desc& d = q->r.mesh;
            const size_t N = size_t(d.nx) * d.ny;
            const size_t nxv = size_t(d.q) * d.ny * (d.nx + 1), nyv = size_t(d.q) * (d.ny + 1) * d.nx;
            double* h = reinterpret_cast<double*>(p);
            const bool okH = copy_finite(h, q->r.H, N);
            const bool okC = copy_finite(h + N, q->r.C, N);
            const bool okx = copy_finite(h + 2 * N, q->r.vnx, nxv);
            const bool oky = copy_finite(h + 2 * N + nxv, q->r.vny, nyv);
            if (!okH || !okC) {
                st = MP_ERR_ARG;
                msg = "Cell averages must be finite";
            } else if (!okx || !oky) {
                st = MP_ERR_ARG;
                msg = "Edge velocity must be finite";
            }
        }
        {
            std::lock_guard<std::mutex> g(mu);
            staged_ahead += (p ? cap : 0) - need;
            q->staged = p;
            q->staged_bytes = p ? cap : 0;
            q->stage_st = st;
            q->stage_msg = msg;
            q->stage = Request::kStaged;
            --staging;
        }
        cv_work.notify_all();
    }
}

// Bring admitted requests onto the device from their staged copies: slot,
// boundary data, state, velocities, then one read of all their advective
// rates.  While epochs are queued, admission stops once the oldest of them
// has completed (after at least kMinAdmit requests) and hands the rest back
// to the front of the queue, so that the next epoch is queued in time.
void mp_server::admit(std::vector<Request*>& batch, int rows_now)
{
    constexpr size_t kMinAdmit = 4;
    std::vector<Request*> ok;
    for (size_t i = 0; i < batch.size(); ++i) {
        if (i >= kMinAdmit && !flights.empty() &&
            hipEventQuery(flight_done[flights.front().set]) == hipSuccess) {
            // back into the queue in arrival (id) order
            std::lock_guard<std::mutex> g(mu);
            std::deque<Request*> merged;
            auto a = batch.begin() + long(i);
            auto b = pending.begin();
            while (a != batch.end() || b != pending.end()) {
                if (b == pending.end() || (a != batch.end() && (*a)->id < (*b)->id)) merged.push_back(*a++);
                else merged.push_back(*b++);
            }
            pending.swap(merged);
            break;
        }
        Request* q = batch[i];
        if (q->stage_st != MP_OK) {
            finish(q, q->stage_st, q->stage_msg);
            continue;
        }
        const mp_mesh_desc& d = q->r.mesh;
        const size_t N = size_t(d.nx) * d.ny, nxv = size_t(d.q) * d.ny * (d.nx + 1);
        const double* h = reinterpret_cast<const double*>(q->staged);
        int k = -1;
        mp_status st = eng.add_mesh(q->r.mesh, &k);
        if (st == MP_OK) {
            q->slot = k;
            bool any_bc = false;
            for (int f = 0; f < 4; ++f) any_bc = any_bc || q->r.bc[f];
            if (any_bc) st = eng.set_boundary(k, q->r.bc);
        }
        if (st == MP_OK) st = eng.upload_async(k, h, h + N, true);
        if (st == MP_OK) st = eng.set_velocity(k, h + 2 * N, h + 2 * N + nxv, false, true);
        if (st == MP_OK && q->r.lithostatic) st = eng.set_lithostatic(k, 1, q->r.surface_y);
        if (st == MP_OK && q->r.limiter)
            st = eng.set_limiter(k, 1, q->r.cell_points_per_axis, q->r.diffusion_samples, q->r.extent_fraction);
        if (st == MP_OK && q->r.thermal_diffusivity != 0.0)
            st = eng.set_diffusion(k, q->r.thermal_diffusivity, q->r.diffusion_samples ? q->r.diffusion_samples : 4,
                                   q->r.extent_fraction != 0.0 ? q->r.extent_fraction : 0.9);
        if (st == MP_OK && q->r.thermal_types)
            st = eng.set_thermal_boundary(k, q->r.thermal_types, q->r.thermal_values);
        if (st != MP_OK) {
            finish(q, st, last_error());
            continue;
        }
        ok.push_back(q);
    }
    if (ok.empty()) return;
    Admission a;
    a.runs = std::move(ok);
    hipError_t e = hipSuccess;
    if (!spare_bits.empty()) {
        a.bits = spare_bits.back();
        spare_bits.pop_back();
    } else {
        e = hipHostMalloc(reinterpret_cast<void**>(&a.bits), sizeof(unsigned long long) * eng.capacity,
                          hipHostMallocDefault);
    }
    if (e == hipSuccess) {
        if (!spare_done.empty()) {
            a.done = spare_done.back();
            spare_done.pop_back();
        } else {
            e = hipEventCreateWithFlags(&a.done, hipEventDisableTiming);
        }
    }
    // the rates follow the admission's uploads and rate kernels on `upload`
    if (e == hipSuccess)
        e = hipMemcpyAsync(a.bits, eng.d_maxbits.ptr, sizeof(unsigned long long) * eng.capacity,
                           hipMemcpyDeviceToHost, eng.stream);
    if (e == hipSuccess) e = hipEventRecord(a.done, eng.stream);
    if (e != hipSuccess) {
        // nothing to wait for: settle the uploads, then fail the runs
        (void)hipStreamSynchronize(eng.stream);
        for (Request* q : a.runs) {
            unstage(q);
            finish(q, MP_ERR_DEVICE, hipGetErrorString(e));
        }
        if (a.bits) spare_bits.push_back(a.bits);
        if (a.done) spare_done.push_back(a.done);
        return;
    }
    for (Request* q : a.runs) q->rows_now = rows_now;
    admitting_runs += int(a.runs.size());
    admitting.push_back(std::move(a));
}

// Make resident the runs of the admissions whose uploads have completed, in
// admission order (wait: block for the oldest one first).
void mp_server::promote(bool wait)
{
    while (!admitting.empty()) {
        Admission& a = admitting.front();
        hipError_t e = wait ? hipEventSynchronize(a.done) : hipEventQuery(a.done);
        if (e == hipErrorNotReady) return;
        wait = false;
        const double t = now();
        for (Request* q : a.runs) {
            unstage(q);    // its copies have run
            if (e != hipSuccess) {
                finish(q, MP_ERR_DEVICE, hipGetErrorString(e));
                continue;
            }
            Slot& s = eng.slots[q->slot];
            std::memcpy(&s.max_rate, &a.bits[q->slot], sizeof s.max_rate);
            if (!(s.max_rate > 0)) {
                finish(q, MP_ERR_STATE, "zero advective rate: the step is unbounded");
                continue;
            }
            s.dt = q->r.cfl / s.max_rate;
            q->res.dt = s.dt;
            q->res.t_admit = t;
            eng.set_rows(q->slot, q->rows_now);
            resident.push_back(q);
            membership_changed = true;
        }
        admitting_runs -= int(a.runs.size());
        spare_bits.push_back(a.bits);
        spare_done.push_back(a.done);
        admitting.pop_front();
    }
}

// Queue a ragged epoch behind the ones in flight: every resident run with
// steps left takes min(remaining, E) steps, so a run never waits for the
// shortest co-resident one.  Each run's state is committed as if the epoch
// succeeds, so the next epoch can be queued before this one ends; a run that
// the epoch fails is retired with its error once nothing queued advances it
// (its later steps are discarded).  Returns false if no run has steps left.
bool mp_server::start_epoch()
{
    Flight f;
    for (Request* q : resident)
        if (q->remaining > 0 && q->fail_st == MP_OK) f.run.push_back(q);
    if (f.run.empty()) return false;
    for (Request* q : f.run) f.E = std::max(f.E, std::min(q->remaining, max_epoch));
    std::vector<int> list;
    for (Request* q : f.run) {
        list.push_back(q->slot);
        f.steps.push_back(std::min(q->remaining, f.E));
    }
    f.set = next_set;
    next_set ^= 1;
    mp_status st = MP_OK;
    // A freed slot may be refilled by a mesh of another size under the same
    // id, so the work list follows membership events, not ids.
    if (membership_changed || list != last_list || f.steps != last_steps) {
        st = eng.set_active(list, f.steps);
        last_list = list;
        last_steps = f.steps;
        if (!st) membership_changed = false;
    }
    if (!st) st = eng.write_tables(scheme);
    if (!st) st = eng.reserve_epochs(f.E);
    if (!st) st = eng.enqueue(f.E, scheme, nullptr);
    if (!st) st = eng.copy_status(f.set, f.E);
    if (!st && hipEventRecord(flight_done[f.set], eng.stream) != hipSuccess)
        st = fail(MP_ERR_DEVICE, "epoch event");
    if (!st) eng.commit(f.E, std::vector<uint32_t>(size_t(eng.capacity), 0u));
    f.st = st;
    for (size_t i = 0; i < f.run.size(); ++i) {
        f.run[i]->remaining -= f.steps[i];
        ++f.run[i]->inflight;
    }
    ++epochs;
    flights.push_back(std::move(f));
    return true;
}

// Wait for the oldest queued epoch and account its runs: steps, totals and
// boundary sums, or the failure that its status reports.
void mp_server::finish_oldest()
{
    Flight f = std::move(flights.front());
    flights.pop_front();
    mp_status st = f.st;
    if (!st && hipEventSynchronize(flight_done[f.set]) != hipSuccess) st = fail(MP_ERR_DEVICE, "epoch wait");
    const uint32_t* errors = eng.copied_errors(f.set);
    const double* totals = eng.copied_totals(f.set);
    double work = 0;
    for (size_t i = 0; i < f.run.size(); ++i) {
        Request* q = f.run[i];
        --q->inflight;
        if (q->fail_st != MP_OK) continue;      // failed in an earlier epoch
        if (st != MP_OK) {
            q->fail_st = st;
            std::snprintf(q->message, sizeof q->message, "%s", last_error());
            continue;
        }
        const uint32_t b = errors[q->slot];
        if (b) {
            q->res.error_bits = b;
            q->fail_st = status_from_bits(b);
            std::snprintf(q->message, sizeof q->message, "%s (device error bits 0x%x) after %d steps",
                          describe_bits(b), b, q->res.steps_done);
            continue;
        }
        const Slot& s = eng.slots[q->slot];
        const int n = f.steps[i];
        for (int k = 0; k < n; ++k) {
            const double* t = totals + (size_t(k) * eng.capacity + q->slot) * 4;
            q->bH += (long double)s.dt * t[2];
            q->bC += (long double)s.dt * t[3];
        }
        const double* t = totals + (size_t(n - 1) * eng.capacity + q->slot) * 4;
        q->res.total_H = t[0];
        q->res.total_C = t[1];
        q->res.steps_done += n;
        work += double(q->r.mesh.nx) * q->r.mesh.ny * n;
    }
    std::lock_guard<std::mutex> g(mu);
    work_done += work;
    const std::pair<double, double> sample(now(), work_done);
    if (progress.size() < kLatencies) progress.push_back(sample);
    else progress[progress_next] = sample;
    progress_next = (progress_next + 1) % kLatencies;
}

// Retire the runs that no queued epoch advances and that are done or failed.
// Outputs of finished runs are copied into pinned blocks on the download
// stream (their epochs have completed; the queued ones do not touch their
// slots); the upload stream waits for those copies before it reuses the
// freed slots, and the finisher thread hands the outputs over.  A run
// admitted with zero steps computes its phase on the engine's stream behind
// the queued epochs, whose first launch clears the error words that the
// phase evaluation reports through, and waits for it there.
void mp_server::retire()
{
    std::vector<Request*> keep, out;
    for (Request* q : resident)
        ((q->inflight == 0 && (q->remaining == 0 || q->fail_st != MP_OK)) ? out : keep).push_back(q);
    if (out.empty()) return;
    resident.swap(keep);
    Retired r;
    for (Request* q : out) {
        if (q->fail_st != MP_OK) {
            finish(q, q->fail_st, nullptr);
            continue;
        }
        const Slot& s = eng.slots[q->slot];
        const size_t N = s.N;
        const bool need_phase = q->r.T || q->r.phi || q->r.Cl || q->r.Cs || q->r.region;
        size_t cap = 0;
        char* blk = pool.get(need_phase ? 6 * N * sizeof(double) + N : 2 * N * sizeof(double), &cap);
        if (!blk) {
            finish(q, MP_ERR_MEMORY, "pinned output allocation failed");
            continue;
        }
        double* h = reinterpret_cast<double*>(blk);
        const hipStream_t main = eng.stream;
        if (!need_phase || s.phase_valid) eng.stream = download;
        mp_status s2 = eng.download_async(q->slot, q->r.H_out ? h : nullptr, q->r.C_out ? h + N : nullptr);
        if (!s2 && need_phase)
            s2 = eng.phase_out(q->slot, q->r.T ? h + 2 * N : nullptr, q->r.phi ? h + 3 * N : nullptr,
                               q->r.Cl ? h + 4 * N : nullptr, q->r.Cs ? h + 5 * N : nullptr,
                               q->r.region ? reinterpret_cast<uint8_t*>(h + 6 * N) : nullptr, true);
        if (!s2 && eng.stream == main && hipStreamSynchronize(main) != hipSuccess)
            s2 = fail(MP_ERR_DEVICE, "phase of an unstepped run");
        eng.stream = main;
        if (s2) {
            (void)hipStreamSynchronize(download);   // copies may already read the block
            pool.put(blk, cap);
            finish(q, s2, last_error());
            continue;
        }
        q->res.boundary_H = double(q->bH);
        q->res.boundary_C = double(q->bC);
        r.runs.push_back(q);
        r.blocks.push_back(blk);
        r.caps.push_back(cap);
    }
    if (r.runs.empty()) return;
    {
        std::lock_guard<std::mutex> g(mu);
        if (!spare_events.empty()) {
            r.copied = spare_events.back();
            spare_events.pop_back();
        }
    }
    hipError_t e = r.copied ? hipSuccess : hipEventCreateWithFlags(&r.copied, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(r.copied, download);
    if (e == hipSuccess) e = hipStreamWaitEvent(upload, r.copied, 0);
    if (e != hipSuccess) {
        // report the runs here, after a full wait for the copies
        const hipError_t e2 = hipStreamSynchronize(download);
        for (size_t i = 0; i < r.runs.size(); ++i) {
            pool.put(r.blocks[i], r.caps[i]);
            finish(r.runs[i], MP_ERR_DEVICE, hipGetErrorString(e2 != hipSuccess ? e2 : e));
        }
        if (r.copied) (void)hipEventDestroy(r.copied);
        return;
    }
    for (Request* q : r.runs) release(q);
    {
        std::lock_guard<std::mutex> g(mu);
        to_finish.push_back(std::move(r));
    }
    cv_finish.notify_one();
}

// Finisher thread: waits for a retired group's copies, moves the outputs into
// the callers' arrays and reports the runs.
void mp_server::finish_loop()
{
    (void)hipSetDevice(device);
    for (;;) {
        Retired r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_finish.wait(lk, [&] { return finisher_stop || !to_finish.empty(); });
            if (to_finish.empty()) return;
            r = std::move(to_finish.front());
            to_finish.pop_front();
        }
        const hipError_t e = hipEventSynchronize(r.copied);
        for (size_t i = 0; i < r.runs.size(); ++i) {
            Request* q = r.runs[i];
            if (e == hipSuccess) {
                const size_t N = size_t(q->r.mesh.nx) * q->r.mesh.ny;
                const double* h = reinterpret_cast<const double*>(r.blocks[i]);
                double* outs[6] = {q->r.H_out, q->r.C_out, q->r.T, q->r.phi, q->r.Cl, q->r.Cs};
                for (int k = 0; k < 6; ++k)
                    if (outs[k]) std::memcpy(outs[k], h + size_t(k) * N, N * sizeof(double));
                if (q->r.region) std::memcpy(q->r.region, h + 6 * N, N);
            }
            pool.put(r.blocks[i], r.caps[i]);
            report(q, e == hipSuccess ? MP_OK : MP_ERR_DEVICE, e == hipSuccess ? nullptr : hipGetErrorString(e));
        }
        std::lock_guard<std::mutex> g(mu);
        spare_events.push_back(r.copied);
    }
}

// Pipelined: with epochs k and k + 1 queued, the worker admits the requests
// that are staged (uploads, face moments and rates on the upload stream),
// queues epochs until two are in flight, then waits for the oldest, retires
// its finished runs, and goes round again.
void mp_server::loop()
{
    (void)hipSetDevice(device);
    for (;;) {
        std::vector<Request*> batch;
        int rows_now = 8;
        {
            std::unique_lock<std::mutex> lk(mu);
            rows_now = rows;
            auto staged_head = [&] { return !pending.empty() && pending.front()->stage == Request::kStaged; };
            cv_work.wait(lk, [&] {
                return stop || staged_head() || !resident.empty() || !flights.empty() || !admitting.empty();
            });
            if (stop && resident.empty() && flights.empty() && admitting.empty()) {
                // no stager may still be writing into a request failed here
                cv_work.wait(lk, [&] { return staging == 0; });
                for (Request* q : pending) {
                    if (q->staged) {
                        pool.put(q->staged, q->staged_bytes);
                        staged_ahead -= q->staged_bytes;
                        q->staged = nullptr;
                    }
                    q->res.status = MP_ERR_STATE;
                    std::snprintf(q->message, sizeof q->message, "server shut down before admission");
                    q->res.t_done = now();
                    q->done = true;
                    ++failed;
                }
                pending.clear();
                to_stage.clear();
                cv_done.notify_all();
                return;
            }
            // Admission: arrival order over the staged requests, first fit,
            // while slots and budget allow.
            int free_slots = eng.capacity - int(resident.size()) - admitting_runs;
            size_t planned = 0;
            for (auto it = pending.begin(); it != pending.end() && free_slots > 0 && !stop;) {
                Request* q = *it;
                if (q->stage != Request::kStaged) break;
                if (eng.arena.in_use() + planned + q->bytes <= eng.arena.budget()) {
                    planned += q->bytes;
                    --free_slots;
                    batch.push_back(q);
                    it = pending.erase(it);
                } else {
                    ++it;
                }
            }
        }
        if (!batch.empty()) {
            // the engine works on the upload stream for the admission, so its
            // synchronisations wait for the uploads only, not for the epochs;
            // the runs join the epochs once promote() has seen their uploads done
            const hipStream_t main = eng.stream;
            eng.stream = upload;
            admit(batch, rows_now);
            eng.stream = main;
        }
        // with nothing else to run, wait for the oldest admission
        bool idle = flights.empty();
        for (Request* q : resident)
            if (q->remaining > 0 && q->fail_st == MP_OK) idle = false;
        promote(idle && !admitting.empty());
        while (flights.size() < 2 && start_epoch()) {
        }
        if (!flights.empty()) finish_oldest();
        retire();
        {
            std::lock_guard<std::mutex> g(mu);
            hbm_in_use = eng.arena.in_use();
            hbm_peak = std::max(hbm_peak, uint64_t(hbm_in_use));
            max_resident = std::max(max_resident, int64_t(resident.size()));
        }
    }
}

// ===================================================================== C ABI

extern "C" {

double mp_server_now(mp_server* s) { return s ? s->now() : 0.0; }

mp_status mp_server_create(int32_t max_slots, const mp_weno_options* weno, const mp_phase_params* phase,
                           int32_t precision, int32_t scheme, uint64_t hbm_budget_bytes,
                           int32_t max_epoch_steps, mp_server** out)
{
    MP_GUARD({
        if (!out) return fail(MP_ERR_ARG, "null argument");
        if (scheme < 1 || scheme > 3) return fail(MP_ERR_ARG, "scheme must be 1 (FE), 2 (SSPRK2) or 3 (SSPRK3)");
        if (max_epoch_steps < 1) return fail(MP_ERR_ARG, "max_epoch_steps must be positive");
        if (hbm_budget_bytes == 0) return fail(MP_ERR_ARG, "an HBM budget is required");
        mp_weno_options w;
        mp_phase_params p;
        mp_default_weno_options(&w);
        mp_default_phase_params(&p);
        if (weno) w = *weno;
        if (phase) p = *phase;
        mp_server* s = new mp_server();
        MP_HIP(hipGetDevice(&s->device));
        mp_status st = s->eng.init(max_slots, precision, w, p, nullptr, size_t(hbm_budget_bytes));
        if (st) { delete s; return st; }
        if (hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&s->download, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->flight_done[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->flight_done[1], hipEventDisableTiming) != hipSuccess) {
            delete s;
            return fail(MP_ERR_DEVICE, "server streams");
        }
        s->scheme = scheme;
        s->max_epoch = max_epoch_steps;
        // the per-step totals of an epoch, reserved before epochs overlap
        st = s->eng.reserve_steps(max_epoch_steps);
        if (!st) st = s->eng.copy_status(0, 0);
        if (!st) st = s->eng.copy_status(1, 0);
        if (!st && hipStreamSynchronize(s->eng.stream) != hipSuccess) st = fail(MP_ERR_DEVICE, "server set-up");
        if (st) { delete s; return st; }
        // One wave per work item: with few resident runs this measured faster
        // than a persistent grid.  Mixed steps run their stages in one launch.
        s->eng.blocks = 0;
        s->eng.step_launch = true;
        s->eng.step_launch_items = Engine::auto_step_items();
        s->worker = std::thread([s] { s->loop(); });
        for (int i = 0; i < mp_server::kStagers; ++i) s->stagers.emplace_back([s] { s->stage_loop(); });
        s->finisher = std::thread([s] { s->finish_loop(); });
        *out = s;
        return MP_OK;
    })
}

mp_status mp_server_submit(mp_server* s, const mp_request* req, int64_t* id)
{
    if (!s || !req || !id) return fail(MP_ERR_ARG, "null argument");
    MP_GUARD({
        if (!req->H || !req->C || !req->vnx || !req->vny) return fail(MP_ERR_ARG, "null input array");
        if (req->nsteps < 0) return fail(MP_ERR_ARG, "nsteps must be nonnegative");
        if (!std::isfinite(req->cfl) || !(req->cfl > 0)) return fail(MP_ERR_ARG, "CFL must be finite and positive");
        MP_CALL(validate_mesh(req->mesh));
        const size_t bytes = Engine::slot_bytes(req->mesh, s->eng.precision);
        if (bytes > s->eng.arena.budget())
            return fail(MP_ERR_MEMORY, "request needs %zu bytes, above the HBM budget %zu", bytes,
                        s->eng.arena.budget());
        Request* q = new Request();
        q->r = *req;
        q->bytes = Arena::round(bytes);
        q->remaining = req->nsteps;
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (s->stop) {
                delete q;
                return fail(MP_ERR_STATE, "server is shutting down");
            }
            q->id = s->next_id++;
            q->res.id = q->id;
            q->res.t_submit = s->now();
            if (s->first_submit < 0) s->first_submit = q->res.t_submit;
            s->all[q->id] = q;
            s->pending.push_back(q);
            s->to_stage.push_back(q);
            ++s->submitted;
            *id = q->id;
        }
        s->cv_stage.notify_one();
        return MP_OK;
    })
}

mp_status mp_server_wait(mp_server* s, int64_t id, double timeout_s, mp_request_result* out)
{
    if (!s) return fail(MP_ERR_ARG, "null handle");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return fail(MP_ERR_ARG, "unknown request id %lld", (long long)id);
    Request* q = it->second;
    auto ready = [&] { return q->done; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return fail(MP_ERR_PENDING, "request %lld not finished", (long long)id);
    }
    if (out) *out = q->res;
    // the result is handed back once: the request is released
    const mp_status st = mp_status(q->res.status);
    char message[sizeof q->message];
    std::memcpy(message, q->message, sizeof message);
    s->all.erase(it);
    delete q;
    if (st != MP_OK) return fail(st, "request %lld: %s", (long long)id, message);
    return MP_OK;
}

mp_status mp_server_drain(mp_server* s)
{
    if (!s) return fail(MP_ERR_ARG, "null handle");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed + s->failed == s->submitted; });
    return MP_OK;
}

mp_status mp_server_stats_get(mp_server* s, mp_server_stats* o)
{
    if (!s || !o) return fail(MP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(o, 0, sizeof *o);
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->failed = s->failed;
    o->queued = int64_t(s->pending.size());
    o->resident = int64_t(s->resident.size());
    o->cell_steps = s->cell_steps;
    o->elapsed_s = s->first_submit >= 0 ? s->last_done - s->first_submit : 0.0;
    std::vector<double> lat(s->latencies);
    if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        auto rank = [&](double p) {
            size_t i = size_t(std::ceil(p * double(lat.size()))) ;
            return lat[std::min(lat.size() - 1, i > 0 ? i - 1 : 0)];
        };
        o->p50_ms = rank(0.50);
        o->p99_ms = rank(0.99);
        double sum = 0;
        for (double v : lat) sum += v;
        o->mean_ms = sum / double(lat.size());
        o->max_ms = lat.back();
    }
    o->hbm_in_use = s->hbm_in_use;
    o->hbm_peak = s->hbm_peak;
    o->hbm_budget = s->eng.arena.budget();
    o->epochs = s->epochs;
    o->max_resident = s->max_resident;
    return MP_OK;
}

mp_status mp_server_progress(mp_server* s, double* t, double* cell_steps, int32_t cap, int32_t* n)
{
    if (!s || !n || cap < 0 || (cap > 0 && (!t || !cell_steps))) return fail(MP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    const size_t count = s->progress.size();
    const size_t first = count < mp_server::kLatencies ? 0 : s->progress_next;
    const size_t m = std::min(count, size_t(cap));
    // the most recent m samples, oldest first
    for (size_t i = 0; i < m; ++i) {
        const auto& p = s->progress[(first + count - m + i) % count];
        t[i] = p.first;
        cell_steps[i] = p.second;
    }
    *n = int32_t(m);
    return MP_OK;
}

mp_status mp_server_set_tile(mp_server* s, int32_t rows_per_item)
{
    if (!s) return fail(MP_ERR_ARG, "null handle");
    if (rows_per_item < 8 || rows_per_item > 1 << 20) return fail(MP_ERR_ARG, "rows per item must be >= 8");
    std::lock_guard<std::mutex> g(s->mu);
    s->rows = rows_per_item;
    return MP_OK;
}

mp_status mp_server_destroy(mp_server* s)
{
    delete s;
    return MP_OK;
}

}  // extern "C"
