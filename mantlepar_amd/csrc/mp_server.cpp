// Serving front end: continuous batching of transport runs on one device.
//
// Callers submit requests from any thread.  One server thread owns the
// engine: while an epoch runs it admits waiting requests (arrival order,
// first fit) while slots and the HBM budget allow, uploading them on a second
// stream; then it waits for the epoch, retires runs that reached their step
// count, and starts the next epoch, which advances every resident run by up
// to E steps (one launch per stage over all of them).  A run's results are
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
    int remaining = 0;
    long double bH = 0, bC = 0;
    mp_request_result res{};
    char message[256] = {0};
    bool done = false;
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
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::map<int64_t, Request*> all;   // submitted and not yet collected by mp_server_wait
    // Latencies (ms) of the most recent completed runs, a ring of kLatencies.
    static constexpr size_t kLatencies = size_t(1) << 16;
    std::vector<double> latencies;
    size_t latency_next = 0;
    std::vector<Request*> resident;
    int64_t next_id = 1, submitted = 0, completed = 0, failed = 0;
    int64_t epochs = 0, max_resident = 0;
    uint64_t hbm_in_use = 0, hbm_peak = 0;
    double cell_steps = 0, first_submit = -1, last_done = 0;
    bool stop = false;
    bool membership_changed = true;   // a slot was filled or freed since the last work list
    std::thread worker;
    // Requests are admitted on `upload` while the previous epoch runs on the
    // engine's stream; the engine's stream waits for `uploaded` before its
    // next work.
    hipStream_t upload = nullptr;
    hipEvent_t uploaded = nullptr;
    hipEvent_t epoch_done = nullptr;   // recorded after the epoch in flight
    // The epoch in flight: its runs, step count and launch status.
    struct Flight {
        bool on = false;
        int E = 0;
        mp_status st = MP_OK;
        std::vector<Request*> run;
    } flight;

    ~mp_server()
    {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        for (auto& kv : all) delete kv.second;
        if (uploaded) (void)hipEventDestroy(uploaded);
        if (epoch_done) (void)hipEventDestroy(epoch_done);
        if (upload) (void)hipStreamDestroy(upload);
    }

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void loop();
    void admit(std::vector<Request*>& batch, int rows_now);
    void start_epoch();
    void finish_epoch();
    void finish(Request* q, mp_status st, const char* msg);
};

void mp_server::finish(Request* q, mp_status st, const char* msg)
{
    if (q->slot >= 0) {
        eng.remove_mesh(q->slot);
        q->slot = -1;
        membership_changed = true;
    }
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
    hbm_in_use = eng.arena.in_use();
    q->r.H = q->r.C = q->r.vnx = q->r.vny = nullptr;
    cv_done.notify_all();
}

// Bring admitted requests onto the device: slot, boundary data, state,
// velocities, then one read of all their advective rates.  An upload takes
// about 0.2 ms of host time, so while an epoch runs, admission stops once that
// epoch has completed (after at least kMinAdmit requests) and hands the rest
// back to the front of the queue: the GPU is never left idle behind a long
// admission.
void mp_server::admit(std::vector<Request*>& batch, int rows_now)
{
    constexpr size_t kMinAdmit = 4;
    std::vector<Request*> ok;
    for (size_t i = 0; i < batch.size(); ++i) {
        if (i >= kMinAdmit && flight.on && hipEventQuery(epoch_done) == hipSuccess) {
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
        int k = -1;
        mp_status st = eng.add_mesh(q->r.mesh, &k);
        if (st == MP_OK) {
            q->slot = k;
            bool any_bc = false;
            for (int f = 0; f < 4; ++f) any_bc = any_bc || q->r.bc[f];
            if (any_bc) st = eng.set_boundary(k, q->r.bc);
        }
        if (st == MP_OK) st = eng.upload_async(k, q->r.H, q->r.C);
        if (st == MP_OK) st = eng.set_velocity(k, q->r.vnx, q->r.vny, false);
        if (st == MP_OK && q->r.lithostatic) st = eng.set_lithostatic(k, 1, q->r.surface_y);
        if (st != MP_OK) {
            finish(q, st, last_error());
            continue;
        }
        ok.push_back(q);
    }
    if (ok.empty()) return;
    std::vector<unsigned long long> bits(size_t(eng.capacity), 0ull);
    hipError_t e = hipMemcpyAsync(bits.data(), eng.d_maxbits.ptr, sizeof(unsigned long long) * eng.capacity,
                                  hipMemcpyDeviceToHost, eng.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(eng.stream);
    const double t = now();
    for (Request* q : ok) {
        if (e != hipSuccess) {
            finish(q, MP_ERR_DEVICE, hipGetErrorString(e));
            continue;
        }
        Slot& s = eng.slots[q->slot];
        std::memcpy(&s.max_rate, &bits[q->slot], sizeof s.max_rate);
        if (!(s.max_rate > 0)) {
            finish(q, MP_ERR_STATE, "zero advective rate: the step is unbounded");
            continue;
        }
        s.dt = q->r.cfl / s.max_rate;
        q->res.dt = s.dt;
        q->res.t_admit = t;
        eng.set_rows(q->slot, rows_now);
        resident.push_back(q);
        membership_changed = true;
    }
}

// Ragged epoch: every resident run takes min(remaining, E) steps, so a run
// never waits for the shortest co-resident one.  The launches are enqueued
// here and left running; finish_epoch collects them.
void mp_server::start_epoch()
{
    Flight f;
    f.on = true;
    for (Request* q : resident) {
        if (q->remaining > 0) f.run.push_back(q);
    }
    for (Request* q : f.run) f.E = std::max(f.E, std::min(q->remaining, max_epoch));
    if (!f.run.empty()) {
        std::vector<int> list, steps;
        for (Request* q : f.run) {
            list.push_back(q->slot);
            steps.push_back(std::min(q->remaining, f.E));
        }
        mp_status st = MP_OK;
        // A freed slot may be refilled by a mesh of another size under the
        // same id, so the work list follows membership events, not ids.
        if (membership_changed || list != last_list || steps != last_steps) {
            st = eng.set_active(list, steps);
            last_list = list;
            last_steps = steps;
        }
        if (!st) membership_changed = false;
        if (!st) st = eng.reserve_steps(f.E);
        if (!st) st = eng.write_tables(scheme);
        if (!st) st = eng.enqueue(f.E, scheme, nullptr);
        if (!st && hipEventRecord(epoch_done, eng.stream) != hipSuccess)
            st = fail(MP_ERR_DEVICE, "epoch event");
        f.st = st;
        ++epochs;
    }
    flight = std::move(f);
}

// Wait for the epoch in flight, commit its runs, and retire those that are
// done (runs admitted with zero steps retire here too, once their upload has
// landed: the engine's stream waits for the upload stream).
void mp_server::finish_epoch()
{
    if (!flight.on) return;
    const std::vector<Request*> run = flight.run;
    const int E = flight.E;
    mp_status st = flight.st;
    flight = Flight();
    std::vector<uint32_t> errors(size_t(eng.capacity), 0u);
    std::vector<double> totals;
    if (!run.empty()) {
        if (!st) st = eng.fetch_status(E, errors, &totals);
        if (!st) eng.commit(E, errors);
    }
    std::vector<Request*> keep, retire;
    for (Request* q : resident) {
        const bool stepped = std::find(run.begin(), run.end(), q) != run.end();
        if (stepped) {
            if (st != MP_OK) {
                finish(q, st, last_error());
                continue;
            }
            const uint32_t b = errors[q->slot];
            if (b) {
                char msg[200];
                std::snprintf(msg, sizeof msg, "%s (device error bits 0x%x) after %d steps",
                              describe_bits(b), b, q->res.steps_done);
                q->res.error_bits = b;
                finish(q, status_from_bits(b), msg);
                continue;
            }
            const Slot& s = eng.slots[q->slot];
            const int n = eng.steps_of(q->slot);
            for (int i = 0; i < n; ++i) {
                const double* t = totals.data() + (size_t(i) * eng.capacity + q->slot) * 4;
                q->bH += (long double)s.dt * t[2];
                q->bC += (long double)s.dt * t[3];
            }
            const double* t = totals.data() + (size_t(n - 1) * eng.capacity + q->slot) * 4;
            q->res.total_H = t[0];
            q->res.total_C = t[1];
            q->remaining -= n;
            q->res.steps_done += n;
        }
        (q->remaining == 0 ? retire : keep).push_back(q);
    }
    resident.swap(keep);
    // Outputs of finished runs: queue every copy, then one synchronisation.
    std::vector<Request*> good;
    for (Request* q : retire) {
        mp_status s2 = eng.download_async(q->slot, q->r.H_out, q->r.C_out);
        if (!s2 && (q->r.T || q->r.phi || q->r.Cl || q->r.Cs || q->r.region))
            s2 = eng.phase_out(q->slot, q->r.T, q->r.phi, q->r.Cl, q->r.Cs, q->r.region, true);
        if (s2) finish(q, s2, last_error());
        else good.push_back(q);
    }
    if (!good.empty()) {
        const hipError_t e = hipStreamSynchronize(eng.stream);
        for (Request* q : good) {
            q->res.boundary_H = double(q->bH);
            q->res.boundary_C = double(q->bC);
            finish(q, e == hipSuccess ? MP_OK : MP_ERR_DEVICE, e == hipSuccess ? nullptr : hipGetErrorString(e));
        }
    }
}

// Pipelined: while epoch k runs on the engine's stream, the worker admits the
// requests that arrived (their uploads, face moments and rates on the upload
// stream), then waits for epoch k, retires its finished runs and starts epoch
// k + 1 with the new runs in it.
void mp_server::loop()
{
    (void)hipSetDevice(device);
    for (;;) {
        std::vector<Request*> batch;
        int rows_now = 8;
        {
            std::unique_lock<std::mutex> lk(mu);
            rows_now = rows;
            cv_work.wait(lk, [&] { return stop || !pending.empty() || !resident.empty(); });
            if (stop && resident.empty() && !flight.on) {
                for (Request* q : pending) {
                    q->res.status = MP_ERR_STATE;
                    std::snprintf(q->message, sizeof q->message, "server shut down before admission");
                    q->res.t_done = now();
                    q->done = true;
                    ++failed;
                }
                pending.clear();
                cv_done.notify_all();
                return;
            }
            // Admission: arrival order, first fit, while slots and budget allow.
            int free_slots = eng.capacity - int(resident.size());
            size_t planned = 0;
            for (auto it = pending.begin(); it != pending.end() && free_slots > 0 && !stop;) {
                Request* q = *it;
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
            // synchronisations wait for the uploads only, not for the epoch
            const hipStream_t main = eng.stream;
            eng.stream = upload;
            admit(batch, rows_now);
            eng.stream = main;
            if (hipEventRecord(uploaded, upload) == hipSuccess)
                (void)hipStreamWaitEvent(main, uploaded, 0);
        }
        finish_epoch();
        {
            std::lock_guard<std::mutex> g(mu);
            hbm_in_use = eng.arena.in_use();
            hbm_peak = std::max(hbm_peak, uint64_t(hbm_in_use));
            max_resident = std::max(max_resident, int64_t(resident.size()));
        }
        if (!resident.empty()) start_epoch();
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
            hipEventCreateWithFlags(&s->uploaded, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->epoch_done, hipEventDisableTiming) != hipSuccess) {
            delete s;
            return fail(MP_ERR_DEVICE, "server upload stream");
        }
        s->scheme = scheme;
        s->max_epoch = max_epoch_steps;
        // One wave per work item: with few resident runs this measured faster
        // than a persistent grid.
        s->eng.blocks = 0;
        s->worker = std::thread([s] { s->loop(); });
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
            ++s->submitted;
            *id = q->id;
        }
        s->cv_work.notify_one();
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
