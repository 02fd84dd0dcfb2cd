// Batched front end: many meshes, one persistent launch per stage, optional
// hipGraph replay, and the schedule autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mp_engine.h"
#include "mp_launch.h"

using namespace mp;

struct mp_batch {
    Engine eng;
    int rows = 64;
    bool dirty = true;                 // membership or schedule changed
    hipGraphExec_t graph = nullptr;
    uint64_t graph_gen = 0;
    int graph_nsteps = 0, graph_scheme = 0, graph_blocks = -1;
    bool graph_step = false;
    // MP_LAUNCH_PER_STAGE, _PER_STEP or _AUTO (the autotuner's choice, per-step
    // launches until it has measured)
    int launch_mode = MP_LAUNCH_AUTO;
    bool tuned = false, tuned_step = true;

    ~mp_batch() { drop_graph(); }
    void drop_graph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    mp_status check_id(int id) const
    {
        if (id < 0 || id >= eng.capacity || !eng.slots[id].live)
            return fail(MP_ERR_ARG, "no resident mesh with id %d", id);
        return MP_OK;
    }
    std::vector<int> live() const
    {
        std::vector<int> v;
        for (int k = 0; k < eng.capacity; ++k)
            if (eng.slots[k].live) v.push_back(k);
        return v;
    }
    mp_status prepare(int nsteps, int scheme);
    mp_status run(int nsteps, int scheme, bool use_graph, std::vector<hipEvent_t>* events);
    mp_status step(int nsteps, int scheme, int flags, mp_step_report* reports);
    mp_status profile(int nsteps, int scheme, double* ms_total, double* ms_stage, double* ms_phase);
    mp_status time_steps(int nsteps, int scheme, double* ms);
    bool want_step_launch() const
    {
        return launch_mode == MP_LAUNCH_PER_STEP || (launch_mode == MP_LAUNCH_AUTO && tuned_step);
    }
    mp_status autotune(int nsteps, int scheme, const char* cache, int* rows_out, int* blocks_out,
                       double* ms_out);
    std::string shape_key(int scheme) const;
};

static mp_status check_scheme(int nsteps, int scheme)
{
    if (scheme < 1 || scheme > 3) return fail(MP_ERR_ARG, "scheme must be 1 (FE), 2 (SSPRK2) or 3 (SSPRK3)");
    if (nsteps < 0) return fail(MP_ERR_ARG, "nsteps must be nonnegative");
    return MP_OK;
}

// Work list, device tables and totals for a call of nsteps.
mp_status mp_batch::prepare(int nsteps, int scheme)
{
    std::vector<int> list = live();
    if (list.empty()) return fail(MP_ERR_STATE, "the batch holds no mesh");
    for (int k : list) {
        const Slot& s = eng.slots[k];
        if (!s.have_velocity) return fail(MP_ERR_STATE, "mesh %d: set face velocities before stepping", k);
        if (!std::isfinite(s.dt) || !(s.dt > 0)) return fail(MP_ERR_STATE, "mesh %d: no time step set", k);
    }
    if (dirty) {
        for (int k : list) eng.set_rows(k, rows);
        MP_CALL(eng.set_active(list));
        dirty = false;
    }
    eng.step_launch = want_step_launch();
    // untuned AUTO follows the item-count rule; a measured or forced choice holds
    eng.step_launch_items = (launch_mode == MP_LAUNCH_AUTO && !tuned) ? Engine::auto_step_items() : 0;
    MP_CALL(eng.reserve_steps(nsteps));
    return eng.write_tables(scheme);
}

mp_status mp_batch::run(int nsteps, int scheme, bool use_graph, std::vector<hipEvent_t>* events)
{
    const bool one = eng.uses_step_launch(events != nullptr);
    if (one) MP_CALL(eng.reserve_epochs(nsteps));
    if (!use_graph) return eng.enqueue(nsteps, scheme, events);
    if (!graph || graph_gen != eng.generation || graph_nsteps != nsteps || graph_scheme != scheme ||
        graph_blocks != eng.blocks || graph_step != one) {
        drop_graph();
        hipGraph_t g = nullptr;
        MP_HIP(hipStreamBeginCapture(eng.stream, hipStreamCaptureModeThreadLocal));
        mp_status s = eng.enqueue(nsteps, scheme, nullptr);
        hipError_t e = hipStreamEndCapture(eng.stream, &g);
        if (s) {
            if (g) (void)hipGraphDestroy(g);
            return s;
        }
        if (e != hipSuccess) return fail(MP_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            graph = nullptr;
            return fail(MP_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e));
        }
        graph_gen = eng.generation;
        graph_nsteps = nsteps;
        graph_scheme = scheme;
        graph_blocks = eng.blocks;
        graph_step = one;
    }
    MP_HIP(hipGraphLaunch(graph, eng.stream));
    return MP_OK;
}

mp_status mp_batch::step(int nsteps, int scheme, int flags, mp_step_report* reports)
{
    MP_CALL(check_scheme(nsteps, scheme));
    if (nsteps == 0) {
        if (reports)
            for (int k : live()) {
                std::memset(&reports[k], 0, sizeof reports[k]);
                reports[k].dt = eng.slots[k].dt;
                reports[k].max_rate = eng.slots[k].max_rate;
            }
        return MP_OK;
    }
    MP_CALL(prepare(nsteps, scheme));
    MP_CALL(run(nsteps, scheme, (flags & MP_BATCH_GRAPH) != 0, nullptr));
    std::vector<uint32_t> errors;
    std::vector<double> totals;
    MP_CALL(eng.fetch_status(nsteps, errors, reports ? &totals : nullptr));
    eng.commit(nsteps, errors);
    int first_bad = -1;
    for (int k : eng.active) {
        if (errors[k] && first_bad < 0) first_bad = k;
        if (reports) eng.make_report(k, nsteps, totals, errors[k], &reports[k]);
    }
    if (first_bad >= 0)
        return fail(status_from_bits(errors[first_bad]), "mesh %d: %s (device error bits 0x%x); its state is unchanged",
                    first_bad, describe_bits(errors[first_bad]), errors[first_bad]);
    return MP_OK;
}

mp_status mp_batch::profile(int nsteps, int scheme, double* ms_total, double* ms_stage, double* ms_phase)
{
    MP_CALL(check_scheme(nsteps, scheme));
    if (nsteps < 1) return fail(MP_ERR_ARG, "profile needs at least one step");
    MP_CALL(prepare(nsteps, scheme));
    std::vector<hipEvent_t> ev(size_t(6) * nsteps, nullptr);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    mp_status st = MP_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) st = fail(MP_ERR_DEVICE, "event create failed");
    if (!st && (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess))
        st = fail(MP_ERR_DEVICE, "event create failed");
    if (!st && hipEventRecord(t0, eng.stream) != hipSuccess) st = fail(MP_ERR_DEVICE, "event record failed");
    if (!st) st = eng.enqueue(nsteps, scheme, &ev);
    if (!st && hipEventRecord(t1, eng.stream) != hipSuccess) st = fail(MP_ERR_DEVICE, "event record failed");
    if (!st && hipStreamSynchronize(eng.stream) != hipSuccess) st = fail(MP_ERR_DEVICE, "synchronize failed");
    double plain = 0, fused = 0;
    int np = 0, nf = 0;
    float total = 0;
    for (int s = 0; s < nsteps && !st; ++s)
        for (int k = 0; k < scheme; ++k) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev[2 * (s * 3 + k)], ev[2 * (s * 3 + k) + 1]) != hipSuccess)
                st = fail(MP_ERR_DEVICE, "event timing failed");
            if (k == scheme - 1) { fused += ms; ++nf; } else { plain += ms; ++np; }
        }
    if (!st && hipEventElapsedTime(&total, t0, t1) != hipSuccess) st = fail(MP_ERR_DEVICE, "event timing failed");
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    MP_CALL(st);
    // No commit: the stepped states land in the slots' spare buffers.
    if (ms_total) *ms_total = total;
    if (ms_stage) *ms_stage = np ? plain / np : fused / nf;
    if (ms_phase) *ms_phase = fused / nf;
    return MP_OK;
}

// Eager wall time of nsteps steps on the device (events around the enqueue);
// no commit, like profile.
mp_status mp_batch::time_steps(int nsteps, int scheme, double* ms)
{
    MP_CALL(prepare(nsteps, scheme));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    mp_status st = MP_OK;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
        st = fail(MP_ERR_DEVICE, "event create failed");
    if (!st && hipEventRecord(t0, eng.stream) != hipSuccess) st = fail(MP_ERR_DEVICE, "event record failed");
    if (!st) st = run(nsteps, scheme, false, nullptr);
    if (!st && hipEventRecord(t1, eng.stream) != hipSuccess) st = fail(MP_ERR_DEVICE, "event record failed");
    if (!st && hipStreamSynchronize(eng.stream) != hipSuccess) st = fail(MP_ERR_DEVICE, "synchronize failed");
    float t = 0;
    if (!st && hipEventElapsedTime(&t, t0, t1) != hipSuccess) st = fail(MP_ERR_DEVICE, "event timing failed");
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    MP_CALL(st);
    *ms = t;
    return MP_OK;
}

std::string mp_batch::shape_key(int scheme) const
{
    int dev = 0;
    hipDeviceProp_t prop;
    std::memset(&prop, 0, sizeof prop);
    if (hipGetDevice(&dev) == hipSuccess) (void)hipGetDeviceProperties(&prop, dev);
    long long cells = 0;
    int meshes = 0, qmax = 0;
    for (int k : live()) {
        cells += (long long)eng.slots[k].N;
        ++meshes;
        qmax = std::max(qmax, eng.slots[k].desc.q);
    }
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s/cu%d/prec%d/s%d/q%d/m%d/c%lld", prop.gcnArchName,
                  prop.multiProcessorCount, eng.precision, scheme, qmax, meshes, cells);
    return std::string(buf);
}

// Cache file: one JSON object whose keys are shape keys, values
// {"rows": r, "blocks": b, "step": s, "ms_per_step": t} (s: one launch per
// step, 0 or 1; absent: 0).  Parsed by key lookup only.
static bool cache_lookup(const char* path, const std::string& key, int& rows, int& blocks, int& step,
                         double& ms)
{
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const std::string tag = "\"" + key + "\"";
    const size_t at = text.find(tag);
    if (at == std::string::npos) return false;
    const size_t open = text.find('{', at), close = text.find('}', at);
    if (open == std::string::npos || close == std::string::npos || close < open) return false;
    const std::string body = text.substr(open, close - open + 1);
    const size_t r = body.find("\"rows\":"), b = body.find("\"blocks\":"), m = body.find("\"ms_per_step\":");
    if (r == std::string::npos || b == std::string::npos) return false;
    rows = std::atoi(body.c_str() + r + 7);
    blocks = std::atoi(body.c_str() + b + 9);
    const size_t sp = body.find("\"step\":");
    step = sp == std::string::npos ? 0 : std::atoi(body.c_str() + sp + 7);
    ms = m == std::string::npos ? 0.0 : std::atof(body.c_str() + m + 14);
    return rows >= 8 && blocks >= 0;
}

static void cache_store(const char* path, const std::string& key, int rows, int blocks, int step, double ms)
{
    std::string text;
    {
        std::ifstream f(path);
        if (f) {
            std::stringstream ss;
            ss << f.rdbuf();
            text = ss.str();
        }
    }
    char entry[512];
    std::snprintf(entry, sizeof entry,
                  "  \"%s\": {\"rows\": %d, \"blocks\": %d, \"step\": %d, \"ms_per_step\": %.6f}",
                  key.c_str(), rows, blocks, step, ms);
    const size_t end = text.rfind('}');
    if (text.empty() || end == std::string::npos) {
        text = std::string("{\n") + entry + "\n}\n";
    } else {
        const bool has_entries = text.find(':') != std::string::npos && text.find(':') < end;
        text = text.substr(0, end);
        while (!text.empty() && (text.back() == '\n' || text.back() == ' ')) text.pop_back();
        text += has_entries ? ",\n" : "\n";
        text += entry;
        text += "\n}\n";
    }
    std::ofstream o(path, std::ios::trunc);
    o << text;
}

mp_status mp_batch::autotune(int nsteps, int scheme, const char* cache, int* rows_out, int* blocks_out,
                             double* ms_out)
{
    MP_CALL(check_scheme(nsteps, scheme));
    if (nsteps < 1) return fail(MP_ERR_ARG, "autotune needs at least one step");
    if (live().empty()) return fail(MP_ERR_STATE, "the batch holds no mesh");
    const std::string key = shape_key(scheme);
    int br = rows, bb = eng.blocks, bs = 0;
    double best = -1;
    if (cache && cache_lookup(cache, key, br, bb, bs, best)) {
        rows = br;
        eng.blocks = bb;
        tuned_step = bs != 0;
        tuned = true;
        dirty = true;
    } else {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        // rows per item x grid: one wave per item (0) or a persistent grid of
        // 1..3 resident blocks per CU taking items from a queue
        const int row_set[] = {8, 12, 16, 24, 32, 64};
        const int grid_mult[] = {0, 1, 2, 3};
        tuned = true;                       // measured choices from here on
        tuned_step = false;                 // the per-stage launches are timed one by one
        MP_CALL(prepare(nsteps, scheme));   // warm the code objects and tables
        MP_CALL(eng.enqueue(1, scheme, nullptr));
        MP_HIP(hipStreamSynchronize(eng.stream));
        for (int r : row_set)
            for (int g : grid_mult) {
                rows = r;
                eng.blocks = g * cus;
                dirty = true;
                double t = 0, best_rep = -1;
                for (int rep = 0; rep < 2; ++rep) {
                    MP_CALL(profile(nsteps, scheme, &t, nullptr, nullptr));
                    if (best_rep < 0 || t < best_rep) best_rep = t;
                }
                if (best < 0 || best_rep < best) {
                    best = best_rep;
                    br = r;
                    bb = eng.blocks;
                }
            }
        best /= nsteps;
        rows = br;
        eng.blocks = bb;
        dirty = true;
        // mixed, one wave per item: per-stage against per-step launches, by
        // the wall time of whole steps
        if (eng.precision == MP_PREC_MIXED && bb == 0) {
            double t[2] = {-1, -1};
            for (int rep = 0; rep < 3; ++rep)
                for (int one = 0; one < 2; ++one) {
                    tuned_step = one != 0;
                    const int save = launch_mode;
                    launch_mode = MP_LAUNCH_AUTO;
                    double ms = 0;
                    mp_status st = time_steps(nsteps, scheme, &ms);
                    launch_mode = save;
                    MP_CALL(st);
                    if (t[one] < 0 || ms < t[one]) t[one] = ms;
                }
            bs = t[1] < t[0] ? 1 : 0;
            best = t[bs] / nsteps;
        }
        tuned_step = bs != 0;
        if (cache) cache_store(cache, key, br, bb, bs, best);
    }
    drop_graph();
    if (rows_out) *rows_out = br;
    if (blocks_out) *blocks_out = bb;
    if (ms_out) *ms_out = best;
    return MP_OK;
}

// ===================================================================== C ABI

extern "C" {

mp_status mp_batch_create(int32_t capacity, const mp_weno_options* weno, const mp_phase_params* phase,
                          int32_t precision, uint64_t hbm_budget_bytes, void* stream, mp_batch** out)
{
    MP_GUARD({
        if (!out) return fail(MP_ERR_ARG, "null argument");
        mp_weno_options w;
        mp_phase_params p;
        mp_default_weno_options(&w);
        mp_default_phase_params(&p);
        if (weno) w = *weno;
        if (phase) p = *phase;
        mp_batch* b = new mp_batch();
        const size_t budget = hbm_budget_bytes ? size_t(hbm_budget_bytes) : ~size_t(0) / 2;
        mp_status s = b->eng.init(capacity, precision, w, p, stream, budget);
        if (s) { delete b; return s; }
        // Defaults before any tuning: 12-row items, one wave per item and per
        // workgroup (the autotuner's choice for the 256-mesh batch on MI355X,
        // in both precisions).
        b->rows = 12;
        b->eng.blocks = 0;
        *out = b;
        return MP_OK;
    })
}

mp_status mp_batch_destroy(mp_batch* b)
{
    delete b;
    return MP_OK;
}

mp_status mp_batch_add_mesh(mp_batch* b, const mp_mesh_desc* mesh, int32_t* id)
{
    if (!b || !mesh || !id) return fail(MP_ERR_ARG, "null argument");
    MP_GUARD({
        int k = -1;
        MP_CALL(b->eng.add_mesh(*mesh, &k));
        *id = k;
        b->dirty = true;
        return MP_OK;
    })
}

mp_status mp_batch_remove_mesh(mp_batch* b, int32_t id)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({
        MP_HIP(hipStreamSynchronize(b->eng.stream));
        b->eng.remove_mesh(id);
        b->dirty = true;
        return MP_OK;
    })
}

mp_status mp_batch_upload_state(mp_batch* b, int32_t id, const double* H, const double* C)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.upload(id, H, C); })
}

mp_status mp_batch_download_state(mp_batch* b, int32_t id, double* H, double* C)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.download(id, H, C); })
}

mp_status mp_batch_set_face_velocity(mp_batch* b, int32_t id, const double* vnx, const double* vny)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.set_velocity(id, vnx, vny, true); })
}

mp_status mp_batch_set_boundary(mp_batch* b, int32_t id, const uint8_t* bottom, const uint8_t* right,
                                const uint8_t* top, const uint8_t* left)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    const uint8_t* sides[4] = {bottom, right, top, left};
    MP_GUARD({ return b->eng.set_boundary(id, sides); })
}

mp_status mp_batch_set_inflow(mp_batch* b, int32_t id, const double* inflow_H, const double* inflow_C)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.set_inflow(id, inflow_H, inflow_C); })
}

mp_status mp_batch_stable_dt(mp_batch* b, int32_t id, double cfl, double* dt)
{
    if (!b || !dt) return fail(MP_ERR_ARG, "null argument");
    MP_CALL(b->check_id(id));
    const Slot& s = b->eng.slots[id];
    if (!s.have_velocity) return fail(MP_ERR_STATE, "set face velocities first");
    if (!std::isfinite(cfl) || !(cfl > 0)) return fail(MP_ERR_ARG, "CFL must be finite and positive");
    if (!(s.max_rate > 0)) return fail(MP_ERR_STATE, "zero advective rate: the step is unbounded");
    *dt = cfl / s.max_rate;
    return MP_OK;
}

mp_status mp_batch_set_dt(mp_batch* b, int32_t id, double dt)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    if (!std::isfinite(dt) || !(dt > 0)) return fail(MP_ERR_ARG, "Time step does not advance finite time");
    b->eng.slots[id].dt = dt;
    return MP_OK;
}

mp_status mp_batch_step(mp_batch* b, int32_t nsteps, int32_t scheme, int32_t flags, mp_step_report* reports)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return b->step(nsteps, scheme, flags, reports); })
}

mp_status mp_batch_set_lithostatic(mp_batch* b, int32_t id, int32_t enable, double surface_y)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.set_lithostatic(id, enable, surface_y); })
}

mp_status mp_batch_set_limiter(mp_batch* b, int32_t id, int32_t enable, int32_t cell_points_per_axis,
                              int32_t diffusion_samples, double extent_fraction)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({
        return b->eng.set_limiter(id, enable, cell_points_per_axis, diffusion_samples, extent_fraction);
    })
}

mp_status mp_batch_set_diffusion(mp_batch* b, int32_t id, double kappa, int32_t diffusion_samples,
                                double extent_fraction)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.set_diffusion(id, kappa, diffusion_samples, extent_fraction); })
}

mp_status mp_batch_set_thermal_boundary(mp_batch* b, int32_t id, const uint8_t* types, const double* values)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.set_thermal_boundary(id, types, values); })
}

mp_status mp_batch_phase(mp_batch* b, int32_t id, double* T, double* phi, double* Cl, double* Cs,
                         uint8_t* region)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_CALL(b->check_id(id));
    MP_GUARD({ return b->eng.phase_out(id, T, phi, Cl, Cs, region, false); })
}

mp_status mp_batch_set_schedule(mp_batch* b, int32_t rows_per_item, int32_t blocks)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    if (rows_per_item < 8 || rows_per_item > 1 << 20) return fail(MP_ERR_ARG, "rows per item must be >= 8");
    if (blocks < 0 || blocks > 1 << 20) return fail(MP_ERR_ARG, "blocks must be in [0, 2^20]");
    b->rows = rows_per_item;
    b->eng.blocks = blocks;
    b->dirty = true;
    return MP_OK;
}

mp_status mp_batch_get_schedule(mp_batch* b, int32_t* rows_per_item, int32_t* blocks)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    if (rows_per_item) *rows_per_item = b->rows;
    if (blocks) *blocks = b->eng.blocks;
    return MP_OK;
}

mp_status mp_batch_set_balance_tolerance(mp_batch* b, double tol)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    if (!std::isfinite(tol) || tol < 0) return fail(MP_ERR_ARG, "balance tolerance must be finite and >= 0");
    b->eng.balance_tol = tol;
    ++b->eng.generation;      // a captured graph holds the old value
    return MP_OK;
}

mp_status mp_batch_set_launch(mp_batch* b, int32_t mode)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    if (mode != MP_LAUNCH_PER_STAGE && mode != MP_LAUNCH_PER_STEP && mode != MP_LAUNCH_AUTO)
        return fail(MP_ERR_ARG, "launch mode must be MP_LAUNCH_PER_STAGE, _PER_STEP or _AUTO");
    b->launch_mode = mode;
    return MP_OK;
}

mp_status mp_batch_get_launch(mp_batch* b, int32_t* per_step)
{
    if (!b || !per_step) return fail(MP_ERR_ARG, "null argument");
    Engine& e = b->eng;
    long long items = 0;
    bool limited = false;
    for (int k : b->live()) {
        items += e.slots[k].nitems;
        limited = limited || e.slots[k].md.limit != 0;
    }
    const bool rule = b->launch_mode == MP_LAUNCH_AUTO && !b->tuned;
    *per_step = (b->want_step_launch() && e.precision == MP_PREC_MIXED && e.blocks == 0 && !limited &&
                 (!rule || items < Engine::auto_step_items())) ? 1 : 0;
    return MP_OK;
}

mp_status mp_batch_autotune(mp_batch* b, int32_t nsteps, int32_t scheme, const char* cache_path,
                            int32_t* rows_per_item, int32_t* blocks, double* ms_per_step)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({
        int r = 0, g = 0;
        MP_CALL(b->autotune(nsteps, scheme, cache_path, &r, &g, ms_per_step));
        if (rows_per_item) *rows_per_item = r;
        if (blocks) *blocks = g;
        return MP_OK;
    })
}

mp_status mp_batch_profile(mp_batch* b, int32_t nsteps, int32_t scheme, double* ms_total, double* ms_stage,
                           double* ms_phase_stage)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    MP_GUARD({ return b->profile(nsteps, scheme, ms_total, ms_stage, ms_phase_stage); })
}

mp_status mp_batch_info(mp_batch* b, int32_t* meshes, int64_t* cells, int64_t* items, uint64_t* hbm_in_use)
{
    if (!b) return fail(MP_ERR_ARG, "null handle");
    int32_t m = 0;
    int64_t c = 0, it = 0;
    for (int k : b->live()) {
        Slot tmp = b->eng.slots[k];
        ++m;
        c += int64_t(tmp.N);
        const int strips = (tmp.desc.nx + kStripCells - 1) / kStripCells;
        int r = std::min(std::max(b->rows, 8), std::max(tmp.desc.ny, 8));
        it += int64_t(strips) * ((tmp.desc.ny + r - 1) / r);
    }
    if (meshes) *meshes = m;
    if (cells) *cells = c;
    if (items) *items = it;
    if (hbm_in_use) *hbm_in_use = b->eng.arena.in_use();
    return MP_OK;
}

}  // extern "C"
