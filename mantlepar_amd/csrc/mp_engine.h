// Multi-mesh engine shared by the batched (mp_batch_*) and serving
// (mp_server_*) front ends.
//
// Every resident mesh occupies one slot: one contiguous block of an HBM
// arena holding its states, face data, phase outputs and partials, plus a
// MeshDev record in a device array.  A step over any set of live slots is one
// persistent launch per stage (batch_sweep_kernel) over a work list of
// (slot, item) pairs, followed by one per-slot reduction; the stage pointer
// records live in device memory at fixed addresses, so a captured hipGraph
// stays valid while the host rewrites them between calls.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "mp_host.h"

namespace mp {

// First-fit sub-allocator over large device chunks; never frees chunks, so
// releasing a block never synchronises the device.
class Arena {
public:
    Arena() = default;
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    ~Arena();
    void configure(size_t budget_bytes, size_t chunk_bytes);
    mp_status alloc(size_t bytes, void** out);
    void release(void* p);
    size_t in_use() const { return in_use_; }
    size_t budget() const { return budget_; }
    bool fits(size_t bytes) const { return in_use_ + round(bytes) <= budget_; }
    static size_t round(size_t b) { return (b + 255) & ~size_t(255); }

private:
    struct Chunk {
        char* base = nullptr;
        size_t size = 0;
        std::map<size_t, size_t> free;   // offset -> size
    };
    std::vector<Chunk> chunks_;
    std::map<void*, std::pair<size_t, size_t>> live_;   // ptr -> (chunk, size)
    size_t budget_ = 0, chunk_ = size_t(1) << 30, in_use_ = 0;
};

// Pinned host buffers for uploads, used in turn.  The host writes a buffer
// only after the device has run the copies that read it the last time, so an
// upload never waits for the work queued before it.
class PinnedRing {
public:
    PinnedRing() = default;
    PinnedRing(const PinnedRing&) = delete;
    PinnedRing& operator=(const PinnedRing&) = delete;
    ~PinnedRing();
    // The next buffer, at least `bytes` long, free for the host to write.
    mp_status next(size_t bytes, char** out);
    // The buffer that next() returned is read by the copies enqueued on
    // `stream` so far.
    mp_status issued(hipStream_t stream);

private:
    static constexpr int kBufs = 3;
    struct Buf {
        char* p = nullptr;
        size_t cap = 0;
        hipEvent_t read = nullptr;
        bool pending = false;
    };
    Buf bufs_[kBufs];
    int cur_ = 0;
};

struct Slot {
    bool live = false;
    mp_mesh_desc desc{};
    MeshDev md{};
    size_t N = 0;
    void* block = nullptr;
    double* H[5] = {};
    double* C[5] = {};
    double* phase[2] = {};     // T, phi, Cl, Cs (N each) then region bytes
    double* partials = nullptr;   // two step-parity sets of three stages
    uint8_t* bc = nullptr;     // bottom(nx) right(ny) top(nx) left(ny) policies
    double* inflow = nullptr;  // H then C, each [face][q] in side order
    TepPair* tep = nullptr;    // lithostatic {Tep, Tmp} per row (md.tep_row when enabled)
    uint8_t* tbc = nullptr;    // thermal boundary policies [face][q] in side order
    double* tval = nullptr;    // thermal boundary data [face][q]
    std::vector<TepPair> tep_host;   // its host source, kept until the slot is reused
    int cur = 0, phase_cur = 0;
    bool phase_valid = false, have_velocity = false;
    double max_rate = 0, dt = 0;
    int rows = 64, strips = 1, nitems = 1;
    ItemLayout layout{};
};

class Engine {
public:
    Engine() = default;
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    ~Engine();

    mp_status init(int capacity, int32_t precision, const mp_weno_options& weno,
                   const mp_phase_params& phase, void* stream, size_t budget_bytes);
    static size_t slot_bytes(const mp_mesh_desc& d, int32_t precision, int rows_min = 8);
    mp_status add_mesh(const mp_mesh_desc& d, int* slot);
    void remove_mesh(int slot);
    mp_status upload(int slot, const double* H, const double* C);
    // checked: the caller has already rejected non-finite values (the
    // server's stagers do it while they copy the inputs).
    mp_status upload_async(int slot, const double* H, const double* C, bool checked = false);
    mp_status download(int slot, double* H, double* C);
    mp_status download_async(int slot, double* H, double* C);
    mp_status set_velocity(int slot, const double* vnx, const double* vny, bool wait = true,
                           bool checked = false);
    mp_status set_boundary(int slot, const uint8_t* const side[4]);
    mp_status set_inflow(int slot, const double* H, const double* C);
    mp_status set_lithostatic(int slot, int enable, double surface_y);
    // Bound limiter of one mesh (StateReconstruction::Limit); enable = 0: off.
    mp_status set_limiter(int slot, int enable, int cell_points, int samples, double extent);
    // Thermal diffusion of H and thermal boundary data of one mesh, as the
    // single-mesh handle's; both refresh the mesh's CFL rate (diffusion terms).
    mp_status set_diffusion(int slot, double kappa, int samples, double extent);
    mp_status set_thermal_boundary(int slot, const uint8_t* types, const double* values);
    // Whether any active slot carries diffusion (the launches' diffusion instance).
    bool any_diffusion() const;
    mp_status refresh_rate(int slot);
    // The limiter argument of the stage launches over the active slots: 0 none
    // limited, 2 every limited mesh on the default tables, else 1.
    int limit_mode() const;
    mp_status phase_out(int slot, double* T, double* phi, double* Cl, double* Cs, uint8_t* region,
                        bool async = false);
    void set_rows(int slot, int rows);
    // Work lists over the given slots (called when membership, tiles or the
    // step pattern change).  steps[i] (optional) is how many of the next
    // enqueued steps slot list[i] takes: at step s only slots with steps > s
    // run, so runs of different lengths share an epoch without padding.
    mp_status set_active(const std::vector<int>& slots, const std::vector<int>& steps = {});
    int steps_of(int slot) const;
    // Enqueue nsteps of `scheme` over the active slots, each with its own dt;
    // a mesh whose step fails the heat balance gets kErrBalance.
    // write_tables uploads the per-slot stage pointers (not capturable);
    // enqueue then issues kernels only, reading device tables, so it may be
    // captured into a graph.  totals receive [step][slot][4] per-step sums.
    mp_status write_tables(int scheme);
    mp_status reserve_steps(int nsteps);
    mp_status enqueue(int nsteps, int scheme, std::vector<hipEvent_t>* events = nullptr);
    // After the enqueued work completes: commit each slot whose error word is 0.
    void commit(int nsteps, const std::vector<uint32_t>& errors);
    mp_status fetch_status(int nsteps, std::vector<uint32_t>& errors, std::vector<double>* totals);
    // Without waiting: the error words and the per-step totals of the steps
    // enqueued so far, copied on the stream into pinned set `set` (0 or 1).
    // They are valid once the stream has passed this point, and stay valid
    // until the next copy into the same set.
    mp_status copy_status(int set, int nsteps);
    // A [scheme-1][nflags] flag array for the current lists (zeroed when it
    // grows; a reallocation synchronises).
    mp_status ensure_flags();
    const uint32_t* copied_errors(int set) const { return h_errors[set]; }
    const double* copied_totals(int set) const { return h_totals[set]; }
    // Per-slot report from the per-step totals (same arithmetic as the
    // single-mesh handle).
    void make_report(int slot, int nsteps, const std::vector<double>& totals, uint32_t bits,
                     mp_step_report* r) const;
    // Work list generation: bumped whenever device tables or buffers move.
    uint64_t generation = 0;

    int32_t precision = MP_PREC_FP64;
    mp_weno_options weno{};
    mp_phase_params phase_params{};
    PhaseHost phase{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // Mixed precision: each step's phase kernel and totals run on `side`,
    // beside the next step's stages (stage3: the step's last stage is done;
    // side_done: its phase and totals are done; by step parity).
    hipStream_t side = nullptr;
    hipEvent_t stage3[2] = {nullptr, nullptr}, side_done[2] = {nullptr, nullptr};
    hipEvent_t side_join = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    std::vector<int> active;         // slots, ordered by steps (descending)
    std::vector<int> active_steps;   // steps per entry of active (INT32_MAX: all)
    std::vector<int> slot_steps;     // per slot: steps in the current pattern (0: idle)
    // Work lists of a prefix of the ordered slots: offsets and counts in the
    // item array, in queue order and in XCD slabs (padded with (-1, 0)).
    struct List {
        int queue_off = 0, queue_count = 0, slab_off = 0, slab_count = 0;
    };
    static constexpr int kXcds = 8;
    std::map<int, List> lists;   // prefix length -> lists
    int capacity = 0;
    int blocks = 0;            // persistent grid in blocks of 4 waves (0: one wave per item)
    // Mixed precision with one wave per item: each step's stages run in one
    // launch (batch_step32_kernel) that orders them with per-item completion
    // flags, instead of one launch per stage.  Profiling keeps the per-stage
    // launches, which it times one by one.
    bool step_launch = false;
    // > 0: only steps whose work list is shorter take the one launch (the
    // single-mesh rule: fewer than eight items per wave slot of the chip;
    // with more, the waits cost what the shared ramp saves)
    int step_launch_items = 0;
    bool uses_step_launch(bool profiling) const
    {
        return step_launch && precision == MP_PREC_MIXED && blocks == 0 && !profiling;
    }
    bool one_launch_for(int list_len, bool profiling) const
    {
        return uses_step_launch(profiling) && (step_launch_items <= 0 || list_len < step_launch_items) &&
               limit_mode() == 0 && !any_diffusion();
    }
    // eight items per wave slot (three waves per SIMD, four SIMDs per CU)
    static int auto_step_items();
    // Before enqueueing or replaying nsteps steps: restarts the epoch counter
    // (and clears the flags) when it would wrap.  Not capturable.
    mp_status reserve_epochs(int nsteps);
    DeviceBuffer d_flags, d_flag_off, d_epoch;
    int nflags = 0;            // flags per stage of the current work lists
    uint64_t epochs_used = 0;  // counter advances since it was last cleared
    int nitems_total = 0;
    int steps_capacity = 0;    // per-step totals reserved
    int q_common = 0;          // q shared by all active slots, or 0
    DeviceBuffer d_meshes, d_stages, d_items, d_nitems, d_active, d_errors, d_totals, d_maxbits;
    DeviceBuffer d_queues;     // per step and stage: work counter of the persistent grid
    // Every step's totals are checked against the totals before it (the
    // preheat's heat balance, src/couple/preheat.cpp:268-273) at this relative
    // tolerance; the totals before a call's first step are formed at its start.
    double balance_tol = 1e-10;
    DeviceBuffer d_init_totals, d_init_part;
    std::vector<int> h_nitems;
    // Uploads of the work lists and of the stage tables (which hold only the
    // active slots' entries; the kernels read no others).
    PinnedRing ring_lists, ring_tables;
    uint32_t* h_errors[2] = {nullptr, nullptr};
    double* h_totals[2] = {nullptr, nullptr};
    int status_steps = 0;      // per-step totals that h_totals holds
};

}  // namespace mp
