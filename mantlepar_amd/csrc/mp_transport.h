// The single-mesh handle behind mp_transport_* (defined at global scope to
// match the opaque C type).
#pragma once

#include <vector>

#include "mp_host.h"

struct mp_transport {
    mp_mesh_desc mesh{};
    mp_weno_options weno{};
    mp_phase_params phase_params{};
    int32_t precision = MP_PREC_FP64;
    mp::PhaseHost phase{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;               // HIP device current at create time
    mp::MeshDev md{};
    mp::DeviceBuffer d_md;
    size_t N = 0;
    // 0..2: full states (the call's starting state is never written), 3..4: stage buffers
    mp::DeviceBuffer H[5], C[5];
    int cur = 0;
    mp::DeviceBuffer vnx, vny;
    mp::DeviceBuffer fx, fy;   // per-face flux moments
    mp::DeviceBuffer rx, ry;   // and their one-sided 16-byte form
    mp::DeviceBuffer rx64, ry64;   // FP64 handles: one-sided FP64 records
    bool have_velocity = false;
    double max_rate = 0;
    mp::DeviceBuffer bc[4], inflowH, inflowC;
    mp::DeviceBuffer tbc, tval;   // thermal boundary policies and data
    mp::DeviceBuffer tep;         // lithostatic {Tep, Tmp} per cell row
    // phase outputs: T, phi, Cl, Cs (N doubles each) then region (N bytes)
    mp::DeviceBuffer phase_buf[2];
    int phase_cur = 0;
    bool phase_valid = false;
    mp::DeviceBuffer partials, totals, error;
    // One launch per step (mixed precision, see step_fused): completion
    // flags of the handle's own work memory and their epoch counter; the
    // stage records of a call's three step kinds (step 0, odd, even >= 2) in
    // device memory, staged through pinned host memory.
    mp::DeviceBuffer flags;
    uint32_t epoch = 0;
    int32_t launch_mode = MP_LAUNCH_AUTO;
    double balance_tol = 1e-10;   // relative tolerance of a step's balance check
    mp::DeviceBuffer d_tab;
    mp::StagePtrs* h_tab = nullptr;
    hipEvent_t tab_copied = nullptr;
    // Mixed path: each step's phase kernel and totals run on `side`, beside the
    // next step's stages on `stream` (stage3: the step's last stage is done;
    // side_done: its phase and totals are done; by step parity).
    hipStream_t side = nullptr;
    hipEvent_t stage3[2] = {nullptr, nullptr}, side_done[2] = {nullptr, nullptr};
    hipEvent_t side_join = nullptr;   // the main stream waits on it when a call ends
    int rows = 64, strips = 1, nitems = 1;
    mp::ItemLayout layout{};

    mp_transport() = default;
    mp_transport(const mp_transport&) = delete;
    mp_transport& operator=(const mp_transport&) = delete;
    ~mp_transport();

    mp_status init(const mp_mesh_desc& m, const mp_weno_options& w, const mp_phase_params& p,
                   int32_t prec, void* stream);
    mp_status sync_md();
    void set_tile(int rows_per_item);
    int default_rows() const;
    mp_status ensure_scratch(int nsteps);
    mp_status upload(const double* h, const double* c);
    mp_status download(double* h, double* c);
    mp_status set_velocity(const double* x, const double* y);
    mp_status compute_rate();
    mp_status set_boundary(const uint8_t* const side[4]);
    mp_status set_inflow(const double* h, const double* c);
    // Device memory a step works in beyond the states: stage partials of two
    // consecutive steps plus the step reduction's block results (partials),
    // per-step totals [step][4] and the error word.
    // The one-launch step also needs completion flags [3][nitems] (flag_words
    // words in all, holding only epochs) and the host counter of the epochs
    // already used on those flags.
    struct StepWork {
        double* partials;
        double* totals;
        uint32_t* error;
        uint32_t* flags;
        size_t flag_words;
        uint32_t* epoch;
    };
    StepWork own_work();
    static size_t work_bytes(int nitems, int nsteps);
    mp::StagePtrs stage_ptrs(const StepWork& w, int src, int kin, int out, int scheme, int st,
                             double dt, bool phase_stage, int phase_set, int part_stage) const;
    mp_status enqueue_steps(const StepWork& w, int nsteps, int scheme, double dt, int& final_state,
                            std::vector<hipEvent_t>* events);
    mp_status check_step_args(int nsteps, int scheme, double dt);
    mp_status step(int nsteps, int scheme, double dt, mp_step_report* report);
    // step() with its work memory in the library's static device scratch
    // (no device allocation; see mp_transport_step_noalloc).
    mp_status step_noalloc(int nsteps, int scheme, double dt, mp_step_report* report);
    mp_status finish_step(const StepWork& w, int nsteps, int scheme, double dt, int final_state,
                          mp_step_report* report);
    mp_status state_totals(double& th, double& tc);
    mp_status profile(int nsteps, int scheme, double dt, double* ms_total, double* ms_stage,
                      double* ms_phase_stage);
    mp_status phase_out(double* T, double* phi, double* Cl, double* Cs, uint8_t* region);
    mp_status rhs(double* rh, double* rc, double* boundary);
    mp_status reconstruct(int field, double* sigma, double* omega, double* traces);
    mp_status evaluate(int64_t n, const int32_t* cells, const double* x, const double* y,
                       double* outH, double* outC);
    mp_status set_limiter(int enable, int cell_points, int samples, double extent);
    mp_status limits(double* thetaH, double* thetaC);
    mp_status set_diffusion(double kappa, int samples, double extent);
    mp_status set_thermal_boundary(const uint8_t* types, const double* values);
    mp_status set_lithostatic(int enable, double surface_y);
    // thermal diffusion or thermal boundary data are set
    bool has_diffusion() const;
    // this call's steps run as one launch each
    bool step_fused(int scheme) const;
};
