// fvm_engine.hpp -- host runtime: a set of cases resident on one GPU, their
// work lists, the launch sequence of one step and its HIP-graph capture.
// The plain solver (fvm_solver) is an engine with one case; the batched
// engine (fvm_batch) holds hundreds.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/fvm/fvm.h"
#include "fvm_device.hpp"
#include "fvm_kernels.hpp"

namespace fvm {

int set_error(int code, const std::string& msg);
// Validates a snapshot file's header and reads its parameters (p may be null).
int read_snapshot_header(const char* path, fvm_snapshot_header* h, fvm_params* p);
int hip_fail(hipError_t e, const char* what);

// Byte offset of the FP64 CFL slots inside a case's 256-byte slot block (the
// FP32 slots are at 0).
constexpr size_t kSlots64Off = 64;

#define FVM_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

// First-fit arena over one device allocation (256-byte granules).  Cases are
// carved out of it so that admission and release never call hipMalloc/hipFree
// while other cases are running.
class Arena {
public:
    ~Arena();
    int init(size_t bytes);
    void* alloc(size_t bytes);
    void release(void* p);
    size_t capacity() const { return cap_; }
    size_t in_use() const { return used_; }
    bool ready() const { return base_ != nullptr; }

private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    std::map<size_t, size_t> live_;  // offset -> size
};

struct HostCase {
    fvm_params p{};
    int id = -1;
    bool active = false;
    int nsets = 0;
    size_t bytes = 0;
    void* mem = nullptr;
    std::vector<int> ghosts;      // cell index of every ghost cell
    void* ghost_mem = nullptr;    // GhostInfo<T> table, then the ghost index map of the bounding box
    size_t ghost_cap = 0;         // bytes
    int gbox[4] = {0, 0, 0, 0};   // ghost bounding box: i0, j0, width, height
    size_t gmap_off = 0;          // byte offset of the index map in ghost_mem
    int tiles_x = 0, tiles_y = 0;
    int steps_total = 0;   // batch: steps to run before completion (0 = unbounded)
    long long steps_done = 0;
    bool finished = false;
    int frozen_set = -1;   // set holding the final state once finished (-1: follows the step parity)
    hipEvent_t ev_frozen = nullptr;   // recorded after the case's last step (batch)
};

struct KernelStat {
    double total_ms = 0;
    long long launches = 0;
};

class Engine {
public:
    Engine(int device, bool fp64, bool kahan);
    ~Engine();

    int init_stream();
    int set_tile(int tx, int ty);
    // Viscous step schedule: false = the first half step as a launch of its
    // own, true = fused into stage 1 (where the tile and precision have it).
    void set_fused_d1(bool on);
    bool fused_d1() const;
    int tile_x() const { return tx_; }
    int tile_y() const { return ty_; }

    // case lifetime
    static size_t case_bytes(const fvm_params& p, bool fp64, bool kahan, bool all_sets);
    int add_case(const fvm_params& p, int steps_total, bool all_sets, int* id);
    int remove_case(int id);
    int reset_case(int id, const fvm_params& p);
    int update_params(int id, const fvm_params& p);   // constants only, same geometry
    int update_obstacle(int id, const fvm_params& p);
    int set_host_dt(int id, float dt);

    // stepping
    int step_direct(int nsteps);        // launches without a graph
    int step_graph(int nsteps);         // replays captured 64-, 16- and 2-step graphs
    int advance(int nsteps, bool use_graph);
    // Builds the work lists and captures the step graphs that step_graph
    // replays (both parities, 2, LONG_STEPS and XL_STEPS steps) without stepping, so a
    // later advance only launches.
    int prepare_graphs();
    int sync();

    // per-case access
    HostCase* get(int id);
    int current_set(const HostCase& c) const;   // set holding U^n
    int latest_slot() const { return phase_; }  // slot holding the maxima of the current state
    int read_state(int id, void* host, bool as_f64);
    int write_state(int id, const void* host);
    int read_cell_types(int id, uint8_t* host);
    int read_sdf(int id, float* host);
    int read_time(int id, double* t, long long* steps);
    // CFL maxima the next step's dt comes from (FP64 maxima for dt64 engines,
    // else the FP32 ones widened)
    int read_maxima(int id, double* max_speed, double* max_nu);
    int queue_maxima(int id);
    int field(int id, int field, float* dev_out);
    int stat(int id, int mode, float* out);
    // generateVectorArrows (solver.cu:3809-3845) into caller device memory;
    // ordered after the caller's null-stream work, complete on return
    int arrows(int id, float* dev_out, int max_vertices, int step, float u_inf, float max_len, float head_angle,
               float head_ratio, int* num_vertices);
    int conservation(int id, double out[4]);
    int forces(int id, double out[3]);   // Fx, Fy, number of wall faces
    int save_case(int id, const char* path);   // snapshot (fvm_snapshot_header layout)
    int load_case(int id, const char* path);   // case must have the snapshot's size and precision

    int active_count() const;
    int case_slots() const { return (int)cases_.size(); }
    void mark_lists_dirty() { lists_dirty_ = true; }
    int retire_finished();
    // Batched serving: admissions and read-backs of finished cases on a side
    // stream, so that neither waits for the steps queued on the main stream.
    void set_side_admission(bool on) { side_admission_ = on; }
    // Keeps at most two advance calls queued: waits for the one before the
    // previous (call at the start of an advance), marks this one's end.
    int throttle_begin();
    int throttle_end();
    size_t bytes_in_use() const;
    int set_arena(size_t bytes);
    int reserve_cases(int n);
    bool fp64() const { return fp64_; }
    bool kahan() const { return kahan_; }
    // CFL maxima and dt formula in FP64 (FP64 state and compensated variant)
    bool dt64() const { return fp64_ || kahan_; }
    hipStream_t stream() const { return stream_; }

    void enable_timing(bool on) { timing_ = on; }
    // Per-kernel durations: events recorded between the launches of a step,
    // with no host wait inside the step; the events of a step are read two
    // steps later (or at flush), so the queue never drains while timing.
    int flush_timings();
    const std::vector<std::pair<std::string, KernelStat>>& timings() const { return timing_list_; }
    void clear_timings();

private:
    int build_consts(HostCase& hc, const fvm_params& p, void* desc_out);
    int upload_desc(HostCase& hc);
    int mark_and_lists(HostCase& hc, int fixup);
    int rebuild_lists();
    int enqueue_step(int phase);
    int capture_graphs();
    int capture_long_graphs();
    int capture_xl_graphs();
    int capture_sequence(int first_phase, int nsteps, hipGraphExec_t* out);
    void drop_graphs();
    void* dev_alloc(size_t bytes);
    void dev_free(void* p);
    int upload_items(const std::vector<WorkItem>& v, WorkItem** buf, size_t* cap, int* d_count, int count_slot);
    // Pinned host copy of one device list's last upload: the copy runs in
    // stream order, so the host rewrites it only after that copy has run.
    struct UploadStaging {
        char* host = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        bool used = false;
    };
    std::map<WorkItem**, UploadStaging> staging_;
    int timed_begin();
    // Start / stop events for the next timed launch (null when not timing).
    void timed_events(hipEvent_t* ev0, hipEvent_t* ev1);
    int timed(const char* name, int rc_before);
    int flush_timed_set(int set);

    int device_;
    bool fp64_, kahan_;
    int tx_ = 64, ty_ = 8;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;       // admissions and read-backs of finished cases
    hipStream_t main_ = nullptr;       // the steps' stream (stream_ outside an admission)
    bool side_admission_ = false;
    hipEvent_t ev_ring_[2] = {nullptr, nullptr};
    bool ring_used_[2] = {false, false};
    int ring_ = 0;
    hipEvent_t ev_null_ = nullptr;   // marks the caller's null-stream work before a readback
    std::vector<HostCase*> cases_;
    int max_cases_ = 0;
    void* d_cases_ = nullptr;        // CaseDesc<T>[max_cases_]
    size_t desc_size_ = 0;
    std::vector<char> h_descs_;      // host mirror

    WorkItem* d_tiles_all_ = nullptr;
    WorkItem* d_tiles_visc_ = nullptr;
    WorkItem* d_ghosts_ = nullptr;
    WorkItem* d_case_list_ = nullptr;   // (case, chunk of REDUCE_CHUNK tiles) items of the per-tile maxima reduction
    WorkItem* scr_tiles_ = nullptr;     // one case's lists for the reset-path launches
    WorkItem* scr_ghosts_ = nullptr;
    int* scr_cnt_ = nullptr;
    size_t scr_tiles_cap_ = 0, scr_ghosts_cap_ = 0;
    int* d_counts_ = nullptr;        // [0] all tiles, [1] viscous tiles, [2] ghosts, [3] active cases
    size_t cap_tiles_ = 0, cap_visc_ = 0, cap_ghosts_ = 0, cap_case_list_ = 0;
    int n_tiles_ = 0, n_visc_ = 0, n_ghosts_ = 0;
    int single_case_ = -1;              // the one active case of a non-persistent engine, else -1
    int grid_tiles_ = 0, grid_visc_ = 0, grid_ghosts_ = 0, grid_reduce_ = 1;
    bool any_visc_ = false;
    bool fuse_d1_ = false;
    bool persistent_ = false;
    int persistent_grid_ = 2048;         // persistent grid when the occupancy query fails
    int cu_count_ = 0;

    int phase_ = 0;                  // step parity: inviscid current set and dt slot
    bool lists_dirty_ = true;
    // Captured step sequences per starting parity: two steps, LONG_STEPS and
    // XL_STEPS steps (captured on first use), all even so the parity returns to
    // its start.  Consecutive graph launches leave a gap of ~8 us on the
    // device, so long runs take the longest graph that fits.
    static constexpr int LONG_STEPS = 16;
    static constexpr int XL_STEPS = 64;
    hipGraphExec_t graph_[2] = {nullptr, nullptr};
    hipGraphExec_t graph_long_[2] = {nullptr, nullptr};
    hipGraphExec_t graph_xl_[2] = {nullptr, nullptr};
    bool graph_valid_ = false;
    bool long_valid_ = false;
    bool xl_valid_ = false;

    Arena arena_;
    bool use_arena_ = false;

    bool timing_ = false;
    static constexpr int kTimedMax = 8;
    struct TimedSet {
        hipEvent_t ev[2 * kTimedMax] = {};   // start / stop of each timed launch
        const char* name[kTimedMax] = {};
        int n = 0;
        bool pending = false;
    };
    TimedSet tset_[2];
    int tcur_ = 0;
    std::vector<std::pair<std::string, KernelStat>> timing_list_;

    unsigned* d_stat_ = nullptr;
    double* d_cons_ = nullptr;
    float* h_pinned_ = nullptr;
};

}  // namespace fvm
