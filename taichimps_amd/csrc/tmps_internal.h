// Internal structures shared by the HIP kernels and the host orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tmps.h"

#define TMPS_BLOCK 256
// items per block of the rebuild's exclusive scans
#define TMPS_SCAN_TILE (TMPS_BLOCK * 4)

// ---------------------------------------------------------------- device state

// Domain device fields (domain.py:55-68), orthogonal box.
struct DevBox {
    double lo[3], hi[3], prd[3];
    double h_rate[6];
    int per[3];
    int vremap;
};

// FixDeformPressure state held on the device (deform_pressure.py:216-226).
struct DevServo {
    tmps_servo_params p;
    double lo_start[3], hi_start[3], cum[3];
    int nsteps, step_count, nsteps_total;
};

// A pending affine remap (remap_positions, deform_pressure.py:493-530): the
// box before and after the end-of-step update of `step`.
struct RemapRec {
    int step;  // -1: none pending
    int pad;
    double old_lo[3], old_prd[3], oh[3], new_lo[3], new_prd[3];
};

// Per-step constants of the FP32-split position encoding
//   x = lo + (prd / g) * (c + 0.5 + u)      (c: integer cell, u: FP32)
struct SplitConst {
    float cw[3];       // prd / g of the current box
    float inv_cw[3];   // g / prd (length -> bin units)
    float dlo[3];      // lo - lo_build           (displacement check)
    float dcw[3];      // (prd - prd_build) / g   (displacement check)
    float cw_build[3]; // prd_build / g
    float g[3];        // grid dimensions as floats
    int gi[3];
    float h_rate[6];
    int pad;
};

struct GridP {
    int g[3];
    double csize[3];
    int ncell;
};

struct Ctl {
    DevBox box;
    DevServo servo;
    RemapRec remap;
    SplitConst sc;
    double tensor[6];   // pressure tensor used by the last servo update
    double vsum[6];     // last reduced virial(+kinetic) sum, or KE pair
    double lo_build[3], prd_build[3];
    int timestep;       // last completed step
    int trigger;        // step whose displacement check asked for a rebuild
    int abort;          // nonzero: the rest of the segment is skipped
    int error;          // TMPS_ERR_BOX after a collapsed box
    int last_build_step;
    int delay, every, check;
    double trigger_sq;
    int moved;          // stage-mode displacement flag
    int ntouch_acc;     // touching slots counted by the force pass
    int ntouch;
    int servo_on;       // end-of-step servo enabled for fused steps
    int fast_img[3];    // per dim: list bins >= 3 at the last build, so the
                        // contact kernels may take image counts by reciprocal
    // an ensemble member's own progress (batch.hip): the buffer set its next
    // step reads, the step that ends its run, a pending segment-start
    // half-kick, and whether its run is complete
    int bpar, bend, bkick, bdone;
    // a list build decided on the device (a run's rebuild redo): its bins
    // and slot count
    GridP lgrid;
    int nslots;
    int lpad;
};

// The control block's error code of a device-decided rebuild that needs the
// host (a list beyond the capacity of its buffers): the run redoes the
// step's list build on the host and continues. Never returned to a caller.
#define TMPS_NEED_HOST -100

// One ensemble member as the batched close kernel sees it.
struct CloseEntry {
    Ctl* ctl;
    const double* partial;
    int nparts, servo_on, natoms, pad;
};

// Pair parameters as the kernels read them.
struct PairP {
    int style, limit_damping;
    double kn, kt, gamman, gammat, xmu;
    int normal_model, damping_model, tangential_model, uses_history;
    double g_kn, g_damp, g_kt, g_xt, g_mu;
};

// One wall as a step applies it: its settings, the plane's position and
// velocities (fixed for a fused run; a cylinder ignores them), and whether it
// comes before fix damping/cundall among the fixes.
struct WallSpec {
    tmps_wall_params p;
    double coord, vwall_n, vwall_s;
    double dt;  // the run's timestep (the history advances by v dt in FP64)
    int before_cundall;
    int pad;
};

// The walls of a fused step, at most TMPS_RUN_WALLS: settings in device
// memory, and per wall the shear history the step reads and the one it
// writes (the two buffer sets of the step, or the same buffer in place).
// nz[i] holds, per wall q and buffer set b, bit wall_nz_bit(q, b): clear only
// where that set's entry of atom i is zero (a set bit may cover a zero).
#define TMPS_RUN_WALLS 2
struct WallRefs {
    const WallSpec* spec;
    const double4* h_in[TMPS_RUN_WALLS];
    double4* h_out[TMPS_RUN_WALLS];
    uint8_t* nz;
    int nwall;
    int in, out;  // the buffer sets h_in and h_out belong to
};

__host__ __device__ constexpr unsigned wall_nz_bit(int wall, int set) {
    return 1u << (2 * wall + set);
}

// One FP32-split ensemble member whose step is redone with a list rebuild,
// as the batched redo kernels read it (the single-system redo's buffers,
// grid and decisions). The host fills it in two passes: before the list
// build, and after the build's slot count is read back (long_rows, codes,
// step, fast, gi, done, and the list's buffers, which the build can move).
struct RedoEntry {
    Ctl* ctl;
    int n, ncell, old_n, onb_cap;
    int copy_blocks, save_blocks, max_nb, servo_on;
    int long_rows, codes, step, done;
    int to_x64, pad;
    int fast[3], gi[3];
    GridP grid;
    double skin;
    PairP pair;
    double* partial;
    double4 *xr, *vm, *x0, *bin_x;
    float4 *up, *up2, *vmf, *vmf2, *omf, *omf2, *ff, *tqf, *u0;
    uint32_t* cell;
    double* vir;
    int *cell_of, *cell_count, *cell_start, *cell_fill, *cell_particles, *scan_tmp;
    int *full_count, *half_count, *flags, *row_tmp;
    int *nb_start, *nb_j, *onb_start, *onb_j;
    uint8_t *nb_flag, *nb_flag2, *onb_flag;
    float4 *nb_hist, *nb_hist2, *onb_hist;
    uint16_t* nb_code;
};

// Block maps of a batched redo: (entry, local block) pairs, per kind of
// item. atoms: nblocks_for(n) per entry; cells: nblocks_for(ncell + 1);
// cell_tiles / atom_tiles: scan tiles over ncell / n; copy: the entries that
// resume from buffer set 1, copy_blocks each; save: save_blocks per entry;
// style: the atoms blocks of the entries with each pair style.
struct RedoMaps {
    int entries, max_parts;
    const int2 *atoms, *cells, *cell_tiles, *atom_tiles, *copy, *save, *style[3];
    int n_atoms, n_cells, n_cell_tiles, n_atom_tiles, n_copy, n_save, n_style[3];
};

// ---------------------------------------------------------------- host system

struct System {
    int device = 0;
    hipStream_t stream = nullptr;      // where the system's work runs
    hipStream_t own_stream = nullptr;  // the stream it created (stream, unless a caller's)
    // the host mirror of the control block equals the device's: set by a
    // getter's pull or the end of a run, cleared by every other entry point
    bool hctl_fresh = false;
    int precision = TMPS_FP64;
    int max_atoms = 0, max_nb = 64, n = 0;

    // FP64 per-atom state (double4 packing: x|r, v|m, omega|0)
    double4 *xr = nullptr, *vm = nullptr, *om = nullptr, *f = nullptr, *tq = nullptr, *x0 = nullptr;
    // second v/omega buffers: a fused force pass reads one and writes the other,
    // since other atoms' contacts still read this step's half-step velocities
    double4 *vm2 = nullptr, *om2 = nullptr;
    double4* xr2 = nullptr;  // second position buffer of the one-kernel step
    double* vir = nullptr;  // [n][6]
    int* tag = nullptr;
    // FP32-split per-atom state
    float4 *up = nullptr, *vmf = nullptr, *omf = nullptr, *ff = nullptr, *tqf = nullptr, *u0 = nullptr;
    float4 *vmf2 = nullptr, *omf2 = nullptr;
    float4* up2 = nullptr;
    uint32_t* cell = nullptr;

    // full neighbor list (CSR) and its history; o* = previous list for the remap
    int *nb_start = nullptr, *nb_j = nullptr, *onb_start = nullptr, *onb_j = nullptr;
    uint8_t *nb_flag = nullptr, *onb_flag = nullptr;
    void *nb_hist = nullptr, *onb_hist = nullptr;  // double4 (FP64) or float4 (split)
    // second touch-flag / history buffers: a one-kernel step reads one set
    // and writes the other, so a rebuild request leaves the first intact
    uint8_t* nb_flag2 = nullptr;
    void* nb_hist2 = nullptr;
    // FP32-split: per slot, i's cell minus j's (nearest image), 3 bits per axis
    uint16_t* nb_code = nullptr;
    size_t nb_cap = 0, onb_cap = 0;
    int nslots = 0, onslots = 0, onb_n = 0;
    bool have_old = false;
    // the o* copy holds a history seeded from a restart file: the next save
    // is skipped so that the following restore carries it onto the new list
    bool seed_pending = false;
    bool x64_valid = false;  // FP32-split: xr holds the authoritative positions

    // bins
    int *cell_of = nullptr, *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr,
        *cell_particles = nullptr;
    int cell_cap = 0;
    // the radix binning's keys, values and (digit, tile) counts
    int *rs_k1 = nullptr, *rs_k2 = nullptr, *rs_v1 = nullptr, *rs_v2 = nullptr;
    int *rs_hist = nullptr, *rs_offs = nullptr;
    size_t rs_na = 0, rs_nh = 0;
    int *full_count = nullptr, *half_count = nullptr, *scan_tmp = nullptr, *flags = nullptr;
    int scan_tmp_cap = 0;

    // reductions
    double* partial = nullptr;  // [blocks][8]
    int partial_cap = 0;
    double* level = nullptr;  // [groups of 4096 partials][8], large systems only
    int level_cap = 0;

    Ctl* ctl = nullptr;   // device control block
    Ctl* hctl = nullptr;  // pinned host mirror
    Ctl edit_base{};      // the mirror as a ctl_edit found it

    // host mirrors
    std::vector<double> h_radius;
    double max_r = 0.0;
    bool max_r_valid = false;
    int* row_tmp = nullptr;  // single-pass list build: slot-major candidate rows
    double4* bin_x = nullptr;  // list build: positions in bin order
    size_t row_tmp_cap = 0;
    DevBox box{};
    bool box_on_device_newer = false;

    double skin = 0.001;
    int delay = 0, every = 1, check = 1, max_grid_cells = 200000;
    GridP grid{};      // bins of the last list build (also the split encoding's bins)
    GridP ref_grid{};  // the reference's grid (max_grid_cells cap), reported by the API
    int has_built = 0, build_count = 0, last_build_step = -1, current_step = 0;
    int timestep = 0, setup_done = 0, rebuilds = 0, npairs = 0;

    PairP pair{};
    bool have_pair = false;
    tmps_servo_params servo{};
    bool have_servo = false;
    int servo_step_count = 0, servo_nsteps = 0;

    // captured step segments, keyed by their launch configuration
    static const int kGraphSlots = 16;
    hipGraphExec_t seg_exec[kGraphSlots] = {};
    uint64_t seg_key[kGraphSlots] = {};
    int seg_next = 0;
    int* hstatus = nullptr;  // pinned: [slot][timestep, trigger, abort, error]
    double* hsum = nullptr;  // pinned: 8 doubles, a reduction result read back
    hipEvent_t seg_event[kGraphSlots] = {};
    // the longest segment a run launches, from the steps between its recent
    // list rebuilds: a rebuild request leaves the rest of its segment and
    // the one queued behind it to return at once, so frequent rebuilds call
    // for short segments (and rare ones for long)
    int seg_max = 16;
    int last_rebuild_step = -1;
    double rebuild_gap = 64.0;

    int launch_mode = TMPS_LAUNCH_GRAPH;

    // walls a device-resident run applies in its fused step (tmps_run_set_walls):
    // their settings (host copy and device copy) and per wall the shear
    // history of buffer set 0 (the wall's own) and of set 1
    int nrun_walls = 0;
    WallSpec run_spec[TMPS_RUN_WALLS] = {};
    WallSpec* d_run_spec = nullptr;
    double4* run_wh[TMPS_RUN_WALLS][2] = {};
    uint8_t* run_wnz = nullptr;  // [max_atoms]: WallRefs::nz of the run's walls
    double* d_radius = nullptr;  // [max_atoms]: h_radius staged for tmps_set_field_device

    // the adjoint engine's checkpoints and work buffers (adjoint.hip)
    struct AdjContext* adj = nullptr;
    // steps at which the list was built while log_builds is set (the adjoint
    // forward pass logs them for its replays)
    std::vector<int> build_log;
    bool log_builds = false;

    // outputs of the per-contact computes, grown on first use: per-atom
    // counts, per-atom stress [n][6], and the reduced sums (8 doubles)
    int* cmp_count = nullptr;
    double* cmp_stress = nullptr;
    double* cmp_sum = nullptr;
    int cmp_count_cap = 0, cmp_stress_cap = 0;

    // profiling
    bool profiling = false;
    // one-shot: the next step kernel records its own start / stop time
    hipEvent_t kstart = nullptr, kstop = nullptr;
    // per class (kick, step pass, segment tail, step close): summed ms, count
    double prof_ms[TMPS_PROF_CLASSES] = {0, 0, 0, 0};
    int prof_cnt[TMPS_PROF_CLASSES] = {0, 0, 0, 0};
};

// Host-side stages of api.hip used by the adjoint engine (adjoint.hip).
namespace tmps_api {
int ctl_pull(System& s);
int ctl_push(System& s);
int check_run_error(System& s);
int ensure_split(System& s);
// pull_after: read the control block back (timestep, error) when the redo
// is done; without it the host mirror is current only in the box and step
int redo_step(System& s, int par, double dt, int use_cundall, double gl, double ga,
              bool pull_after = true);
int redo_tail(System& s, int stp, double dt, int use_cundall, double gl, double ga,
              bool pull_after);
// A run's rebuild redo with the list build decided on the device (api.hip).
struct DevRedo {
    bool ok = false;       // the system's buffers allow it this run
    bool built = false;    // some redo of this run built a list on the device
    bool pending = false;  // the last such redo's outcome is not yet read
    double max_cutoff = 0.0;
    int max_cells = 0;
    // the host record before the last redo, taken back should it fail
    int stp = 0, build_count = 0, last_build_step = 0, has_built = 0;
    bool seed_pending = false, have_old = false;
};
int init_simulation(System& s, double dt, int use_cundall, double gl, double ga);
int reneighbor(System& s, bool force, int* rebuilt);
// reneighbor's build without its decision: a rebuild at step s.timestep
int reneighbor_now(System& s);
int clear_forces(System& s);
int pair_compute(System& s, double dt, int shearupdate, int accumulate);
int servo_end_of_step(System& s, double dt, double* tensor_out);
// servo_end_of_step's kernels only: the host mirror is not read back and a
// collapsed box is left in the control block's error field
int servo_end_of_step_launch(System& s, double dt);
int run(System& s, int nsteps, double dt, int run_pre, int use_cundall, double gl, double ga);
// redo_step's host work for an FP32-split ensemble member whose step the
// batched redo kernels redo: redo_batch_begin up to the list build (from the
// control block in s.hctl), redo_batch_list after it, given the build's
// flags (overflow, long row, slot count).
int redo_batch_begin(System& s, RedoEntry& e);
int redo_batch_list(System& s, int4 flags, RedoEntry& e);
}  // namespace tmps_api

// grows the list buffers to hold `need` slots (k_rebuild.hip)
int ensure_list_capacity(System& s, size_t need);
// sizes the single-pass list build's scratch rows for s.n atoms (k_rebuild.hip)
int list_reserve_rows(System& s);
// frees the adjoint engine's buffers (adjoint.hip)
void adjoint_free(System& s);

// ---------------------------------------------------------------- launchers

#define HIP_TRY(expr)                                                   \
    do {                                                                \
        hipError_t _e = (expr);                                         \
        if (_e != hipSuccess) return tmps_fail_hip(_e, #expr, __LINE__); \
    } while (0)

int tmps_fail_hip(hipError_t e, const char* what, int line);
int tmps_fail(int code, const std::string& msg);

// A copy ordered on the given stream and complete on return. Host code never
// uses the null stream: HIP refuses its synchronous copies and fills
// (hipErrorStreamCaptureImplicit) while any thread of the process captures a
// graph, and systems driven from several host threads capture at any time.
inline hipError_t copy_sync(hipStream_t st, void* dst, const void* src, size_t bytes,
                            hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    return e != hipSuccess ? e : hipStreamSynchronize(st);
}

inline int nblocks_for(int n) { return (n + TMPS_BLOCK - 1) / TMPS_BLOCK; }

// MI355X hands workgroup w to XCD (w mod 8). Blocks of consecutive atoms share
// neighbours (and so L2 lines), so each XCD is given one contiguous range of
// blocks: a bijection of [0, nb) onto itself.
__device__ __forceinline__ int xcd_block(int w, int nb) {
    const int q = nb >> 3, r = nb & 7, x = w & 7, k = w >> 3;
    return x * q + (x < r ? x : r) + k;
}

// The fused run's walls as a step that reads buffer set `in` and writes set
// `out` sees them (in == out: in place).
inline WallRefs run_walls(const System& s, int in, int out) {
    WallRefs W{};
    W.spec = s.d_run_spec;
    W.nwall = s.nrun_walls;
    W.nz = s.run_wnz;
    W.in = in;
    W.out = out;
    for (int q = 0; q < s.nrun_walls; ++q) {
        W.h_in[q] = s.run_wh[q][in];
        W.h_out[q] = s.run_wh[q][out];
    }
    return W;
}

// rebuild (k_rebuild.hip)
int launch_pbc(System& s);
int launch_bins(System& s, const double4* xr);
// wrapped: every position lies inside the box on its periodic axes (the PBC
// wrap ran since the positions last changed)
int launch_list(System& s, const double4* xr, int* overflow_out, int* total_out, bool wrapped);
int launch_hist_restore(System& s);
// the current list, touch flags and history into the o* copy (save_state)
int launch_list_save(System& s);
int launch_store_x0(System& s);
// exclusive, out has n+1 (out[n] = total, also written to total_copy when set)
int launch_scan(System& s, const int* in, int* out, int n, int* total_copy);
int launch_half_pairs(System& s, int* d_pair_i, int* d_pair_j, int* d_partner, double* d_shear,
                      int* d_counts);
int launch_set_half_history(System& s, const int* d_partner, const double* d_shear,
                            const int* d_half_start);
int launch_hist_clear(System& s);

// FP64 path (k_fp64.hip)
int launch_clear_f64(System& s);
int launch_nve_initial_f64(System& s, double dt);
int launch_nve_final_f64(System& s, double dt);
int launch_cundall_f64(System& s, double gl, double ga);
int launch_pair_f64(System& s, double dt, int shearupdate, int accumulate);
int launch_check_f64(System& s);
int launch_leaves_f64(System& s, int keflag);
int launch_ke_f64(System& s, double* out2);
int launch_step_begin_f64(System& s, double dt, int kick, int buf);
int launch_step_force_f64(System& s, double dt, int final_mode, int cundall, double gl, double ga,
                          int in_buf, int out_buf);
// copy v/omega buffer 1 -> buffer 0 (skipped on abort when honor_abort)
int launch_copy_vw(System& s, int honor_abort);

// one kernel per step (k_fp64.hip, k_split.hip): in_buf selects the buffer
// set the step reads (the other one receives its results)
int launch_fused_f64(System& s, int in_buf, double dt, int final_mode, int cundall, double gl,
                     double ga);
int launch_fused_split(System& s, int in_buf, double dt, int final_mode, int cundall, double gl,
                       double ga);
int launch_step_close(System& s, double dt, int servo_on);

// ensemble (batch.hip): one launch advances every member that is running
size_t batch_split_entry_size();
int batch_split_entry(System& s, void* host_entry);  // fills one table entry
int launch_batch_split_kick(const void* table, const int2* blocks, int nblocks, float dt,
                            hipStream_t st);
int launch_batch_split_step(int style, const void* table, const int2* blocks, int nblocks,
                            float dt, hipStream_t st);
int launch_batch_split_tail(const void* table, const int2* blocks, int nblocks, hipStream_t st);
int launch_batch_close(const CloseEntry* table, int ninst, int max_parts, double dt,
                       hipStream_t st);
// The batched redo of ensemble members' rebuild steps, in redo_step's order:
// resume (buffer set 1 -> 0 where needed, trigger and abort cleared), the
// step's start and the FP64 positions, the list save, PBC wrap, bins, list
// rows with their flags read back (the host then sizes the lists), the CSR
// rows and x0, the build record, the split encoding, the history restore,
// the force pass, the step's end, v / omega into set 0, and the rejoin.
int launch_redo_resume(const RedoEntry* T, const RedoMaps& M, hipStream_t st);      // k_fp64
int launch_redo_begin_split(const RedoEntry* T, const RedoMaps& M, float dt,
                            hipStream_t st);                                         // k_split
int launch_redo_save(const RedoEntry* T, const RedoMaps& M, hipStream_t st);        // k_rebuild
int launch_redo_pbc(const RedoEntry* T, const RedoMaps& M, hipStream_t st);         // k_rebuild
int launch_redo_bins(const RedoEntry* T, const RedoMaps& M, hipStream_t st);        // k_rebuild
int launch_redo_rows(const RedoEntry* T, const RedoMaps& M, int4* flags_out,
                     hipStream_t st);                                                // k_rebuild
int launch_redo_list(const RedoEntry* T, const RedoMaps& M, hipStream_t st);        // k_rebuild
int launch_redo_built(const RedoEntry* T, const RedoMaps& M, hipStream_t st);       // k_fp64
int launch_redo_encode_split(const RedoEntry* T, const RedoMaps& M, hipStream_t st);  // k_split
int launch_redo_restore(const RedoEntry* T, const RedoMaps& M, hipStream_t st);     // k_rebuild
int launch_redo_force_split(const RedoEntry* T, const RedoMaps& M, float dt,
                            hipStream_t st);                                         // k_split
int launch_redo_end(const RedoEntry* T, const RedoMaps& M, double dt, hipStream_t st);  // k_fp64
int launch_redo_copy_vw_split(const RedoEntry* T, const RedoMaps& M, hipStream_t st);  // k_split
int launch_redo_rejoin(const RedoEntry* T, const RedoMaps& M, hipStream_t st);      // k_fp64
// the initial_integrate half-kick of a segment's first step, in buffer 0
int launch_kick_f64(System& s, double dt);
int launch_kick_split(System& s, double dt);
// buffer set 1 -> buffer set 0 (positions, v, omega, touch flags, history);
// skipped on abort when honor_abort
int launch_copy_state(System& s, int honor_abort);
// WallRefs::nz of the run's walls from their histories (a run's start)
int launch_walls_nz(System& s);

// shared (k_fp64.hip)
int launch_step_end(System& s, double dt, int nparts, int advance, int apply_servo);
int launch_remap_apply(System& s);
// clears the pending remap record (kept when honor_abort and the step aborted:
// the redo of that step still needs it)
int launch_remap_clear(System& s, int honor_abort = 0);
// a redone step resumes after its rebuild (FP32-split: with the split
// constants of the current box)
int launch_ctl_rebuilt(System& s, int step);
// clears trigger and abort before a redo
int launch_ctl_resume(System& s);
int launch_ctl_store(System& s);
// records a list build (last_build_step, fast_img; FP32-split: sc.gi and the
// split constants with the current box as the skin check's reference)
int launch_ctl_built(System& s, int step, const int fast[3], const int gi[3], int split);
// the same for a list build decided on the device (its grid from the control
// block; nothing on a recorded error)
int launch_ctl_built_dev(System& s, int step);
// the binning's work buffers for the current atom count (k_rebuild.hip)
int bins_reserve(System& s);
// a list build decided on the device (k_rebuild.hip)
int launch_build_dev(System& s, double max_cutoff, int max_cells, bool store_x0);
// the split encoding after it (k_split.hip)
int launch_x64_to_split_dev(System& s);
int launch_remap_apply_fused(System& s);
void drop_graphs(System& s);
int launch_final_tree(System& s, int nc, int nparts);
// per-contact computes on the FP64 positions in xr (k_fp64.hip)
int launch_contact_atom(System& s, int* d_count);
int launch_fabric(System& s, double* d_sum7);
// launch_fabric with the block partials in `partial` (nblocks_for(n) x 8
// doubles, at most CLOSE_SPAN blocks) instead of the system's buffer
int launch_fabric_into(System& s, double* partial, double* d_sum7);
int launch_stress_atom(System& s, int keflag, double* d_stress, double* d_sum6);
// fix wall/gran on an FP64 system; wh: per-atom shear history + touch flag
int launch_wall_gran(System& s, double4* wh, const tmps_wall_params& p, double dt, double coord,
                     double vwall_n, double vwall_s, int history_update);

// FP32-split path (k_split.hip)
int launch_split_to_x64(System& s);
int launch_x64_to_split(System& s);
int launch_step_begin_split(System& s, double dt, int kick, int buf);
// a redone step's start on an FP32-split system: launch_step_begin_split (no
// kick), launch_split_to_x64 and launch_pbc in one pass
int launch_redo_begin_split(System& s, double dt);
// the same on an FP64 system: launch_step_begin_f64 (no kick) and launch_pbc
int launch_redo_begin_f64(System& s, double dt);
// field records to and from a caller's device buffer (tmps_*_field_device)
int launch_field_pack(System& s, const double4* d4, const float4* f4, int n, int scalar,
                      double* out);
int launch_field_unpack(System& s, double4* d4, float4* f4, int n, const double* in,
                        const double* wsrc);
int launch_step_force_split(System& s, double dt, int final_mode, int cundall, double gl, double ga,
                            int in_buf, int out_buf);
int launch_copy_vw_split(System& s, int honor_abort);
int launch_pair_split(System& s, double dt, int shearupdate, int accumulate);
int launch_nve_initial_split(System& s, double dt);
int launch_nve_final_split(System& s, double dt);
int launch_cundall_split(System& s, double gl, double ga);
int launch_check_split(System& s);
int launch_leaves_split(System& s, int keflag);
int launch_ke_split(System& s, double* out2);
int launch_clear_split(System& s);
int launch_split_consts(System& s, int set_build);
int launch_split_fields_to_f64(System& s, int field, double* d_out);
int launch_split_fields_from_f64(System& s, int field, const double* d_in);

// Orders LDS accesses of one wavefront: its lanes' earlier writes are visible
// to its later reads.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-step constants of the split encoding from the current box; with
// set_build the current box also becomes the reference of the skin check.
__device__ __forceinline__ void update_split_consts(Ctl* ctl, int set_build) {
    SplitConst sc = ctl->sc;
    const DevBox B = ctl->box;
    double lo_b[3], prd_b[3];
    for (int k = 0; k < 3; ++k) {
        lo_b[k] = set_build ? B.lo[k] : ctl->lo_build[k];
        prd_b[k] = set_build ? B.prd[k] : ctl->prd_build[k];
    }
    for (int k = 0; k < 3; ++k) {
        const double g = (double)(sc.gi[k] > 0 ? sc.gi[k] : 1);
        sc.g[k] = (float)g;
        sc.cw[k] = (float)(B.prd[k] / g);
        sc.inv_cw[k] = (float)(g / B.prd[k]);
        sc.dlo[k] = (float)(B.lo[k] - lo_b[k]);
        sc.dcw[k] = (float)((B.prd[k] - prd_b[k]) / g);
        sc.cw_build[k] = (float)(prd_b[k] / g);
    }
    ctl->sc = sc;
    if (set_build)
        for (int k = 0; k < 3; ++k) {
            ctl->lo_build[k] = lo_b[k];
            ctl->prd_build[k] = prd_b[k];
        }
}
