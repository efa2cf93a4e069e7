// C ABI of libtmps_hip.so: system lifetime, field transfer, the stage entry
// points of the DEM path, and the device-resident fused runner.
//
// Host arithmetic that feeds the kernels (the bin grid of
// NeighborList.setup_grid, the strain-style h_rate of FixDeformPressure.setup)
// is built with -ffp-contract=off like the kernels, so it rounds as the
// reference's Python does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tmps_internal.h"

// The longest captured segment of a device-resident run's steps (a run's
// last segment may take up to twice as many). A rebuild request wastes on
// average half a segment plus the one queued behind it (their kernels exit
// at once), while each segment costs one status read.
#define TMPS_SEG_LEN 16

namespace {
thread_local std::string g_err;

// An entry point's system. Every entry point but the getters may change the
// device's control block, so it marks the host mirror stale.
System* sys(tmps_system* h) {
    System* s = reinterpret_cast<System*>(h);
    s->hctl_fresh = false;
    return s;
}

// The system of an entry point that enqueues no work that changes the
// device's control block (getters, host-only settings), or that keeps track
// of the host mirror itself (ctl_edit / ctl_commit, run): the mirror's
// freshness stays as it is.
System* sys_view(tmps_system* h) { return reinterpret_cast<System*>(h); }
}  // namespace

int tmps_fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int tmps_fail_hip(hipError_t e, const char* what, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at api line %d", (int)e, hipGetErrorString(e),
             what, line);
    g_err = buf;
    return TMPS_ERR_HIP;
}

namespace tmps_api {

template <typename T>
int dalloc(hipStream_t st, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc((void**)p, count * sizeof(T)));
    HIP_TRY(hipMemsetAsync(*p, 0, count * sizeof(T), st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int ctl_pull(System& s) {
    HIP_TRY(hipStreamSynchronize(s.stream));
    HIP_TRY(copy_sync(s.stream, s.hctl, s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    return 0;
}

// ctl_pull for the getters: no round trip while the mirror is known fresh
// (no entry point that could have changed the device's copy ran since).
int ctl_pull_view(System& s) {
    if (s.hctl_fresh) return 0;
    int rc = ctl_pull(s);
    if (!rc) s.hctl_fresh = true;
    return rc;
}

int ctl_push(System& s) {
    HIP_TRY(copy_sync(s.stream, s.ctl, s.hctl, sizeof(Ctl), hipMemcpyHostToDevice));
    return 0;
}

// An edit of the control block without a round trip while the host mirror is
// fresh: ctl_edit brings the mirror up to date only when enqueued work may
// have changed the device's copy; ctl_commit stores the edited mirror into
// the device's copy in stream order (a kernel argument, no synchronisation),
// after which the mirror is still the device's state at the end of the
// enqueued work.
int ctl_edit(System& s) {
    const int rc = ctl_pull_view(s);
    if (!rc) std::memcpy(&s.edit_base, s.hctl, sizeof(Ctl));
    return rc;
}

// (an edit that changed nothing stores nothing)
int ctl_commit(System& s) {
    if (std::memcmp(&s.edit_base, s.hctl, sizeof(Ctl)) == 0) return 0;
    const int rc = launch_ctl_store(s);
    if (rc) s.hctl_fresh = false;
    return rc;
}

bool split(const System& s) { return s.precision == TMPS_FP32_SPLIT; }

// NeighborList.setup_grid, neighbor.py:143-189 (orthogonal perpendicular
// widths, :122-141), evaluated in the reference's order.
void setup_grid(System& s, const DevBox& B, double max_cutoff) {
    double cut = max_cutoff + s.skin;
    if (cut <= 0) cut = 1.0;
    const double xprd = B.prd[0], yprd = B.prd[1], zprd = B.prd[2];
    const double vol = xprd * yprd * zprd;
    const double c0 = yprd * zprd - 0.0 * 0.0;
    const double ca = zprd * xprd - 0.0 * 0.0;
    const double c2 = xprd * yprd - 0.0 * 0.0;
    const double w0 = vol / std::sqrt(c0 * c0);
    const double w1 = vol / std::sqrt(ca * ca);
    const double w2 = vol / std::sqrt(c2 * c2);
    int gx = std::max(1, (int)std::floor(w0 / cut));
    int gy = std::max(1, (int)std::floor(w1 / cut));
    int gz = std::max(1, (int)std::floor(w2 / cut));
    if (B.per[0] == 1 && gx == 2) gx = 1;
    if (B.per[1] == 1 && gy == 2) gy = 1;
    if (B.per[2] == 1 && gz == 2) gz = 1;
    const long long total = (long long)gx * gy * gz;
    if (total > s.max_grid_cells) {
        const double factor = std::pow((double)s.max_grid_cells / (double)total, 1.0 / 3.0);
        gx = std::max(1, (int)(gx * factor));
        gy = std::max(1, (int)(gy * factor));
        gz = std::max(1, (int)(gz * factor));
        // The scaling can bring a periodic axis back to 2 bins, where the -1
        // and +1 stencil offsets wrap onto the same bin and every candidate
        // there would be listed twice. Collapse it again (the reference
        // collapses only before scaling, neighbor.py:166-178).
        if (B.per[0] == 1 && gx == 2) gx = 1;
        if (B.per[1] == 1 && gy == 2) gy = 1;
        if (B.per[2] == 1 && gz == 2) gz = 1;
    }
    s.grid.g[0] = gx;
    s.grid.g[1] = gy;
    s.grid.g[2] = gz;
    s.grid.csize[0] = 1.0 / gx;
    s.grid.csize[1] = 1.0 / gy;
    s.grid.csize[2] = 1.0 / gz;
    s.grid.ncell = gx * gy * gz;
}

int ensure_cells(System& s, int ncell) {
    if (ncell + 1 <= s.cell_cap) return 0;
    dfree(s.cell_count);
    dfree(s.cell_start);
    dfree(s.cell_fill);
    // slack, so that a box that grows a little under the servo does not
    // reallocate at a rebuild
    int cap = ncell + ncell / 4 + 1;
    int rc = dalloc(s.stream, &s.cell_count, cap);
    if (!rc) rc = dalloc(s.stream, &s.cell_start, cap);
    if (!rc) rc = dalloc(s.stream, &s.cell_fill, cap);
    if (rc) return rc;
    s.cell_cap = cap;
    const int need_scan = (std::max(cap, s.max_atoms + 1) + 1023) / 1024 + 64;
    if (need_scan > s.scan_tmp_cap) {
        dfree(s.scan_tmp);
        rc = dalloc(s.stream, &s.scan_tmp, need_scan);
        if (rc) return rc;
        s.scan_tmp_cap = need_scan;
    }
    return 0;
}

void invalidate_graphs(System& s) { drop_graphs(s); }

// max radius of the local atoms (NeighborList.build, neighbor.py:419-420),
// cached until a radius or the atom count changes
double max_radius(System& s) {
    if (s.n <= 0) return 1.0;
    if (!s.max_r_valid) {
        double m = s.h_radius[0];
        for (int i = 1; i < s.n; ++i) m = std::max(m, s.h_radius[i]);
        s.max_r = m;
        s.max_r_valid = true;
    }
    return s.max_r;
}

// FP32-split: bring the split encoding up to date with xr when xr holds the
// authoritative positions (after a position upload, a PBC wrap or a box set).
int ensure_split(System& s) {
    if (!split(s) || !s.x64_valid) return 0;
    int rc = ctl_pull(s);
    if (rc) return rc;
    if (s.grid.ncell == 0) {
        setup_grid(s, s.hctl->box, 2.0 * max_radius(s));
        rc = ensure_cells(s, s.grid.ncell);
        if (rc) return rc;
    }
    if (s.grid.g[0] > 2047 || s.grid.g[1] > 2047 || s.grid.g[2] > 1023)
        return tmps_fail(TMPS_ERR_UNSUPPORTED,
                         "FP32-split encoding holds at most 2047 x 2047 x 1023 bins");
    for (int k = 0; k < 3; ++k) s.hctl->sc.gi[k] = s.grid.g[k];
    const bool first = !s.has_built;
    rc = ctl_push(s);
    if (rc) return rc;
    if (first) {
        rc = launch_store_x0(s);
        if (!rc) rc = launch_split_consts(s, 1);
        if (rc) return rc;
    } else {
        rc = launch_split_consts(s, 0);
        if (rc) return rc;
    }
    rc = launch_x64_to_split(s);
    if (rc) return rc;
    s.x64_valid = false;
    s.hctl_fresh = false;  // the split constants were set on the device
    return 0;
}

// FP32-split: make xr hold the current positions (FP64) without changing
// which representation is authoritative.
int split_positions_to_xr(System& s) {
    if (!split(s) || s.x64_valid) return 0;
    return launch_split_to_x64(s);
}

// ------------------------------------------------------------ neighbor build

// The bins of a list build, from the box in the host mirror: the reference's
// grid (reported by the API) and the build's own, with bin arrays that hold
// it.
int neigh_grid(System& s) {
    const DevBox B = s.hctl->box;
    const double maxcut = 2.0 * max_radius(s);
    setup_grid(s, B, maxcut);
    s.ref_grid = s.grid;
    // The list is built on the same lamda grid without the reference's
    // 200 000-bin cap (up to 4 bins per atom): every bin is still at least
    // one cutoff wide, so the 27-bin stencil finds the same neighbor set while
    // scanning ~4x fewer candidates at 1M atoms.
    {
        const int keep = s.max_grid_cells;
        s.max_grid_cells = std::max(keep, (int)std::min<long long>(4LL * s.n, 1LL << 26));
        setup_grid(s, B, maxcut);
        s.max_grid_cells = keep;
    }
    return ensure_cells(s, s.grid.ncell);
}

// NeighborList.build's overflow error (neighbor.py:426-431)
int list_overflow_error(const System& s) {
    char buf[256];
    snprintf(buf, sizeof buf,
             "Neighbor list overflow: more than %d neighbors for at least one particle. "
             "Increase max_neighbors_per_atom, or reduce the neighbor skin.",
             s.max_nb);
    return tmps_fail(TMPS_ERR_OVERFLOW, buf);
}

// The host's record of a list build of `total` slots at s.current_step
// (counters, host mirror); fast / gi receive what the control block records.
int neigh_record(System& s, int total, int fast[3], int gi[3]) {
    s.npairs = total / 2;
    s.has_built = 1;
    s.build_count += 1;
    s.last_build_step = s.current_step;
    if (s.log_builds) s.build_log.push_back(s.current_step);
    // the copy of the list the next rebuild saves for the history remap,
    // allocated with the list (not inside a later rebuild)
    if (!s.onb_j && s.nb_cap > 0) {
        const size_t hsz = split(s) ? sizeof(float4) : sizeof(double4);
        HIP_TRY(hipMalloc(&s.onb_j, s.nb_cap * sizeof(int)));
        HIP_TRY(hipMalloc(&s.onb_flag, s.nb_cap));
        HIP_TRY(hipMalloc(&s.onb_hist, s.nb_cap * hsz));
        s.onb_cap = s.nb_cap;
    }
    // the build's record in the control block (device and host mirror)
    for (int k = 0; k < 3; ++k) {
        fast[k] = s.grid.g[k] >= 3 ? 1 : 0;
        gi[k] = s.grid.g[k];
    }
    if (split(s) && (s.grid.g[0] > 2047 || s.grid.g[1] > 2047 || s.grid.g[2] > 1023))
        return tmps_fail(TMPS_ERR_UNSUPPORTED,
                         "FP32-split encoding holds at most 2047 x 2047 x 1023 bins");
    s.hctl->last_build_step = s.last_build_step;
    for (int k = 0; k < 3; ++k) {
        s.hctl->fast_img[k] = fast[k];
        if (split(s)) s.hctl->sc.gi[k] = gi[k];
    }
    return 0;
}

// box_known: the host mirror of the control block is current (the caller
// pulled it and nothing since changed the box), so no round trip is needed.
// wrapped: the PBC wrap ran just before (reneighbor), so every position lies
// inside the box on its periodic axes.
int neigh_build(System& s, bool box_known = false, bool wrapped = false) {
    if (s.n <= 0) return 0;
    int rc = 0;
    if (split(s)) {
        rc = split_positions_to_xr(s);
        if (rc) return rc;
        s.x64_valid = true;
    }
    if (!box_known) {
        rc = ctl_pull(s);
        if (rc) return rc;
    }
    rc = neigh_grid(s);
    if (rc) return rc;
    rc = launch_bins(s, s.xr);
    if (rc) return rc;
    int overflow = 0, total = 0;
    rc = launch_list(s, s.xr, &overflow, &total, wrapped);
    if (rc) return rc;
    if (overflow) return list_overflow_error(s);
    rc = launch_hist_clear(s);
    if (rc) return rc;
    rc = launch_store_x0(s);
    if (rc) return rc;
    int fast[3], gi[3];
    rc = neigh_record(s, total, fast, gi);
    if (rc) return rc;
    rc = launch_ctl_built(s, s.last_build_step, fast, gi, split(s) ? 1 : 0);
    if (rc) return rc;
    if (split(s)) {
        rc = launch_x64_to_split(s);
        if (rc) return rc;
        s.x64_valid = false;
    }
    return 0;
}

// The copy of the list that the history restore merges with, sized for the
// current list.
int history_reserve(System& s) {
    const size_t need = (size_t)std::max(s.nslots, 0);
    const size_t hsz = split(s) ? sizeof(float4) : sizeof(double4);
    if (need > s.onb_cap || s.onb_j == nullptr) {
        dfree(s.onb_j);
        dfree(s.onb_flag);
        if (s.onb_hist) (void)hipFree(s.onb_hist);
        s.onb_hist = nullptr;
        const size_t cap = need + need / 4 + 1024;
        int rc = dalloc(s.stream, &s.onb_j, cap);
        if (!rc) rc = dalloc(s.stream, &s.onb_flag, cap);
        if (rc) return rc;
        HIP_TRY(hipMalloc(&s.onb_hist, cap * hsz));
        s.onb_cap = cap;
    }
    return 0;
}

int history_save(System& s) {
    // a seeded history stands in for the save (save_state returns early while
    // setup_pending, contact_history.py:141-145)
    if (s.seed_pending) {
        s.seed_pending = false;
        return 0;
    }
    // keep a copy of the current list: the restore merges the new list with it
    const size_t need = (size_t)std::max(s.nslots, 0);
    int rc = history_reserve(s);
    if (!rc) rc = launch_list_save(s);
    if (rc) return rc;
    s.onslots = (int)need;
    s.onb_n = s.n;
    s.have_old = true;
    return 0;
}

int history_restore(System& s) {
    if (!s.have_old) return 0;
    int rc = launch_hist_restore(s);
    s.have_old = false;
    s.seed_pending = false;
    return rc;
}

// ctl_fresh: the host mirror of the control block is current.
int pbc(System& s, bool ctl_fresh = false) {
    if (s.n <= 0) return 0;
    if (split(s)) {
        if (!ctl_fresh) {
            int rc = ctl_pull(s);
            if (rc) return rc;
        }
        int rc = 0;
        if (s.hctl->box.vremap)
            return tmps_fail(TMPS_ERR_UNSUPPORTED, "remap v is not supported with FP32-split");
        rc = split_positions_to_xr(s);
        if (rc) return rc;
        s.x64_valid = true;
    }
    return launch_pbc(s);
}

int neigh_check(System& s, int* moved) {
    if (s.n <= 0) {
        *moved = 0;
        return 0;
    }
    int rc = ensure_split(s);
    if (rc) return rc;
    rc = ctl_pull(s);
    if (rc) return rc;
    s.hctl->moved = 0;
    s.hctl->trigger_sq = 0.25 * s.skin * s.skin;
    rc = ctl_push(s);
    if (rc) return rc;
    rc = split(s) ? launch_check_split(s) : launch_check_f64(s);
    if (rc) return rc;
    rc = ctl_pull(s);
    if (rc) return rc;
    *moved = s.hctl->moved;
    return 0;
}

// NeighborList.decide, neighbor.py:371-403
int neigh_decide(System& s, int timestep, int have_timestep, int* rebuild) {
    *rebuild = 0;
    if (!s.has_built) {
        *rebuild = 1;
        return 0;
    }
    if (have_timestep) s.current_step = timestep;
    else s.current_step += 1;
    const int since = s.current_step - s.last_build_step;
    if (since < s.delay) return 0;
    if (since % s.every != 0) return 0;
    if (!s.check) {
        *rebuild = 1;
        return 0;
    }
    if (s.n == 0) return 0;
    int moved = 0;
    int rc = neigh_check(s, &moved);
    if (rc) return rc;
    *rebuild = moved;
    return 0;
}

// Simulation.reneighbor_if_needed, simulation.py:122-159
int reneighbor(System& s, bool force, int* rebuilt) {
    *rebuilt = 0;
    int dec = 0;
    int rc = neigh_decide(s, s.timestep, 1, &dec);
    if (rc) return rc;
    if (!dec && !force) return 0;
    if (s.has_built) {
        rc = history_save(s);
        if (rc) return rc;
    }
    rc = pbc(s);
    if (!rc) rc = neigh_build(s, false, true);
    if (!rc) rc = history_restore(s);
    if (rc) return rc;
    *rebuilt = 1;
    return 0;
}

int reneighbor_now(System& s) {
    s.current_step = s.timestep;
    int rc = 0;
    if (s.has_built) rc = history_save(s);
    if (!rc) rc = pbc(s);
    if (!rc) rc = neigh_build(s, false, true);
    if (!rc) rc = history_restore(s);
    return rc;
}

int clear_forces(System& s) { return split(s) ? launch_clear_split(s) : launch_clear_f64(s); }

int pair_compute(System& s, double dt, int shearupdate, int accumulate) {
    if (!s.have_pair) return tmps_fail(TMPS_ERR_ARG, "no pair style set");
    if (s.n <= 0) return 0;
    int rc = ensure_split(s);
    if (rc) return rc;
    return split(s) ? launch_pair_split(s, dt, shearupdate, accumulate)
                    : launch_pair_f64(s, dt, shearupdate, accumulate);
}

int cundall(System& s, double gl, double ga) {
    return split(s) ? launch_cundall_split(s, gl, ga) : launch_cundall_f64(s, gl, ga);
}

// FixDeformPressure._strain_h_rate, deform_pressure.py:447-472 (x/y/z)
void strain_h_rate(const tmps_servo_params& S, const DevServo& SV, double dt, double h[6]) {
    for (int k = 0; k < 6; ++k) h[k] = 0.0;
    const double delt = SV.nsteps_total * dt;
    for (int i = 0; i < 3; ++i) {
        const double length = SV.hi_start[i] - SV.lo_start[i];
        const int st = S.style[i];
        if (st == 1 || st == 2) h[i] = S.rate[i] * length;
        else if (st == 3) h[i] = S.vel[i];
        else if (st == 4) h[i] = 2.0 * M_PI / S.tper[i] * S.amp[i];
        else if ((st == 5 || st == 6 || st == 7) && delt > 0.0) {
            double stop;
            if (st == 5) stop = S.fhi[i] - S.flo[i];
            else if (st == 6) stop = length + S.dhi[i] - S.dlo[i];
            else stop = S.scale[i] * length;
            h[i] = (stop - length) / delt;
        }
    }
}

int servo_end_of_step_launch(System& s, double dt) {
    if (!s.have_servo || !s.servo.enabled) return 0;
    int rc = ensure_split(s);
    if (rc) return rc;
    if (split(s)) rc = launch_leaves_split(s, 1);
    else rc = launch_leaves_f64(s, 1);
    if (rc) return rc;
    rc = launch_step_end(s, dt, nblocks_for(std::max(s.n, 1)), 0, 1);
    if (rc) return rc;
    if (!split(s)) {
        rc = launch_remap_apply(s);
        if (rc) return rc;
    }
    return launch_remap_clear(s);
}

int servo_end_of_step(System& s, double dt, double* tensor_out) {
    if (!s.have_servo || !s.servo.enabled) return 0;
    int rc = servo_end_of_step_launch(s, dt);
    if (rc) return rc;
    rc = ctl_pull(s);
    if (rc) return rc;
    s.servo_step_count = s.hctl->servo.step_count;
    s.servo_nsteps = s.hctl->servo.nsteps;
    if (tensor_out)
        for (int k = 0; k < 6; ++k) tensor_out[k] = s.hctl->tensor[k];
    if (s.hctl->error) {
        const int e = s.hctl->error;
        s.hctl->error = 0;
        s.hctl->abort = 0;
        rc = ctl_push(s);
        if (rc) return rc;
        char buf[256];
        const DevBox& B = s.hctl->box;
        snprintf(buf, sizeof buf,
                 "fix deform/pressure drove a box dimension to non-positive (box [%g %g %g]); "
                 "check the gain, max/rate or strain rate",
                 B.prd[0], B.prd[1], B.prd[2]);
        return tmps_fail(e, buf);
    }
    return 0;
}

// Verlet::setup as Simulation.init_simulation does it, simulation.py:170-203
// A fused run's walls that come before (pass 0) or after (pass 1)
// damping/cundall, as post_force calls them in setup (history_update off).
static int setup_walls(System& s, int pass, double dt) {
    for (int q = 0; q < s.nrun_walls; ++q) {
        const WallSpec& w = s.run_spec[q];
        if ((w.before_cundall ? 0 : 1) != pass) continue;
        int rc = split_positions_to_xr(s);
        if (!rc) rc = launch_wall_gran(s, s.run_wh[q][0], w.p, dt, w.coord, w.vwall_n, w.vwall_s, 0);
        if (rc) return rc;
    }
    return 0;
}

int init_simulation(System& s, double dt, int use_cundall, double gl, double ga) {
    s.setup_done = 1;
    int rc = clear_forces(s);
    if (rc) return rc;
    int rebuilt = 0;
    rc = reneighbor(s, true, &rebuilt);
    if (rc) return rc;
    s.rebuilds += rebuilt;
    rc = pair_compute(s, dt, 0, 1);
    if (rc) return rc;
    rc = setup_walls(s, 0, dt);
    if (rc) return rc;
    if (use_cundall) rc = cundall(s, gl, ga);
    if (rc) return rc;
    return setup_walls(s, 1, dt);
}

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
    return h;
}

uint64_t dbits(double d) {
    uint64_t u;
    std::memcpy(&u, &d, sizeof u);
    return u;
}

// One fused step on v/omega buffer `buf`: the force pass writes the other
// buffer. Returns the buffer that holds v/omega afterwards.
// One fused step of a segment: it reads buffer set `buf`, writes 1 - buf.
// The step pass writes the block partials; k_step_close reduces them, runs the
// servo and advances the step (or records a rebuild request).
int fused_step(System& s, int buf, int fin, double dt, int use_cundall, double gl, double ga,
               int servo_on) {
    const int rc = split(s) ? launch_fused_split(s, buf, dt, fin, use_cundall, gl, ga)
                            : launch_fused_f64(s, buf, dt, fin, use_cundall, gl, ga);
    return rc ? rc : launch_step_close(s, dt, servo_on);
}

// The end of a segment whose last step wrote buffer set `buf`: the results
// back into set 0, the last step's remap applied to the stored positions.
// Both keep their hands off an aborted segment: its redo needs set p and the
// pending remap record as the aborted step found them.
int segment_tail(System& s, int buf) {
    int rc = 0;
    if (buf == 1) rc = launch_copy_state(s, 1);
    if (rc) return rc;
    if (!split(s)) {
        rc = launch_remap_apply_fused(s);
        if (rc) return rc;
    }
    return launch_remap_clear(s, 1);
}

// Enqueue a segment of `len` steps (no synchronisation). `kick`: the segment
// opens a run (or follows a rebuild) and applies the first step's half-kick
// from the stored forces. `last`: it closes the run, so its last step stores
// f/torque/virial instead of the next half-kick and the tail brings the
// results to buffer set 0. Inner segments have even length and need neither.
int enqueue_steps(System& s, int len, int kick, int last, double dt, int use_cundall, double gl,
                  double ga) {
    const int servo_on = (s.have_servo && s.servo.enabled) ? 1 : 0;
    int rc = 0;
    if (kick) rc = split(s) ? launch_kick_split(s, dt) : launch_kick_f64(s, dt);
    for (int k = 0; k < len && !rc; ++k)
        rc = fused_step(s, k & 1, last && k == len - 1, dt, use_cundall, gl, ga, servo_on);
    if (rc || !last) return rc;
    return segment_tail(s, len & 1);
}

// Same as enqueue_steps, timed with HIP events on the system's stream: class 0
// the kick, class 1 each step pass kernel from its own start to its own stop
// (hipExtLaunchKernel events), class 3 each step's close kernel, class 2 the
// tail. A segment cut short by a rebuild request is not counted (its later
// kernels exit at once).
int profiled_steps(System& s, int len, int kick, int last, double dt, int use_cundall, double gl,
                   double ga) {
    const int servo_on = (s.have_servo && s.servo.enabled) ? 1 : 0;
    std::vector<hipEvent_t> ev(4 + 3 * (size_t)std::max(len, 0), nullptr);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    hipEvent_t* se = ev.data() + 4;  // per step: pass start, pass stop, close stop
    HIP_TRY(hipEventRecord(ev[0], s.stream));
    int rc = 0;
    if (kick) rc = split(s) ? launch_kick_split(s, dt) : launch_kick_f64(s, dt);
    if (!rc) rc = hipEventRecord(ev[1], s.stream) == hipSuccess ? 0 : TMPS_ERR_HIP;
    for (int k = 0; k < len && !rc; ++k) {
        const int fin = last && k == len - 1;
        s.kstart = se[3 * k];
        s.kstop = se[3 * k + 1];
        rc = split(s) ? launch_fused_split(s, k & 1, dt, fin, use_cundall, gl, ga)
                      : launch_fused_f64(s, k & 1, dt, fin, use_cundall, gl, ga);
        s.kstart = s.kstop = nullptr;
        if (!rc) rc = launch_step_close(s, dt, servo_on);
        if (!rc) rc = hipEventRecord(se[3 * k + 2], s.stream) == hipSuccess ? 0 : TMPS_ERR_HIP;
    }
    if (!rc) rc = hipEventRecord(ev[2], s.stream) == hipSuccess ? 0 : TMPS_ERR_HIP;
    if (!rc && last) rc = segment_tail(s, len & 1);
    if (!rc) rc = hipEventRecord(ev[3], s.stream) == hipSuccess ? 0 : TMPS_ERR_HIP;
    HIP_TRY(hipStreamSynchronize(s.stream));
    if (!rc) {
        float ms = 0.f;
        if (kick) {
            HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
            s.prof_ms[0] += ms;
            s.prof_cnt[0] += 1;
        }
        int aborted = 0;
        HIP_TRY(copy_sync(s.stream, &aborted, &s.ctl->abort, sizeof(int), hipMemcpyDeviceToHost));
        if (len > 0 && !aborted) {
            for (int k = 0; k < len; ++k) {
                HIP_TRY(hipEventElapsedTime(&ms, se[3 * k], se[3 * k + 1]));
                s.prof_ms[1] += ms;
                HIP_TRY(hipEventElapsedTime(&ms, se[3 * k + 1], se[3 * k + 2]));
                s.prof_ms[3] += ms;
            }
            s.prof_cnt[1] += len;
            s.prof_cnt[3] += len;
        }
        if (last) {
            HIP_TRY(hipEventElapsedTime(&ms, ev[2], ev[3]));
            s.prof_ms[2] += ms;
            s.prof_cnt[2] += 1;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}

// The graph slot holding the work `enqueue` puts on the system's stream for
// `key`, captured now if no slot holds it.
template <class Enqueue>
int graph_slot(System& s, uint64_t key, Enqueue&& enqueue, int* slot_out) {
    int slot = -1;
    for (int k = 0; k < System::kGraphSlots; ++k)
        if (s.seg_exec[k] && s.seg_key[k] == key) slot = k;
    if (slot < 0) {
        slot = s.seg_next;
        s.seg_next = (s.seg_next + 1) % System::kGraphSlots;
        if (s.seg_exec[slot]) {
            // the slot's graph may still be in flight
            HIP_TRY(hipStreamSynchronize(s.stream));
            (void)hipGraphExecDestroy(s.seg_exec[slot]);
            s.seg_exec[slot] = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s.stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return tmps_fail_hip(e, "hipStreamEndCapture", __LINE__);
        e = hipGraphInstantiate(&s.seg_exec[slot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return tmps_fail_hip(e, "hipGraphInstantiate", __LINE__);
        s.seg_key[slot] = key;
    }
    *slot_out = slot;
    return 0;
}

// The graph slot holding the captured segment for these settings, captured
// now if none does.
int segment_graph(System& s, int len, int kick, int last, double dt, int use_cundall, double gl,
                  double ga, int* slot_out) {
    uint64_t key = 1469598103934665603ULL;
    key = mix(key, (uint64_t)len);
    key = mix(key, (uint64_t)(kick ? 1 : 0) | (uint64_t)(last ? 2 : 0));
    key = mix(key, dbits(dt));
    key = mix(key, (uint64_t)use_cundall);
    key = mix(key, dbits(gl));
    key = mix(key, dbits(ga));
    key = mix(key, (uint64_t)s.pair.style);
    key = mix(key, (uint64_t)s.n);
    key = mix(key, (uint64_t)(s.have_servo && s.servo.enabled));
    key = mix(key, (uint64_t)(uintptr_t)s.nb_j);
    key = mix(key, (uint64_t)(uintptr_t)s.nb_hist);
    key = mix(key, (uint64_t)s.nrun_walls);
    for (int q = 0; q < s.nrun_walls; ++q)
        key = mix(key, (uint64_t)(uintptr_t)s.run_wh[q][0]);
    if (s.nrun_walls > 0) key = mix(key, (uint64_t)(uintptr_t)s.run_wnz);
    return graph_slot(
        s, key, [&] { return enqueue_steps(s, len, kick, last, dt, use_cundall, gl, ga); },
        slot_out);
}


int launch_segment(System& s, int len, int kick, int last, double dt, int use_cundall, double gl,
                   double ga) {
    if (s.profiling) return profiled_steps(s, len, kick, last, dt, use_cundall, gl, ga);
    if (s.launch_mode == TMPS_LAUNCH_DIRECT)
        return enqueue_steps(s, len, kick, last, dt, use_cundall, gl, ga);
    int slot = -1;
    int rc = segment_graph(s, len, kick, last, dt, use_cundall, gl, ga, &slot);
    if (rc) return rc;
    HIP_TRY(hipGraphLaunch(s.seg_exec[slot], s.stream));
    return 0;
}

// The segment length policy: a power of two from 4 to TMPS_SEG_LEN, at most
// a quarter of the typical number of steps between list rebuilds (an
// exponential mean of the gaps, raised again by a long stretch without one).
void seg_policy(System& s) {
    int m = 4;
    while (m < TMPS_SEG_LEN && 2.0 * m <= 0.5 * s.rebuild_gap) m *= 2;
    s.seg_max = m;
}

void note_rebuild(System& s, int step) {
    if (s.last_rebuild_step >= 0 && step > s.last_rebuild_step)
        s.rebuild_gap = 0.5 * s.rebuild_gap + 0.5 * (double)(step - s.last_rebuild_step);
    s.last_rebuild_step = step;
    seg_policy(s);
}

void note_quiet(System& s, int step) {
    if (s.last_rebuild_step >= 0 && step - s.last_rebuild_step > s.rebuild_gap) {
        s.rebuild_gap = (double)(step - s.last_rebuild_step);
        seg_policy(s);
    }
}

// Segment lengths of a run of nsteps from step t (no rebuild requests): up
// to seg_max steps, even except the last, which ends the run.
// With rare rebuilds, a rest of at most twice TMPS_SEG_LEN steps goes in
// one segment.
int next_segment_len(int t, int t_end, int seg_max, int* last) {
    int len = std::min(seg_max, t_end - t);
    if (seg_max == TMPS_SEG_LEN && t_end - t <= 2 * TMPS_SEG_LEN) len = t_end - t;
    *last = t + len == t_end;
    if (!*last && (len & 1)) --len;
    return len;
}

int check_run_error(System& s) {
    if (!s.hctl->error) return 0;
    const int e = s.hctl->error;
    s.hctl->error = 0;
    s.hctl->abort = 0;
    int rc = ctl_push(s);
    if (rc) return rc;
    // a list build decided on the device records its errors here too
    if (e == TMPS_ERR_OVERFLOW) return list_overflow_error(s);
    if (e == TMPS_ERR_UNSUPPORTED)
        return tmps_fail(e, "FP32-split encoding holds at most 2047 x 2047 x 1023 bins");
    return tmps_fail(e,
                     "fix deform/pressure drove a box dimension to non-positive; check the gain, "
                     "max/rate or strain rate");
}

// The redo of a step whose skin check asked for a rebuild (after the
// control block was pulled and the step's abort seen). Its pass wrote only
// buffer set 1 - par; set par still holds the previous step (positions
// before its remap, half-kicked v, omega, touch flags, history). That set is
// brought to buffer 0, the step's start redone there with the stage kernel
// (remap, drift), the list rebuilt, then the step's force pass,
// final_integrate and end run. On return the step is complete with its
// results in buffer 0 and f / torque stored for the next step's half-kick.
int redo_step(System& s, int par, double dt, int use_cundall, double gl, double ga,
              bool pull_after) {
    const int stp = s.hctl->timestep + 1;
    int rc = 0;
    if (par == 1) {
        rc = launch_copy_state(s, 0);
        if (rc) return rc;
    }
    // The caller pulled the control block after the abort and nothing below
    // changes the box before the build, so the host mirror serves the PBC
    // wrap and the grid; the abort is cleared on the device.
    s.hctl->abort = 0;
    s.hctl->trigger = -1;
    rc = launch_ctl_resume(s);
    if (rc) return rc;
    rc = split(s) ? launch_step_begin_split(s, dt, 0, 0) : launch_step_begin_f64(s, dt, 0, 0);
    if (rc) return rc;
    s.current_step = stp;
    rc = history_save(s);
    if (!rc) rc = pbc(s, true);
    if (rc) return rc;
    return redo_tail(s, stp, dt, use_cundall, gl, ga, pull_after);
}

// redo_step from the list build on: positions wrapped in buffer set 0 (and,
// FP32-split, decoded into xr), the old list saved, the host mirror current.
int redo_tail(System& s, int stp, double dt, int use_cundall, double gl, double ga,
              bool pull_after) {
    const int nparts = nblocks_for(std::max(s.n, 1));
    const int servo_on = (s.have_servo && s.servo.enabled) ? 1 : 0;
    int rc = neigh_build(s, true, true);
    if (!rc) rc = history_restore(s);
    if (!rc) rc = ensure_split(s);
    if (rc) return rc;
    s.rebuilds += 1;
    rc = launch_ctl_rebuilt(s, stp);
    if (rc) return rc;
    rc = split(s) ? launch_step_force_split(s, dt, 1, use_cundall, gl, ga, 0, 1)
                  : launch_step_force_f64(s, dt, 1, use_cundall, gl, ga, 0, 1);
    if (!rc) rc = launch_step_end(s, dt, nparts, 1, servo_on);
    // the force pass left x, flags and history in set 0 and v, omega in set 1
    if (!rc) rc = split(s) ? launch_copy_vw_split(s, 1) : launch_copy_vw(s, 1);
    if (!rc && !split(s)) rc = launch_remap_apply_fused(s);
    if (!rc) rc = launch_remap_clear(s, 1);
    if (rc) return rc;
    if (!pull_after) {
        // the caller learns of an error from the next status it reads; the
        // step completes as step stp otherwise
        s.hctl->timestep = stp;
        s.timestep = stp;
        s.current_step = stp;
        return 0;
    }
    rc = ctl_pull(s);
    if (rc) return rc;
    rc = check_run_error(s);
    if (rc) return rc;
    s.timestep = s.hctl->timestep;
    s.current_step = s.timestep;
    return 0;
}

// What a run's rebuild redo decided on the device needs of the host, set up
// before the run's first segment: bin arrays for the largest grid the list
// build may choose, the row scratch, and a saved-list copy as large as the
// list's buffers.
int dev_redo_prepare(System& s, DevRedo& R) {
    R.ok = false;
    if (s.n <= 0 || (split(s) && s.hctl->box.vremap)) return 0;
    R.max_cutoff = 2.0 * max_radius(s);
    R.max_cells = std::max(s.max_grid_cells, (int)std::min<long long>(4LL * s.n, 1LL << 26));
    int rc = ensure_cells(s, R.max_cells);
    if (!rc) rc = list_reserve_rows(s);
    if (!rc) rc = bins_reserve(s);
    if (!rc && s.nb_cap > 0 && (s.onb_j == nullptr || s.onb_cap < s.nb_cap)) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        dfree(s.onb_j);
        dfree(s.onb_flag);
        if (s.onb_hist) (void)hipFree(s.onb_hist);
        s.onb_hist = nullptr;
        const size_t hsz = split(s) ? sizeof(float4) : sizeof(double4);
        rc = dalloc(s.stream, &s.onb_j, s.nb_cap);
        if (!rc) rc = dalloc(s.stream, &s.onb_flag, s.nb_cap);
        if (rc) return rc;
        HIP_TRY(hipMalloc(&s.onb_hist, s.nb_cap * hsz));
        s.onb_cap = s.nb_cap;
    }
    if (rc) return rc;
    R.ok = s.nb_cap > 0 && s.onb_cap >= s.nb_cap && s.cell_cap >= R.max_cells + 1;
    return 0;
}

// redo_step with the list build decided on the device: everything is
// enqueued behind the aborted segment (and the one queued after it, whose
// kernels see the abort and return), with no host round trip. The host's
// record of the build is written now; should the device find the build
// beyond its buffers, the run's next status says so and dev_redo_recover
// takes the record back and redoes the build on the host.
// The kernels of a device-decided redo (redo_dev) on the system's stream,
// with the host's flags as they stand (which it updates as it goes). A
// negative step is the step being redone, read on the device: the chain
// then holds no per-rebuild value and is captured once and replayed.
static int redo_dev_chain(System& s, const DevRedo& R, int par, int step, bool last, double dt,
                          int use_cundall, double gl, double ga) {
    const int nparts = nblocks_for(std::max(s.n, 1));
    const int servo_on = (s.have_servo && s.servo.enabled) ? 1 : 0;
    int rc = 0;
    if (par == 1) {
        rc = launch_copy_state(s, 0);
        if (rc) return rc;
    }
    rc = launch_ctl_resume(s);
    if (rc) return rc;
    // the step's start and the pbc of the build in one pass over the atoms
    // (on an FP32-split system whose FP64 positions are stale, with the
    // decode between them); the history save between them touches only the
    // list
    const bool fused_start = !split(s) || !s.x64_valid;
    if (fused_start)
        rc = split(s) ? launch_redo_begin_split(s, dt) : launch_redo_begin_f64(s, dt);
    else
        rc = launch_step_begin_split(s, dt, 0, 0);
    if (rc) return rc;
    // history_save: the list (its slot count read on the device) into the
    // copy, which is as large as the list's buffers
    if (s.seed_pending) {
        s.seed_pending = false;
    } else {
        rc = launch_list_save(s);
        if (rc) return rc;
        s.onb_n = s.n;
        s.have_old = true;
    }
    s.x64_valid = true;
    if (!fused_start) rc = launch_pbc(s);
    // (an FP32-split build's x0 is stored by the encoding pass)
    if (!rc) rc = launch_build_dev(s, R.max_cutoff, R.max_cells, !split(s));
    if (!rc) rc = launch_ctl_built_dev(s, step);
    if (!rc && split(s)) rc = launch_x64_to_split_dev(s);
    if (!rc) rc = launch_hist_restore(s);
    if (rc) return rc;
    s.x64_valid = false;
    s.have_old = false;
    s.seed_pending = false;
    rc = launch_ctl_rebuilt(s, step);
    if (rc) return rc;
    // a run's last step stores f, torque and virial; a step inside it also
    // applies the next step's half-kick (the kick kernel's arithmetic), so
    // the next segment starts without one
    const int final_mode = last ? 1 : 2;
    rc = split(s) ? launch_step_force_split(s, dt, final_mode, use_cundall, gl, ga, 0, 1)
                  : launch_step_force_f64(s, dt, final_mode, use_cundall, gl, ga, 0, 1);
    if (!rc) rc = launch_step_end(s, dt, nparts, 1, servo_on);
    if (!rc) rc = split(s) ? launch_copy_vw_split(s, 1) : launch_copy_vw(s, 1);
    if (!rc && !split(s)) rc = launch_remap_apply_fused(s);
    if (!rc) rc = launch_remap_clear(s, 1);
    return rc;
}

// What a captured redo chain depends on: the host flags that choose its
// kernels, the run's settings, launch shapes, and every buffer that can be
// reallocated between rebuilds (a changed one gives a new capture).
static uint64_t redo_graph_key(const System& s, const DevRedo& R, int par, bool last, double dt,
                               int use_cundall, double gl, double ga) {
    uint64_t key = mix(1469598103934665603ULL, 0x7265646f63686e00ULL);  // not a segment
    key = mix(key, (uint64_t)par | (uint64_t)(s.seed_pending ? 2 : 0) |
                       (uint64_t)(s.x64_valid ? 4 : 0) | (uint64_t)(s.have_old ? 8 : 0) |
                       (uint64_t)(last ? 16 : 0));
    for (double d : {dt, gl, ga, R.max_cutoff, s.skin}) key = mix(key, dbits(d));
    for (long long v : {(long long)use_cundall, (long long)s.pair.style, (long long)s.n,
                        (long long)s.precision, (long long)(s.have_servo && s.servo.enabled),
                        (long long)R.max_cells, (long long)s.max_nb, (long long)s.nb_cap,
                        (long long)s.onb_cap, (long long)s.cell_cap, (long long)s.scan_tmp_cap,
                        (long long)s.rs_na, (long long)s.rs_nh, (long long)s.nrun_walls})
        key = mix(key, (uint64_t)v);
    for (const void* p : {(const void*)s.nb_start, (const void*)s.nb_j, (const void*)s.nb_flag,
                          (const void*)s.nb_hist, (const void*)s.nb_flag2, (const void*)s.nb_hist2,
                          (const void*)s.nb_code, (const void*)s.onb_start, (const void*)s.onb_j,
                          (const void*)s.onb_flag, (const void*)s.onb_hist, (const void*)s.cell_count,
                          (const void*)s.cell_start, (const void*)s.cell_fill,
                          (const void*)s.cell_of, (const void*)s.cell_particles,
                          (const void*)s.bin_x, (const void*)s.row_tmp, (const void*)s.scan_tmp,
                          (const void*)s.rs_k1, (const void*)s.rs_k2, (const void*)s.rs_v1,
                          (const void*)s.rs_v2, (const void*)s.rs_hist, (const void*)s.rs_offs,
                          (const void*)s.run_wh[0][0], (const void*)s.run_wh[1][0],
                          (const void*)s.run_wnz, (const void*)s.d_run_spec})
        key = mix(key, (uint64_t)(uintptr_t)p);
    return key;
}

int redo_dev(System& s, DevRedo& R, int par, int stp, bool last, double dt, int use_cundall,
             double gl, double ga) {
    R.stp = stp;
    R.build_count = s.build_count;
    R.last_build_step = s.last_build_step;
    R.has_built = s.has_built;
    R.seed_pending = s.seed_pending;
    R.have_old = s.have_old;
    const bool save = !s.seed_pending;
    int rc = 0;
    if (s.launch_mode == TMPS_LAUNCH_DIRECT || s.profiling) {
        rc = redo_dev_chain(s, R, par, stp, last, dt, use_cundall, gl, ga);
    } else {
        // one graph launch instead of some forty kernel launches, whose host
        // cost would leave the device idle between the chain's small kernels
        const uint64_t key = redo_graph_key(s, R, par, last, dt, use_cundall, gl, ga);
        int slot = -1;
        rc = graph_slot(
            s, key, [&] { return redo_dev_chain(s, R, par, -1, last, dt, use_cundall, gl, ga); },
            &slot);
        if (!rc) HIP_TRY(hipGraphLaunch(s.seg_exec[slot], s.stream));
    }
    if (rc) return rc;
    // the host flags as the chain leaves them (a replayed graph ran none of
    // its host side)
    if (save) s.onb_n = s.n;
    s.seed_pending = false;
    s.x64_valid = false;
    s.have_old = false;
    s.current_step = stp;
    // neigh_record's bookkeeping (the slot count and grid come back with the
    // run's last control block)
    s.has_built = 1;
    s.build_count += 1;
    s.last_build_step = stp;
    if (s.log_builds) s.build_log.push_back(stp);
    s.rebuilds += 1;
    s.timestep = stp;
    R.built = true;
    R.pending = true;
    return 0;
}

// The device found the redo's list build beyond its buffers (TMPS_NEED_HOST
// in the control block, pulled): the host record goes back to the redo's
// start, and the build and the rest of the step are redone on the host from
// the wrapped positions and the saved list the device left.
int dev_redo_recover(System& s, DevRedo& R, double dt, int use_cundall, double gl, double ga) {
    s.build_count = R.build_count;
    s.last_build_step = R.last_build_step;
    s.has_built = R.has_built;
    if (s.log_builds && !s.build_log.empty() && s.build_log.back() == R.stp) s.build_log.pop_back();
    s.rebuilds -= 1;
    s.timestep = R.stp - 1;
    s.current_step = R.stp;
    s.x64_valid = true;
    s.have_old = R.seed_pending ? R.have_old : true;
    s.seed_pending = false;
    s.nslots = s.hctl->nslots;
    Ctl& H = *s.hctl;
    H.error = 0;
    H.abort = 0;
    H.trigger = -1;
    int rc = ctl_push(s);
    if (rc) return rc;
    R.pending = false;
    return redo_tail(s, R.stp, dt, use_cundall, gl, ga, true);
}

// After a run whose rebuilds were decided on the device (host mirror
// current): the host's slot count and grids from the last build.
void dev_redo_finish(System& s, DevRedo& R) {
    if (!R.built) return;
    const Ctl& H = *s.hctl;
    s.nslots = H.nslots;
    s.npairs = H.nslots / 2;
    s.grid = H.lgrid;
    DevBox B = H.box;
    for (int k = 0; k < 3; ++k) {
        B.lo[k] = H.lo_build[k];
        B.prd[k] = H.prd_build[k];
        B.hi[k] = B.lo[k] + B.prd[k];
    }
    const GridP keep = s.grid;
    setup_grid(s, B, R.max_cutoff);
    s.ref_grid = s.grid;
    s.grid = keep;
    R.built = false;
}

// ------------------------------------------------ batched redo (batch.hip)

// The list's buffers in a redo entry (a build can move them).
void redo_entry_lists(const System& s, RedoEntry& e) {
    e.nb_start = s.nb_start;
    e.nb_j = s.nb_j;
    e.nb_flag = s.nb_flag;
    e.nb_flag2 = s.nb_flag2;
    e.nb_hist = static_cast<float4*>(s.nb_hist);
    e.nb_hist2 = static_cast<float4*>(s.nb_hist2);
    e.nb_code = s.nb_code;
    e.onb_start = s.onb_start;
    e.onb_j = s.onb_j;
    e.onb_flag = s.onb_flag;
    e.onb_hist = static_cast<float4*>(s.onb_hist);
    e.onb_cap = (int)std::min<size_t>(s.onb_cap, (size_t)INT32_MAX);
}

// redo_step (pull_after false) for an FP32-split member, up to the slot
// count of its list build: the same host mirror updates, capacities and
// grid, in the same order; the kernels are the batched ones.
int redo_batch_begin(System& s, RedoEntry& e) {
    if (!split(s) || s.n <= 0) return tmps_fail(TMPS_ERR_ARG, "batched redo of a non-split system");
    Ctl& H = *s.hctl;
    const int stp = H.timestep + 1;
    e = RedoEntry{};
    // the resume (copy of set 1 when H.bpar is 1) and the step's start
    H.abort = 0;
    H.trigger = -1;
    s.current_step = stp;
    // history_save (skipped for a seeded history, as history_save skips it)
    const int old_slots = std::max(s.nslots, 0);
    e.copy_blocks = std::min(nblocks_for(std::max(s.n, old_slots)), 4096);
    int rc = 0;
    if (s.seed_pending) {
        s.seed_pending = false;
        e.save_blocks = 0;
    } else {
        rc = history_reserve(s);
        if (rc) return rc;
        e.save_blocks =
            std::min((std::max(s.n + 1, old_slots) + TMPS_BLOCK - 1) / TMPS_BLOCK, 4096);
        s.onslots = old_slots;
        s.onb_n = s.n;
    }
    s.have_old = true;
    // pbc(s, true)
    if (H.box.vremap)
        return tmps_fail(TMPS_ERR_UNSUPPORTED, "remap v is not supported with FP32-split");
    e.to_x64 = s.x64_valid ? 0 : 1;
    s.x64_valid = true;
    // neigh_build(s, true), up to the list's row counts
    rc = neigh_grid(s);
    if (!rc) rc = list_reserve_rows(s);
    if (rc) return rc;
    e.ctl = s.ctl;
    e.n = s.n;
    e.ncell = s.grid.ncell;
    e.old_n = s.onb_start ? s.n : 0;
    e.max_nb = s.max_nb;
    e.servo_on = (s.have_servo && s.servo.enabled) ? 1 : 0;
    e.grid = s.grid;
    e.skin = s.skin;
    e.pair = s.pair;
    e.partial = s.partial;
    e.xr = s.xr;
    e.vm = s.vm;
    e.x0 = s.x0;
    e.bin_x = s.bin_x;
    e.up = s.up;
    e.up2 = s.up2;
    e.vmf = s.vmf;
    e.vmf2 = s.vmf2;
    e.omf = s.omf;
    e.omf2 = s.omf2;
    e.ff = s.ff;
    e.tqf = s.tqf;
    e.u0 = s.u0;
    e.cell = s.cell;
    e.vir = s.vir;
    e.cell_of = s.cell_of;
    e.cell_count = s.cell_count;
    e.cell_start = s.cell_start;
    e.cell_fill = s.cell_fill;
    e.cell_particles = s.cell_particles;
    e.scan_tmp = s.scan_tmp;
    e.full_count = s.full_count;
    e.half_count = s.half_count;
    e.flags = s.flags;
    e.row_tmp = s.row_tmp;
    redo_entry_lists(s, e);
    return 0;
}

// The rest of redo_step's host work once the list's flags are known.
int redo_batch_list(System& s, int4 flags, RedoEntry& e) {
    if (flags.x) return list_overflow_error(s);
    int rc = ensure_list_capacity(s, (size_t)flags.z);
    if (rc) return rc;
    s.nslots = flags.z;
    int fast[3], gi[3];
    rc = neigh_record(s, flags.z, fast, gi);
    if (rc) return rc;
    s.x64_valid = false;  // the split encoding from the new bins
    s.have_old = false;   // history_restore
    s.seed_pending = false;
    s.rebuilds += 1;
    Ctl& H = *s.hctl;
    const int stp = s.current_step;
    H.timestep = stp;
    s.timestep = stp;
    // the member rejoins the segments
    H.bpar = 0;
    H.bkick = 1;
    H.bdone = stp == H.bend ? 1 : 0;
    e.long_rows = flags.y ? 1 : 0;
    e.codes = (s.nslots > 0 && s.nb_code) ? 1 : 0;
    e.step = stp;
    for (int k = 0; k < 3; ++k) {
        e.fast[k] = fast[k];
        e.gi[k] = gi[k];
    }
    e.done = H.bdone;
    redo_entry_lists(s, e);
    return 0;
}

// Simulation.run(steps) with nve/sphere + pair [+ damping/cundall]
// [+ deform/pressure], device-resident.
int run(System& s, int nsteps, double dt, int run_pre, int use_cundall, double gl, double ga) {
    if (!s.have_pair) return tmps_fail(TMPS_ERR_ARG, "tmps_run needs a pair style");
    if (nsteps < 0) return tmps_fail(TMPS_ERR_ARG, "negative step count");
    s.rebuilds = 0;
    int rc = 0;
    if (run_pre || !s.setup_done) {
        s.hctl_fresh = false;  // Verlet::setup changes the device's control block
        rc = init_simulation(s, dt, use_cundall, gl, ga);
        if (rc) return rc;
    }
    if (nsteps == 0 || s.n == 0) {
        s.timestep += nsteps;
        return 0;
    }
    rc = ensure_split(s);
    if (rc) return rc;
    rc = launch_walls_nz(s);
    if (rc) return rc;
    rc = ctl_edit(s);
    if (rc) return rc;
    Ctl& H = *s.hctl;
    H.timestep = s.timestep;
    H.trigger = -1;
    H.abort = 0;
    H.error = 0;
    H.last_build_step = s.last_build_step;
    H.delay = s.delay;
    H.every = s.every;
    H.check = s.check;
    H.trigger_sq = 0.25 * s.skin * s.skin;
    H.remap.step = -1;
    DevRedo R;
    rc = dev_redo_prepare(s, R);
    if (rc) return rc;
    rc = ctl_commit(s);
    // from here the device's copy runs ahead of the mirror until the end of
    // the run brings it back
    s.hctl_fresh = false;
    if (rc) return rc;
    if (split(s)) {
        rc = launch_split_consts(s, 0);
        if (rc) return rc;
    }
    const int t_end = s.timestep + nsteps;
    // Segments run two deep: while one executes, the host waits on the one
    // before it and reads its status (timestep, trigger, abort, error) from
    // pinned memory, so a rebuild request costs at most one segment of
    // skipped kernels.
    struct Seg {
        int start, len, slot, last;
    };
    Seg inflight[2];
    int nin = 0, next_slot = 0;
    int t = s.timestep;
    int kick = 1;
    bool mirror_back = false;  // the last segment's copy of the control block arrived
    while (true) {
        while (nin < 2 && t < t_end) {
            int last = 0;
            const int len = next_segment_len(t, t_end, s.seg_max, &last);
            rc = launch_segment(s, len, kick, last, dt, use_cundall, gl, ga);
            kick = 0;
            if (rc) return rc;
            const int slot = next_slot;
            next_slot = (next_slot + 1) % System::kGraphSlots;
            HIP_TRY(hipMemcpyAsync(s.hstatus + 4 * slot, &s.ctl->timestep, 4 * sizeof(int),
                                   hipMemcpyDeviceToHost, s.stream));
            // the run's last segment also brings the whole control block
            // back into the (pinned) mirror, so the end of the run needs no
            // further round trip; the host touches the mirror only after
            // this segment's event or a stream synchronisation
            if (last)
                HIP_TRY(hipMemcpyAsync(s.hctl, s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                                       s.stream));
            HIP_TRY(hipEventRecord(s.seg_event[slot], s.stream));
            inflight[nin++] = Seg{t, len, slot, last};
            t += len;
        }
        if (nin == 0) break;
        const Seg sg = inflight[0];
        inflight[0] = inflight[1];
        --nin;
        HIP_TRY(hipEventSynchronize(s.seg_event[sg.slot]));
        const int* st4 = s.hstatus + 4 * sg.slot;
        const int done = st4[0], aborted = st4[2], err = st4[3];
        if (!aborted && !err) {
            s.timestep = done;
            if (s.has_built) s.current_step = s.timestep;
            mirror_back = sg.last != 0;
            R.pending = false;
            note_quiet(s, done);
            continue;
        }
        if (R.ok && aborted && !err && st4[1] == done + 1) {
            // the skin check of step done + 1 asked for a rebuild: the step
            // is redone with the list build decided on the device, queued
            // behind the segment that follows (its kernels see the abort and
            // return), and the run goes on from that step with no wait
            nin = 0;
            s.timestep = done;
            if (s.has_built) s.current_step = done;
            const int stp = done + 1;
            const bool last = stp == t_end;
            rc = redo_dev(s, R, ((stp - 1) - sg.start) % 2, stp, last, dt, use_cundall, gl, ga);
            if (rc) return rc;
            note_rebuild(s, stp);
            t = stp;
            kick = last ? 1 : 0;  // inside the run the redo applied the next half-kick
            mirror_back = false;
            continue;
        }
        // skip whatever is still queued; it sees the abort and does nothing
        HIP_TRY(hipStreamSynchronize(s.stream));
        nin = 0;
        rc = ctl_pull(s);
        if (rc) return rc;
        if (s.hctl->error == TMPS_NEED_HOST && R.pending) {
            rc = dev_redo_recover(s, R, dt, use_cundall, gl, ga);
            if (rc) return rc;
            R.built = false;  // the host built the list last
            rc = dev_redo_prepare(s, R);  // its buffers may have grown
            if (rc) return rc;
            t = s.timestep;
            kick = 1;
            mirror_back = false;
            continue;
        }
        rc = check_run_error(s);
        if (rc) return rc;
        s.timestep = s.hctl->timestep;
        const int stp = s.hctl->timestep + 1;
        rc = redo_step(s, ((stp - 1) - sg.start) % 2, dt, use_cundall, gl, ga);
        if (rc) return rc;
        note_rebuild(s, stp);
        R.built = false;
        rc = dev_redo_prepare(s, R);
        if (rc) return rc;
        t = s.timestep;
        kick = 1;  // the redone step ended with final_integrate
        mirror_back = false;
    }
    if (!mirror_back) {
        rc = ctl_pull(s);
        if (rc) return rc;
    }
    if (s.hctl->error == TMPS_NEED_HOST && R.pending) {
        // the run's last step was a redo the device could not build
        rc = dev_redo_recover(s, R, dt, use_cundall, gl, ga);
        if (rc) return rc;
        R.built = false;
    } else if (s.hctl->error) {
        return check_run_error(s);
    }
    dev_redo_finish(s, R);
    s.hctl_fresh = true;
    s.servo_step_count = s.hctl->servo.step_count;
    s.servo_nsteps = s.hctl->servo.nsteps;
    return 0;
}

// Captures the segment graphs a run of nsteps (started now, without
// rebuild requests) launches, without running it: a run's first use of a
// segment shape otherwise pays the capture and instantiation inside it.
int prepare_run(System& s, int nsteps, double dt, int use_cundall, double gl, double ga) {
    if (!s.have_pair) return tmps_fail(TMPS_ERR_ARG, "tmps_prepare_run needs a pair style");
    if (nsteps < 0) return tmps_fail(TMPS_ERR_ARG, "negative step count");
    if (nsteps == 0 || s.n == 0 || !s.setup_done || s.launch_mode == TMPS_LAUNCH_DIRECT)
        return 0;
    const int t_end = s.timestep + nsteps;
    int t = s.timestep, kick = 1;
    while (t < t_end) {
        int last = 0;
        const int len = next_segment_len(t, t_end, s.seg_max, &last);
        int slot = -1;
        const int rc = segment_graph(s, len, kick, last, dt, use_cundall, gl, ga, &slot);
        if (rc) return rc;
        kick = 0;
        t += len;
    }
    return 0;
}

// --------------------------------------------------------------- field I/O

int set_field(System& s, int field, const void* host, int count) {
    if (count < 0 || count > s.max_atoms) return tmps_fail(TMPS_ERR_CAPACITY, "field count exceeds max_atoms");
    const int n = count;
    if (n == 0) return 0;
    HIP_TRY(hipStreamSynchronize(s.stream));
    const double* hd = static_cast<const double*>(host);
    if (field == TMPS_TAG) {
        HIP_TRY(copy_sync(s.stream, s.tag, host, sizeof(int) * n, hipMemcpyHostToDevice));
        return 0;
    }
    if (field == TMPS_VIRIAL) {
        HIP_TRY(copy_sync(s.stream, s.vir, host, sizeof(double) * 6 * n, hipMemcpyHostToDevice));
        return 0;
    }
    if (field == TMPS_X0) return tmps_fail(TMPS_ERR_ARG, "x0 is read-only");
    if (!split(s)) {
        std::vector<double4> buf(n);
        double4* dst = nullptr;
        switch (field) {
            case TMPS_X: dst = s.xr; break;
            case TMPS_V: dst = s.vm; break;
            case TMPS_OMEGA: dst = s.om; break;
            case TMPS_F: dst = s.f; break;
            case TMPS_TORQUE: dst = s.tq; break;
            case TMPS_RADIUS: dst = s.xr; break;
            case TMPS_RMASS: dst = s.vm; break;
            default: return tmps_fail(TMPS_ERR_ARG, "unknown field");
        }
        HIP_TRY(copy_sync(s.stream, buf.data(), dst, sizeof(double4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            if (field == TMPS_RADIUS) {
                buf[i].w = hd[i];
                s.h_radius[i] = hd[i];
                s.max_r_valid = false;
            } else if (field == TMPS_RMASS) {
                buf[i].w = hd[i];
            } else {
                buf[i].x = hd[3 * i];
                buf[i].y = hd[3 * i + 1];
                buf[i].z = hd[3 * i + 2];
            }
        }
        HIP_TRY(copy_sync(s.stream, dst, buf.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
        return 0;
    }
    // FP32-split
    if (field == TMPS_X) {
        std::vector<double4> buf(n);
        for (int i = 0; i < n; ++i)
            buf[i] = make_double4(hd[3 * i], hd[3 * i + 1], hd[3 * i + 2], s.h_radius[i]);
        if (!s.x64_valid && s.n > n) {
            // keep the atoms beyond `count` where they are
            int rc = launch_split_to_x64(s);
            if (rc) return rc;
            HIP_TRY(hipStreamSynchronize(s.stream));
        }
        HIP_TRY(copy_sync(s.stream, s.xr, buf.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
        s.x64_valid = true;
        return 0;
    }
    std::vector<float4> buf(n);
    float4* dst = nullptr;
    switch (field) {
        case TMPS_V: dst = s.vmf; break;
        case TMPS_OMEGA: dst = s.omf; break;
        case TMPS_F: dst = s.ff; break;
        case TMPS_TORQUE: dst = s.tqf; break;
        case TMPS_RADIUS: dst = s.up; break;
        case TMPS_RMASS: dst = s.vmf; break;
        default: return tmps_fail(TMPS_ERR_ARG, "unknown field");
    }
    HIP_TRY(copy_sync(s.stream, buf.data(), dst, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        if (field == TMPS_RADIUS) {
            buf[i].w = (float)hd[i];
            s.h_radius[i] = hd[i];
            s.max_r_valid = false;
        } else if (field == TMPS_RMASS) {
            buf[i].w = (float)hd[i];
        } else {
            buf[i].x = (float)hd[3 * i];
            buf[i].y = (float)hd[3 * i + 1];
            buf[i].z = (float)hd[3 * i + 2];
        }
    }
    HIP_TRY(copy_sync(s.stream, dst, buf.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    if (field == TMPS_RADIUS) {
        // the FP64 scratch positions carry the radius too
        std::vector<double4> xb(n);
        HIP_TRY(copy_sync(s.stream, xb.data(), s.xr, sizeof(double4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) xb[i].w = hd[i];
        HIP_TRY(copy_sync(s.stream, s.xr, xb.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    }
    if (field == TMPS_RADIUS || field == TMPS_RMASS) {
        // omega.w = 1 / (0.4 m r^2)
        std::vector<float4> om(n), vm(n);
        HIP_TRY(copy_sync(s.stream, om.data(), s.omf, sizeof(float4) * n, hipMemcpyDeviceToHost));
        HIP_TRY(copy_sync(s.stream, vm.data(), s.vmf, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            const double r = s.h_radius[i], m = vm[i].w;
            om[i].w = (m > 0 && r > 0) ? (float)(1.0 / (0.4 * m * r * r)) : 0.f;
        }
        HIP_TRY(copy_sync(s.stream, s.omf, om.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    }
    return 0;
}

int get_field(System& s, int field, void* host, int count) {
    if (count < 0 || count > s.max_atoms) return tmps_fail(TMPS_ERR_CAPACITY, "field count exceeds max_atoms");
    const int n = count;
    if (n == 0) return 0;
    HIP_TRY(hipStreamSynchronize(s.stream));
    double* hd = static_cast<double*>(host);
    if (field == TMPS_TAG) {
        HIP_TRY(copy_sync(s.stream, host, s.tag, sizeof(int) * n, hipMemcpyDeviceToHost));
        return 0;
    }
    if (field == TMPS_VIRIAL) {
        HIP_TRY(copy_sync(s.stream, host, s.vir, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
        return 0;
    }
    if (field == TMPS_RADIUS) {
        for (int i = 0; i < n; ++i) hd[i] = s.h_radius[i];
        return 0;
    }
    if (!split(s) || field == TMPS_X || field == TMPS_X0) {
        double4* src = nullptr;
        if (split(s) && field == TMPS_X) {
            int rc = split_positions_to_xr(s);
            if (rc) return rc;
            HIP_TRY(hipStreamSynchronize(s.stream));
        }
        switch (field) {
            case TMPS_X: src = s.xr; break;
            case TMPS_V: src = s.vm; break;
            case TMPS_OMEGA: src = s.om; break;
            case TMPS_F: src = s.f; break;
            case TMPS_TORQUE: src = s.tq; break;
            case TMPS_RMASS: src = s.vm; break;
            case TMPS_X0: src = s.x0; break;
            default: return tmps_fail(TMPS_ERR_ARG, "unknown field");
        }
        std::vector<double4> buf(n);
        HIP_TRY(copy_sync(s.stream, buf.data(), src, sizeof(double4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            if (field == TMPS_RMASS) {
                hd[i] = buf[i].w;
            } else {
                hd[3 * i] = buf[i].x;
                hd[3 * i + 1] = buf[i].y;
                hd[3 * i + 2] = buf[i].z;
            }
        }
        return 0;
    }
    float4* src = nullptr;
    switch (field) {
        case TMPS_V: src = s.vmf; break;
        case TMPS_OMEGA: src = s.omf; break;
        case TMPS_F: src = s.ff; break;
        case TMPS_TORQUE: src = s.tqf; break;
        case TMPS_RMASS: src = s.vmf; break;
        default: return tmps_fail(TMPS_ERR_ARG, "unknown field");
    }
    std::vector<float4> buf(n);
    HIP_TRY(copy_sync(s.stream, buf.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        if (field == TMPS_RMASS) {
            hd[i] = buf[i].w;
        } else {
            hd[3 * i] = buf[i].x;
            hd[3 * i + 1] = buf[i].y;
            hd[3 * i + 2] = buf[i].z;
        }
    }
    return 0;
}

void free_all(System& s) {
    adjoint_free(s);
    invalidate_graphs(s);
    dfree(s.xr); dfree(s.vm); dfree(s.om); dfree(s.f); dfree(s.tq); dfree(s.x0); dfree(s.vir);
    dfree(s.vm2); dfree(s.om2); dfree(s.vmf2); dfree(s.omf2); dfree(s.xr2); dfree(s.up2);
    dfree(s.tag);
    dfree(s.up); dfree(s.vmf); dfree(s.omf); dfree(s.ff); dfree(s.tqf); dfree(s.u0); dfree(s.cell);
    dfree(s.nb_start); dfree(s.nb_j); dfree(s.nb_flag);
    dfree(s.onb_start); dfree(s.onb_j); dfree(s.onb_flag); dfree(s.nb_flag2); dfree(s.nb_code);
    if (s.nb_hist2) (void)hipFree(s.nb_hist2);
    s.nb_hist2 = nullptr;
    if (s.nb_hist) (void)hipFree(s.nb_hist);
    if (s.onb_hist) (void)hipFree(s.onb_hist);
    s.nb_hist = s.onb_hist = nullptr;
    dfree(s.cell_of); dfree(s.cell_count); dfree(s.cell_start); dfree(s.cell_fill);
    dfree(s.rs_k1); dfree(s.rs_k2); dfree(s.rs_v1); dfree(s.rs_v2); dfree(s.rs_hist);
    dfree(s.d_run_spec);
    dfree(s.run_wnz);
    dfree(s.d_radius);
    dfree(s.rs_offs);
    dfree(s.cell_particles);
    dfree(s.bin_x);
    dfree(s.full_count); dfree(s.half_count); dfree(s.scan_tmp); dfree(s.flags);
    dfree(s.row_tmp);
    dfree(s.partial);
    dfree(s.level);
    dfree(s.cmp_count); dfree(s.cmp_stress); dfree(s.cmp_sum);
    dfree(s.ctl);
    if (s.hctl) (void)hipHostFree(s.hctl);
    s.hctl = nullptr;
    if (s.hstatus) (void)hipHostFree(s.hstatus);
    s.hstatus = nullptr;
    if (s.hsum) (void)hipHostFree(s.hsum);
    s.hsum = nullptr;
    for (int k = 0; k < System::kGraphSlots; ++k) {
        if (s.seg_event[k]) (void)hipEventDestroy(s.seg_event[k]);
        s.seg_event[k] = nullptr;
    }
    if (s.own_stream) (void)hipStreamDestroy(s.own_stream);
    s.own_stream = nullptr;
    s.stream = nullptr;
}

int create(System& s, int max_atoms, int max_nb, int precision) {
    s.max_atoms = max_atoms;
    s.max_nb = max_nb;
    s.precision = precision;
    HIP_TRY(hipGetDevice(&s.device));
    HIP_TRY(hipStreamCreateWithFlags(&s.own_stream, hipStreamNonBlocking));
    s.stream = s.own_stream;
    const size_t na = (size_t)max_atoms;
    int rc = 0;
    rc = rc ? rc : dalloc(s.stream, &s.xr, na);
    rc = rc ? rc : dalloc(s.stream, &s.x0, na);
    rc = rc ? rc : dalloc(s.stream, &s.vir, 6 * na);
    rc = rc ? rc : dalloc(s.stream, &s.tag, na);
    if (precision == TMPS_FP64) {
        rc = rc ? rc : dalloc(s.stream, &s.vm, na);
        rc = rc ? rc : dalloc(s.stream, &s.om, na);
        rc = rc ? rc : dalloc(s.stream, &s.vm2, na);
        rc = rc ? rc : dalloc(s.stream, &s.om2, na);
        rc = rc ? rc : dalloc(s.stream, &s.xr2, na);
        rc = rc ? rc : dalloc(s.stream, &s.f, na);
        rc = rc ? rc : dalloc(s.stream, &s.tq, na);
    } else {
        rc = rc ? rc : dalloc(s.stream, &s.up, na);
        rc = rc ? rc : dalloc(s.stream, &s.vmf, na);
        rc = rc ? rc : dalloc(s.stream, &s.omf, na);
        rc = rc ? rc : dalloc(s.stream, &s.vmf2, na);
        rc = rc ? rc : dalloc(s.stream, &s.omf2, na);
        rc = rc ? rc : dalloc(s.stream, &s.up2, na);
        rc = rc ? rc : dalloc(s.stream, &s.ff, na);
        rc = rc ? rc : dalloc(s.stream, &s.tqf, na);
        rc = rc ? rc : dalloc(s.stream, &s.u0, na);
        rc = rc ? rc : dalloc(s.stream, &s.cell, na);
    }
    rc = rc ? rc : dalloc(s.stream, &s.nb_start, na + 1);
    rc = rc ? rc : dalloc(s.stream, &s.onb_start, na + 1);
    rc = rc ? rc : dalloc(s.stream, &s.cell_of, na);
    rc = rc ? rc : dalloc(s.stream, &s.cell_particles, na);
    rc = rc ? rc : dalloc(s.stream, &s.bin_x, na);
    rc = rc ? rc : dalloc(s.stream, &s.full_count, na + 1);
    rc = rc ? rc : dalloc(s.stream, &s.half_count, na + 1);
    rc = rc ? rc : dalloc(s.stream, &s.flags, 4);
    const int nparts = nblocks_for(std::max(max_atoms, 1)) + 1;
    rc = rc ? rc : dalloc(s.stream, &s.partial, (size_t)nparts * 8);
    s.partial_cap = nparts;
    s.level_cap = nparts / 4096 + 1;
    rc = rc ? rc : dalloc(s.stream, &s.level, (size_t)s.level_cap * 8);
    rc = rc ? rc : dalloc(s.stream, &s.ctl, 1);
    if (rc) return rc;
    HIP_TRY(hipHostMalloc((void**)&s.hctl, sizeof(Ctl), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&s.hsum, sizeof(double) * 8, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&s.hstatus, sizeof(int) * 4 * System::kGraphSlots,
                          hipHostMallocDefault));
    for (int k = 0; k < System::kGraphSlots; ++k)
        HIP_TRY(hipEventCreateWithFlags(&s.seg_event[k], hipEventDisableTiming));
    std::memset(s.hctl, 0, sizeof(Ctl));
    s.hctl->remap.step = -1;
    s.hctl->trigger = -1;
    s.hctl->last_build_step = -1;
    s.hctl->every = 1;
    s.hctl->check = 1;
    for (int k = 0; k < 3; ++k) {
        s.hctl->box.hi[k] = 1.0;
        s.hctl->box.prd[k] = 1.0;
        s.hctl->box.per[k] = 1;
        s.hctl->sc.gi[k] = 1;
        s.hctl->prd_build[k] = 1.0;
    }
    s.hctl->servo.p.nevery = 1;
    rc = ctl_push(s);
    if (rc) return rc;
    rc = ensure_cells(s, 1);
    if (rc) return rc;
    s.h_radius.assign(na, 0.0);
    return 0;
}

}  // namespace tmps_api

using namespace tmps_api;

void drop_graphs(System& s) {
    bool any = false;
    for (int k = 0; k < System::kGraphSlots; ++k) any = any || s.seg_exec[k] != nullptr;
    if (!any) return;
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    for (int k = 0; k < System::kGraphSlots; ++k) {
        if (s.seg_exec[k]) (void)hipGraphExecDestroy(s.seg_exec[k]);
        s.seg_exec[k] = nullptr;
        s.seg_key[k] = 0;
    }
    s.seg_next = 0;
}

// ================================================================== C ABI

extern "C" {

const char* tmps_version(void) { return "taichimps_amd 0.1 (gfx950)"; }

const char* tmps_last_error(void) { return g_err.c_str(); }

int tmps_device_name(char* buf, int len) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

int tmps_create(int max_atoms, int max_neighbors_per_atom, int precision, tmps_system** out) {
    if (!out) return tmps_fail(TMPS_ERR_ARG, "out is null");
    *out = nullptr;
    if (max_atoms <= 0 || max_neighbors_per_atom <= 0)
        return tmps_fail(TMPS_ERR_ARG, "max_atoms and max_neighbors_per_atom must be positive");
    if (precision != TMPS_FP64 && precision != TMPS_FP32_SPLIT)
        return tmps_fail(TMPS_ERR_ARG, "unknown precision");
    System* s = new System();
    int rc = create(*s, max_atoms, max_neighbors_per_atom, precision);
    if (rc) {
        free_all(*s);
        delete s;
        return rc;
    }
    *out = reinterpret_cast<tmps_system*>(s);
    return 0;
}

int tmps_destroy(tmps_system* h) {
    if (!h) return 0;
    System* s = sys(h);
    (void)hipStreamSynchronize(s->stream);
    free_all(*s);
    delete s;
    return 0;
}

int tmps_set_stream(tmps_system* h, void* stream) {
    System& s = *sys_view(h);
    hipStream_t next = stream ? static_cast<hipStream_t>(stream) : s.own_stream;
    if (next == s.stream) return 0;
    // work already queued on the old stream finishes first
    HIP_TRY(hipStreamSynchronize(s.stream));
    s.stream = next;
    return 0;
}

int tmps_get_stream(tmps_system* h, void** stream) {
    if (!stream) return tmps_fail(TMPS_ERR_ARG, "stream is null");
    *stream = static_cast<void*>(sys(h)->stream);
    return 0;
}

int tmps_set_nlocal(tmps_system* h, int nlocal) {
    System& s = *sys(h);
    if (nlocal < 0 || nlocal > s.max_atoms)
        return tmps_fail(TMPS_ERR_CAPACITY, "Exceeded max_atoms capacity");
    if (nlocal != s.n) {
        invalidate_graphs(s);
        s.max_r_valid = false;
    }
    s.n = nlocal;
    return 0;
}

int tmps_set_field(tmps_system* h, int field, const void* host, int count) {
    return set_field(*sys(h), field, host, count);
}

int tmps_get_field(tmps_system* h, int field, void* host, int count) {
    return get_field(*sys(h), field, host, count);
}

int tmps_get_field_device(tmps_system* h, int field, void* dev, int count) {
    System& s = *sys_view(h);
    if (count < 0 || count > s.max_atoms)
        return tmps_fail(TMPS_ERR_CAPACITY, "field count exceeds max_atoms");
    const int n = count;
    if (n == 0) return 0;
    if (!dev) return tmps_fail(TMPS_ERR_ARG, "null device buffer");
    double* out = static_cast<double*>(dev);
    switch (field) {
        case TMPS_TAG:
            HIP_TRY(hipMemcpyAsync(dev, s.tag, sizeof(int) * n, hipMemcpyDeviceToDevice, s.stream));
            return 0;
        case TMPS_VIRIAL:
            HIP_TRY(hipMemcpyAsync(dev, s.vir, sizeof(double) * 6 * n, hipMemcpyDeviceToDevice,
                                   s.stream));
            return 0;
        case TMPS_RADIUS:
            // the exact radii are the host's (an FP32-split system's records
            // carry them rounded)
            HIP_TRY(hipMemcpyAsync(dev, s.h_radius.data(), sizeof(double) * n,
                                   hipMemcpyHostToDevice, s.stream));
            return 0;
        case TMPS_X0:
            return launch_field_pack(s, s.x0, nullptr, n, 0, out);
        case TMPS_X:
            if (split(s)) {
                const int rc = split_positions_to_xr(s);
                if (rc) return rc;
            }
            return launch_field_pack(s, s.xr, nullptr, n, 0, out);
        default:
            break;
    }
    const bool sp = split(s);
    switch (field) {
        case TMPS_V: return launch_field_pack(s, sp ? nullptr : s.vm, s.vmf, n, 0, out);
        case TMPS_OMEGA: return launch_field_pack(s, sp ? nullptr : s.om, s.omf, n, 0, out);
        case TMPS_F: return launch_field_pack(s, sp ? nullptr : s.f, s.ff, n, 0, out);
        case TMPS_TORQUE: return launch_field_pack(s, sp ? nullptr : s.tq, s.tqf, n, 0, out);
        case TMPS_RMASS: return launch_field_pack(s, sp ? nullptr : s.vm, s.vmf, n, 1, out);
        default: return tmps_fail(TMPS_ERR_ARG, "unknown field");
    }
}

int tmps_set_field_device(tmps_system* h, int field, const void* dev, int count) {
    System& s = *sys(h);
    if (count < 0 || count > s.max_atoms)
        return tmps_fail(TMPS_ERR_CAPACITY, "field count exceeds max_atoms");
    const int n = count;
    if (n == 0) return 0;
    if (!dev) return tmps_fail(TMPS_ERR_ARG, "null device buffer");
    const double* in = static_cast<const double*>(dev);
    switch (field) {
        case TMPS_TAG:
            HIP_TRY(hipMemcpyAsync(s.tag, dev, sizeof(int) * n, hipMemcpyDeviceToDevice, s.stream));
            return 0;
        case TMPS_VIRIAL:
            HIP_TRY(hipMemcpyAsync(s.vir, dev, sizeof(double) * 6 * n, hipMemcpyDeviceToDevice,
                                   s.stream));
            return 0;
        case TMPS_X0:
            return tmps_fail(TMPS_ERR_ARG, "x0 is read-only");
        case TMPS_RADIUS:
        case TMPS_RMASS: {
            // these also set host state (the radius mirror the neighbor grid
            // reads, an FP32-split system's inverse inertia): through the
            // host setter
            std::vector<double> hv(n);
            HIP_TRY(copy_sync(s.stream, hv.data(), dev, sizeof(double) * n,
                              hipMemcpyDeviceToHost));
            return set_field(s, field, hv.data(), n);
        }
        default:
            break;
    }
    if (!split(s)) {
        switch (field) {
            case TMPS_X: return launch_field_unpack(s, s.xr, nullptr, n, in, nullptr);
            case TMPS_V: return launch_field_unpack(s, s.vm, nullptr, n, in, nullptr);
            case TMPS_OMEGA: return launch_field_unpack(s, s.om, nullptr, n, in, nullptr);
            case TMPS_F: return launch_field_unpack(s, s.f, nullptr, n, in, nullptr);
            case TMPS_TORQUE: return launch_field_unpack(s, s.tq, nullptr, n, in, nullptr);
            default: return tmps_fail(TMPS_ERR_ARG, "unknown field");
        }
    }
    if (field == TMPS_X) {
        // as set_field: the FP64 positions become the authority, each with
        // its exact radius, atoms beyond `count` decoded where they are
        if (!s.x64_valid && s.n > n) {
            const int rc = launch_split_to_x64(s);
            if (rc) return rc;
        }
        if (!s.d_radius) {
            int rc = dalloc(s.stream, &s.d_radius, (size_t)std::max(s.max_atoms, 1));
            if (rc) return rc;
        }
        HIP_TRY(hipMemcpyAsync(s.d_radius, s.h_radius.data(), sizeof(double) * n,
                               hipMemcpyHostToDevice, s.stream));
        const int rc = launch_field_unpack(s, s.xr, nullptr, n, in, s.d_radius);
        if (rc) return rc;
        s.x64_valid = true;
        return 0;
    }
    switch (field) {
        case TMPS_V: return launch_field_unpack(s, nullptr, s.vmf, n, in, nullptr);
        case TMPS_OMEGA: return launch_field_unpack(s, nullptr, s.omf, n, in, nullptr);
        case TMPS_F: return launch_field_unpack(s, nullptr, s.ff, n, in, nullptr);
        case TMPS_TORQUE: return launch_field_unpack(s, nullptr, s.tqf, n, in, nullptr);
        default: return tmps_fail(TMPS_ERR_ARG, "unknown field");
    }
}

int tmps_sync(tmps_system* h) {
    HIP_TRY(hipStreamSynchronize(sys_view(h)->stream));
    return 0;
}

int tmps_set_box(tmps_system* h, const double lo[3], const double hi[3], const int periodic[3],
                 const double h_rate[6], int vremap) {
    System& s = *sys(h);
    if (split(s)) {
        // the atoms stay where they are while the box changes under them
        int rc = split_positions_to_xr(s);
        if (rc) return rc;
        s.x64_valid = true;
    }
    int rc = ctl_pull(s);
    if (rc) return rc;
    DevBox& B = s.hctl->box;
    for (int k = 0; k < 3; ++k) {
        B.lo[k] = lo[k];
        B.hi[k] = hi[k];
        B.prd[k] = B.hi[k] - B.lo[k];
        B.per[k] = periodic[k];
    }
    for (int k = 0; k < 6; ++k) B.h_rate[k] = h_rate ? h_rate[k] : 0.0;
    B.vremap = vremap;
    return ctl_push(s);
}

int tmps_get_box(tmps_system* h, double lo[3], double hi[3], double h_rate[6], int* vremap) {
    System& s = *sys_view(h);
    int rc = ctl_pull_view(s);
    if (rc) return rc;
    const DevBox& B = s.hctl->box;
    for (int k = 0; k < 3; ++k) {
        if (lo) lo[k] = B.lo[k];
        if (hi) hi[k] = B.hi[k];
    }
    if (h_rate)
        for (int k = 0; k < 6; ++k) h_rate[k] = B.h_rate[k];
    if (vremap) *vremap = B.vremap;
    return 0;
}

int tmps_pbc(tmps_system* h) { return pbc(*sys(h)); }

int tmps_clear_forces(tmps_system* h) { return clear_forces(*sys(h)); }

int tmps_neigh_config(tmps_system* h, double skin, int delay, int every, int check,
                      int max_grid_cells, int max_neighbors_per_atom) {
    System& s = *sys_view(h);
    if (every < 1) return tmps_fail(TMPS_ERR_ARG, "neigh_modify every must be >= 1");
    if (max_grid_cells < 1) return tmps_fail(TMPS_ERR_ARG, "max_grid_cells must be >= 1");
    if (max_neighbors_per_atom < 1)
        return tmps_fail(TMPS_ERR_ARG, "max_neighbors_per_atom must be >= 1");
    s.max_nb = max_neighbors_per_atom;
    s.skin = skin;
    s.delay = delay;
    s.every = every;
    s.check = check ? 1 : 0;
    s.max_grid_cells = max_grid_cells;
    return 0;
}

int tmps_neigh_check(tmps_system* h, int* moved) { return neigh_check(*sys(h), moved); }

int tmps_neigh_decide(tmps_system* h, int timestep, int have_timestep, int* rebuild) {
    return neigh_decide(*sys(h), timestep, have_timestep, rebuild);
}

int tmps_neigh_build(tmps_system* h) { return neigh_build(*sys(h)); }

int tmps_neigh_info(tmps_system* h, int gdim[3], double csize[3], int* npairs) {
    System& s = *sys(h);
    for (int k = 0; k < 3; ++k) {
        if (gdim) gdim[k] = s.ref_grid.g[k];
        if (csize) csize[k] = s.ref_grid.csize[k];
    }
    if (npairs) *npairs = s.npairs;
    return 0;
}

int tmps_neigh_get_pairs(tmps_system* h, int* pair_i, int* pair_j, int* partner, double* shear,
                         int cap) {
    System& s = *sys(h);
    if (s.n <= 0 || s.npairs == 0) return 0;
    if (cap < s.npairs) return tmps_fail(TMPS_ERR_ARG, "pair buffer too small");
    const int m = s.npairs;
    int *dpi = nullptr, *dpj = nullptr, *dpa = nullptr, *dst = nullptr;
    double* dsh = nullptr;
    int rc = dalloc(s.stream, &dpi, m);
    if (!rc) rc = dalloc(s.stream, &dpj, m);
    if (!rc) rc = dalloc(s.stream, &dpa, m);
    if (!rc) rc = dalloc(s.stream, &dsh, 3 * (size_t)m);
    if (!rc) rc = dalloc(s.stream, &dst, s.n + 1);
    if (!rc) rc = launch_half_pairs(s, dpi, dpj, dpa, dsh, dst);
    if (!rc && hipStreamSynchronize(s.stream) != hipSuccess) rc = TMPS_ERR_HIP;
    if (!rc) {
        if (pair_i) (void)copy_sync(s.stream, pair_i, dpi, sizeof(int) * m, hipMemcpyDeviceToHost);
        if (pair_j) (void)copy_sync(s.stream, pair_j, dpj, sizeof(int) * m, hipMemcpyDeviceToHost);
        if (partner) (void)copy_sync(s.stream, partner, dpa, sizeof(int) * m, hipMemcpyDeviceToHost);
        if (shear) (void)copy_sync(s.stream, shear, dsh, sizeof(double) * 3 * m, hipMemcpyDeviceToHost);
    }
    dfree(dpi); dfree(dpj); dfree(dpa); dfree(dsh); dfree(dst);
    return rc;
}

int tmps_neigh_get_counts(tmps_system* h, int* num_neighbors, int n) {
    System& s = *sys(h);
    if (n > s.n) n = s.n;
    if (n <= 0) return 0;
    int* dst = nullptr;
    int rc = dalloc(s.stream, &dst, s.n + 1);
    if (!rc) rc = launch_half_pairs(s, nullptr, nullptr, nullptr, nullptr, dst);
    if (!rc && hipStreamSynchronize(s.stream) != hipSuccess) rc = TMPS_ERR_HIP;
    if (!rc && copy_sync(s.stream, num_neighbors, s.half_count, sizeof(int) * n, hipMemcpyDeviceToHost) !=
                   hipSuccess)
        rc = TMPS_ERR_HIP;
    dfree(dst);
    return rc;
}

int tmps_neigh_get_bins(tmps_system* h, int* cell_of, int* cell_start, int* cell_particles) {
    System& s = *sys(h);
    if (s.ref_grid.ncell == 0) return tmps_fail(TMPS_ERR_ARG, "no neighbor list has been built");
    // bin the current positions on the reference grid (the list's own bins
    // are scratch of the build)
    int rc = split_positions_to_xr(s);
    if (rc) return rc;
    const GridP keep = s.grid;
    s.grid = s.ref_grid;
    rc = ensure_cells(s, s.grid.ncell);
    if (!rc) rc = launch_bins(s, s.xr);
    const GridP ref = s.grid;
    s.grid = keep;
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s.stream));
    const GridP& G = ref;
    if (s.n > 0 && cell_of)
        HIP_TRY(copy_sync(s.stream, cell_of, s.cell_of, sizeof(int) * s.n, hipMemcpyDeviceToHost));
    if (cell_start)
        HIP_TRY(copy_sync(s.stream, cell_start, s.cell_start, sizeof(int) * (G.ncell + 1),
                          hipMemcpyDeviceToHost));
    if (s.n > 0 && cell_particles)
        HIP_TRY(copy_sync(s.stream, cell_particles, s.cell_particles, sizeof(int) * s.n, hipMemcpyDeviceToHost));
    return 0;
}

int tmps_history_reset(tmps_system* h) {
    System& s = *sys(h);
    s.have_old = false;
    s.seed_pending = false;
    return launch_hist_clear(s);
}

int tmps_history_save(tmps_system* h) { return history_save(*sys(h)); }

int tmps_history_restore(tmps_system* h) { return history_restore(*sys(h)); }

int tmps_history_set_pairs(tmps_system* h, const int* partner, const double* shear, int npairs) {
    System& s = *sys(h);
    if (npairs != s.npairs) return tmps_fail(TMPS_ERR_ARG, "history length differs from npairs");
    if (npairs == 0) return 0;
    int *dpa = nullptr, *dst = nullptr;
    double* dsh = nullptr;
    int rc = dalloc(s.stream, &dpa, npairs);
    if (!rc) rc = dalloc(s.stream, &dsh, 3 * (size_t)npairs);
    if (!rc) rc = dalloc(s.stream, &dst, s.n + 1);
    if (!rc && copy_sync(s.stream, dpa, partner, sizeof(int) * npairs, hipMemcpyHostToDevice) != hipSuccess)
        rc = TMPS_ERR_HIP;
    if (!rc && copy_sync(s.stream, dsh, shear, sizeof(double) * 3 * npairs, hipMemcpyHostToDevice) != hipSuccess)
        rc = TMPS_ERR_HIP;
    if (!rc) rc = launch_half_pairs(s, nullptr, nullptr, nullptr, nullptr, dst);
    if (!rc) rc = launch_set_half_history(s, dpa, dsh, dst);
    if (!rc && hipStreamSynchronize(s.stream) != hipSuccess) rc = TMPS_ERR_HIP;
    dfree(dpa); dfree(dsh); dfree(dst);
    return rc;
}

int tmps_history_seed(tmps_system* h, int n, const int* start, const int* nbr,
                      const double* shear) {
    System& s = *sys(h);
    if (n != s.n) return tmps_fail(TMPS_ERR_ARG, "a seeded history covers exactly nlocal atoms");
    if (!start || (start[n] > 0 && (!nbr || !shear)))
        return tmps_fail(TMPS_ERR_ARG, "null seeded history arrays");
    if (start[0] != 0) return tmps_fail(TMPS_ERR_ARG, "seeded history rows must start at 0");
    for (int i = 0; i < n; ++i) {
        if (start[i + 1] < start[i])
            return tmps_fail(TMPS_ERR_ARG, "seeded history row offsets must not decrease");
        if (start[i + 1] - start[i] > s.max_nb)
            return tmps_fail(TMPS_ERR_ARG, "an atom has more stored contacts than the history "
                                           "has slots per atom");
        for (int k = start[i]; k < start[i + 1]; ++k) {
            if (nbr[k] < 0 || nbr[k] >= n || nbr[k] == i)
                return tmps_fail(TMPS_ERR_ARG, "seeded history partner out of range");
            if (k > start[i] && nbr[k] <= nbr[k - 1])
                return tmps_fail(TMPS_ERR_ARG, "seeded history rows must be strictly ascending");
        }
    }
    const size_t total = (size_t)start[n];
    const size_t hsz = split(s) ? sizeof(float4) : sizeof(double4);
    if (total > s.onb_cap || s.onb_j == nullptr) {
        dfree(s.onb_j);
        dfree(s.onb_flag);
        if (s.onb_hist) (void)hipFree(s.onb_hist);
        s.onb_hist = nullptr;
        const size_t cap = total + total / 4 + 1024;
        int rc = dalloc(s.stream, &s.onb_j, cap);
        if (!rc) rc = dalloc(s.stream, &s.onb_flag, cap);
        if (rc) return rc;
        HIP_TRY(hipMalloc(&s.onb_hist, cap * hsz));
        s.onb_cap = cap;
    }
    HIP_TRY(hipStreamSynchronize(s.stream));
    HIP_TRY(copy_sync(s.stream, s.onb_start, start, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice));
    if (total > 0) {
        HIP_TRY(copy_sync(s.stream, s.onb_j, nbr, sizeof(int) * total, hipMemcpyHostToDevice));
        std::vector<uint8_t> ones(total, 1);
        HIP_TRY(copy_sync(s.stream, s.onb_flag, ones.data(), total, hipMemcpyHostToDevice));
        if (split(s)) {
            std::vector<float4> hv(total);
            for (size_t k = 0; k < total; ++k)
                hv[k] = make_float4((float)shear[3 * k], (float)shear[3 * k + 1],
                                    (float)shear[3 * k + 2], 0.0f);
            HIP_TRY(copy_sync(s.stream, s.onb_hist, hv.data(), total * sizeof(float4), hipMemcpyHostToDevice));
        } else {
            std::vector<double4> hv(total);
            for (size_t k = 0; k < total; ++k)
                hv[k] = make_double4(shear[3 * k], shear[3 * k + 1], shear[3 * k + 2], 0.0);
            HIP_TRY(copy_sync(s.stream, s.onb_hist, hv.data(), total * sizeof(double4),
                              hipMemcpyHostToDevice));
        }
    }
    s.onslots = (int)total;
    s.onb_n = n;
    s.have_old = true;
    s.seed_pending = true;
    return 0;
}

int tmps_pair_set(tmps_system* h, const tmps_pair_params* p) {
    System& s = *sys_view(h);
    if (!p) return tmps_fail(TMPS_ERR_ARG, "null pair params");
    if (p->style < 0 || p->style > 2) return tmps_fail(TMPS_ERR_ARG, "unknown pair style");
    PairP& q = s.pair;
    q.style = p->style;
    q.limit_damping = p->limit_damping;
    q.kn = p->kn; q.kt = p->kt; q.gamman = p->gamman; q.gammat = p->gammat; q.xmu = p->xmu;
    q.normal_model = p->normal_model;
    q.damping_model = p->damping_model;
    q.tangential_model = p->tangential_model;
    q.uses_history = (p->style != TMPS_PAIR_GRANULAR) || p->tangential_model == 2 ||
                     p->tangential_model == 3;
    q.g_kn = p->g_kn; q.g_damp = p->g_damp; q.g_kt = p->g_kt; q.g_xt = p->g_xt; q.g_mu = p->g_mu;
    s.have_pair = true;
    invalidate_graphs(s);
    return 0;
}

int tmps_pair_compute(tmps_system* h, double dt, int shearupdate, int accumulate) {
    return pair_compute(*sys(h), dt, shearupdate, accumulate);
}

int tmps_nve_initial(tmps_system* h, double dt) {
    System& s = *sys(h);
    int rc = ensure_split(s);
    if (rc) return rc;
    return split(s) ? launch_nve_initial_split(s, dt) : launch_nve_initial_f64(s, dt);
}

int tmps_nve_final(tmps_system* h, double dt) {
    System& s = *sys(h);
    return split(s) ? launch_nve_final_split(s, dt) : launch_nve_final_f64(s, dt);
}

int tmps_cundall(tmps_system* h, double gamma_lin, double gamma_ang) {
    return cundall(*sys(h), gamma_lin, gamma_ang);
}

int tmps_servo_set(tmps_system* h, const tmps_servo_params* p) {
    System& s = *sys_view(h);
    if (!p) return tmps_fail(TMPS_ERR_ARG, "null servo params");
    for (int i = 0; i < 3; ++i)
        if (p->style[i] < 0 || p->style[i] > 9)
            return tmps_fail(TMPS_ERR_UNSUPPORTED, "unknown deform style code");
    if (p->couple < 0 || p->couple > 4) return tmps_fail(TMPS_ERR_ARG, "unknown couple code");
    if (p->remap < 0 || p->remap > 2) return tmps_fail(TMPS_ERR_ARG, "unknown remap code");
    if (split(s) && p->enabled && p->remap != 0) {
        for (int i = 0; i < 3; ++i)
            if (p->style[i] != 0)
                return tmps_fail(TMPS_ERR_UNSUPPORTED,
                                 "FP32-split supports a deforming box only with remap x");
    }
    tmps_servo_params q = *p;
    if (q.nevery < 1) q.nevery = 1;
    // a captured segment bakes in only whether the servo runs
    const bool was_on = s.have_servo && s.servo.enabled;
    s.servo = q;
    s.have_servo = true;
    if (was_on != (s.servo.enabled != 0)) invalidate_graphs(s);
    int rc = ctl_edit(s);
    if (rc) return rc;
    s.hctl->servo.p = s.servo;
    return ctl_commit(s);
}

int tmps_servo_setup(tmps_system* h, int nsteps_total, double dt) {
    System& s = *sys_view(h);
    int rc = ctl_edit(s);
    if (rc) return rc;
    DevServo& SV = s.hctl->servo;
    SV.p = s.servo;
    SV.nsteps = 0;
    SV.nsteps_total = nsteps_total;
    DevBox& B = s.hctl->box;
    for (int i = 0; i < 3; ++i) {
        SV.lo_start[i] = B.lo[i];
        SV.hi_start[i] = B.hi[i];
        SV.cum[i] = 0.0;
    }
    double hr[6];
    strain_h_rate(s.servo, SV, dt, hr);
    for (int k = 0; k < 6; ++k) B.h_rate[k] = hr[k];
    B.vremap = s.servo.remap == 1 ? 1 : 0;
    return ctl_commit(s);
}

int tmps_servo_end_of_step(tmps_system* h, double dt, double tensor_out[6]) {
    return servo_end_of_step(*sys(h), dt, tensor_out);
}

int tmps_virial_reduce(tmps_system* h, int keflag, double out6[6]) {
    System& s = *sys(h);
    for (int k = 0; k < 6; ++k) out6[k] = 0.0;
    if (s.n <= 0) return 0;
    int rc = split(s) ? launch_leaves_split(s, keflag) : 0;
    if (!split(s)) {
        rc = launch_leaves_f64(s, keflag);
        if (!rc) rc = launch_final_tree(s, 6, nblocks_for(s.n));
    }
    if (rc) return rc;
    rc = ctl_pull(s);
    if (rc) return rc;
    for (int k = 0; k < 6; ++k) out6[k] = s.hctl->vsum[k];
    return 0;
}

int tmps_ke(tmps_system* h, double* ke_trans, double* ke_rot) {
    System& s = *sys(h);
    *ke_trans = *ke_rot = 0.0;
    if (s.n <= 0) return 0;
    int rc = split(s) ? launch_ke_split(s, nullptr) : launch_ke_f64(s, nullptr);
    if (rc) return rc;
    rc = ctl_pull(s);
    if (rc) return rc;
    *ke_trans = s.hctl->vsum[0];
    *ke_rot = s.hctl->vsum[1];
    return 0;
}

// ---------------------------------------------------------------- walls

struct tmps_wall {
    System* sys = nullptr;
    tmps_wall_params p{};
    double4* hist = nullptr;  // [max_atoms]: shear xyz, touch flag in w
    // the history of a fused run's second buffer set (a step reads one set
    // and writes the other); a run leaves its results in hist
    double4* hist2 = nullptr;
    int cap = 0;
};

int tmps_wall_create(tmps_system* h, const tmps_wall_params* p, tmps_wall** out) {
    System& s = *sys(h);
    if (!p || !out) return tmps_fail(TMPS_ERR_ARG, "null wall arguments");
    *out = nullptr;
    if (p->geometry != TMPS_WALL_PLANE && p->geometry != TMPS_WALL_CYLINDER)
        return tmps_fail(TMPS_ERR_ARG, "unknown wall geometry");
    if (p->axis < 0 || p->axis > 2) return tmps_fail(TMPS_ERR_ARG, "wall_axis must be 0, 1 or 2");
    if (p->side != -1 && p->side != 1)
        return tmps_fail(TMPS_ERR_ARG, "wall_side must be -1 (low) or +1 (high)");
    if (p->shear_axis == p->axis)
        return tmps_fail(TMPS_ERR_ARG,
                         "shear must act in the plane of the wall, not along its normal");
    if (p->shear_axis < -1 || p->shear_axis > 2)
        return tmps_fail(TMPS_ERR_ARG, "shear axis must be -1, 0, 1 or 2");
    tmps_wall* w = new tmps_wall();
    w->sys = &s;
    w->p = *p;
    w->cap = s.max_atoms;
    int rc = dalloc(s.stream, &w->hist, (size_t)std::max(s.max_atoms, 1));
    if (!rc) rc = dalloc(s.stream, &w->hist2, (size_t)std::max(s.max_atoms, 1));
    if (rc) {
        dfree(w->hist);
        delete w;
        return rc;
    }
    *out = w;
    return 0;
}

int tmps_wall_destroy(tmps_wall* w) {
    if (!w) return 0;
    // touches only the wall's own memory: the system may already be gone
    // (hipFree waits for work that still uses the buffer)
    dfree(w->hist);
    dfree(w->hist2);
    delete w;
    return 0;
}

int tmps_wall_reset(tmps_wall* w) {
    if (!w) return tmps_fail(TMPS_ERR_ARG, "null wall");
    HIP_TRY(hipMemsetAsync(w->hist, 0, sizeof(double4) * (size_t)w->cap, w->sys->stream));
    HIP_TRY(hipMemsetAsync(w->hist2, 0, sizeof(double4) * (size_t)w->cap, w->sys->stream));
    return 0;
}

int tmps_run_set_walls(tmps_system* h, int n, tmps_wall* const* walls, const int* before_cundall,
                       const double* coord, const double* vwall_s, double dt) {
    System& s = *sys_view(h);
    if (n < 0 || n > TMPS_RUN_WALLS)
        return tmps_fail(TMPS_ERR_UNSUPPORTED, "a device-resident run applies at most 2 walls");
    if (n > 0 && (!walls || !before_cundall || !coord || !vwall_s))
        return tmps_fail(TMPS_ERR_ARG, "null wall arguments");
    WallSpec spec[TMPS_RUN_WALLS] = {};
    for (int q = 0; q < n; ++q) {
        const tmps_wall* w = walls[q];
        if (!w || w->sys != &s) return tmps_fail(TMPS_ERR_ARG, "a wall of another system");
        if (s.n > w->cap) return tmps_fail(TMPS_ERR_CAPACITY, "more atoms than the wall's history");
        spec[q].p = w->p;
        spec[q].coord = coord[q];
        spec[q].vwall_n = 0.0;
        spec[q].vwall_s = vwall_s[q];
        spec[q].dt = dt;
        spec[q].before_cundall = before_cundall[q] ? 1 : 0;
    }
    if (n > 0 && !s.d_run_spec) HIP_TRY(hipMalloc(&s.d_run_spec, sizeof(WallSpec) * TMPS_RUN_WALLS));
    if (n > 0 && !s.run_wnz) HIP_TRY(hipMalloc(&s.run_wnz, (size_t)std::max(s.max_atoms, 1)));
    if (n > 0 && std::memcmp(spec, s.run_spec, sizeof(spec)) != 0)
        HIP_TRY(copy_sync(s.stream, s.d_run_spec, spec, sizeof(spec), hipMemcpyHostToDevice));
    std::memcpy(s.run_spec, spec, sizeof(spec));
    s.nrun_walls = n;
    for (int q = 0; q < TMPS_RUN_WALLS; ++q) {
        s.run_wh[q][0] = q < n ? walls[q]->hist : nullptr;
        s.run_wh[q][1] = q < n ? walls[q]->hist2 : nullptr;
    }
    return 0;
}

int tmps_wall_post_force(tmps_wall* w, double dt, double coord, double vwall_n, double vwall_s,
                         int history_update) {
    if (!w) return tmps_fail(TMPS_ERR_ARG, "null wall");
    System& s = *w->sys;
    if (s.n > w->cap) return tmps_fail(TMPS_ERR_CAPACITY, "more atoms than the wall's history");
    // FP32-split: the wall reads FP64 positions (authority unchanged)
    int rc = split_positions_to_xr(s);
    if (rc) return rc;
    return launch_wall_gran(s, w->hist, w->p, dt, coord, vwall_n, vwall_s,
                            history_update ? 1 : 0);
}

int tmps_wall_get_history(tmps_wall* w, double* shear, int* touch, int n) {
    if (!w) return tmps_fail(TMPS_ERR_ARG, "null wall");
    System& s = *w->sys;
    if (n < 0 || n > w->cap) return tmps_fail(TMPS_ERR_ARG, "history count out of range");
    std::vector<double4> hv((size_t)std::max(n, 1));
    HIP_TRY(hipStreamSynchronize(s.stream));
    if (n > 0)
        HIP_TRY(copy_sync(s.stream, hv.data(), w->hist, sizeof(double4) * (size_t)n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        if (shear) {
            shear[3 * i] = hv[i].x;
            shear[3 * i + 1] = hv[i].y;
            shear[3 * i + 2] = hv[i].z;
        }
        if (touch) touch[i] = hv[i].w != 0.0 ? 1 : 0;
    }
    return 0;
}

// ------------------------------------------------------- per-contact computes

// The computes read the FP64 positions in xr (an FP32-split system decodes
// its positions there first) and the current neighbor list; before the
// first build every row is empty, as the reference's zeroed num_neighbors.
static int computes_prepare(System& s) {
    if (!s.cmp_sum) {
        int rc = dalloc(s.stream, &s.cmp_sum, 8);
        if (rc) return rc;
    }
    return split_positions_to_xr(s);
}

int tmps_contact_atom(tmps_system* h, int* counts, long long* total) {
    System& s = *sys(h);
    if (total) *total = 0;
    if (s.n <= 0) return 0;
    std::vector<int> tmp;
    int* out = counts;
    if (!out) {
        tmp.assign(s.n, 0);
        out = tmp.data();
    }
    if (!s.has_built) {
        std::fill(out, out + s.n, 0);
        return 0;
    }
    int rc = computes_prepare(s);
    if (rc) return rc;
    if (s.cmp_count_cap < s.n) {
        dfree(s.cmp_count);
        rc = dalloc(s.stream, &s.cmp_count, (size_t)s.n);
        if (rc) return rc;
        s.cmp_count_cap = s.n;
    }
    rc = launch_contact_atom(s, s.cmp_count);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, s.cmp_count, sizeof(int) * (size_t)s.n, hipMemcpyDeviceToHost,
                           s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    if (total) {
        long long t = 0;
        for (int i = 0; i < s.n; ++i) t += out[i];
        *total = t;
    }
    return 0;
}

int tmps_fabric(tmps_system* h, double phi[6], int* ncontacts) {
    System& s = *sys(h);
    for (int k = 0; k < 6; ++k) phi[k] = 0.0;
    if (ncontacts) *ncontacts = 0;
    if (s.n <= 0 || !s.has_built) return 0;
    int rc = computes_prepare(s);
    if (!rc) rc = launch_fabric(s, s.cmp_sum);
    if (rc) return rc;
    double sum[8];
    HIP_TRY(hipMemcpyAsync(sum, s.cmp_sum, sizeof(double) * 7, hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    const int nc = (int)sum[6];
    if (ncontacts) *ncontacts = nc;
    if (nc > 0) {
        // fabric.py:70-76: phi = sum * (1 / Nc)
        const double inv_nc = 1.0 / (double)nc;
        for (int k = 0; k < 6; ++k) phi[k] = sum[k] * inv_nc;
    }
    return 0;
}

// ---- fabric without allocation

// The no-allocation fabric's scratch: the block partials (up to CLOSE_SPAN
// blocks of 8 doubles, i.e. 1,048,576 atoms) and the 8-double sum. One array
// per GPU, alive until the process ends; every call on a GPU uses it, so
// calls that overlap on one GPU race.
#define TMPS_NOALLOC_PARTIALS 4096
static __device__ __align__(256) double scratch[TMPS_NOALLOC_PARTIALS * 8 + 8];

namespace {
// scratch's device address, per GPU (the array's name is no device address in
// host code): taken once per device with that device current
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

int scratch_ptr(int device, double** out) {
    if (device < 0 || device >= 64) return tmps_fail(TMPS_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != device) HIP_TRY(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) return tmps_fail_hip(e, "hipGetSymbolAddress(scratch)", __LINE__);
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<double*>(g_scratch_ptr[device]);
    return 0;
}
}  // namespace

int tmps_fabric_noalloc(tmps_system* h, double phi[6], int* ncontacts) {
    System& s = *sys(h);
    for (int k = 0; k < 6; ++k) phi[k] = 0.0;
    if (ncontacts) *ncontacts = 0;
    if (s.n <= 0 || !s.has_built) return 0;
    if (nblocks_for(s.n) > TMPS_NOALLOC_PARTIALS)
        return tmps_fail(TMPS_ERR_CAPACITY,
                         "tmps_fabric_noalloc takes at most 1,048,576 atoms (its fixed scratch "
                         "holds 4,096 block partials); use tmps_fabric");
    double* sc = nullptr;
    int rc = scratch_ptr(s.device, &sc);
    if (!rc) rc = split_positions_to_xr(s);
    if (!rc) rc = launch_fabric_into(s, sc, sc + TMPS_NOALLOC_PARTIALS * 8);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(s.hsum, sc + TMPS_NOALLOC_PARTIALS * 8, sizeof(double) * 7,
                           hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    const int nc = (int)s.hsum[6];
    if (ncontacts) *ncontacts = nc;
    if (nc > 0) {
        const double inv_nc = 1.0 / (double)nc;
        for (int k = 0; k < 6; ++k) phi[k] = s.hsum[k] * inv_nc;
    }
    return 0;
}

int tmps_stress_atom(tmps_system* h, int keflag, double* stress, double sum6[6]) {
    System& s = *sys(h);
    if (sum6)
        for (int k = 0; k < 6; ++k) sum6[k] = 0.0;
    if (s.n <= 0) return 0;
    int rc = computes_prepare(s);
    if (rc) return rc;
    if (s.cmp_stress_cap < s.n) {
        dfree(s.cmp_stress);
        rc = dalloc(s.stream, &s.cmp_stress, 6 * (size_t)s.n);
        if (rc) return rc;
        s.cmp_stress_cap = s.n;
    }
    rc = launch_stress_atom(s, keflag ? 1 : 0, s.cmp_stress, s.cmp_sum);
    if (rc) return rc;
    if (stress)
        HIP_TRY(hipMemcpyAsync(stress, s.cmp_stress, sizeof(double) * 6 * (size_t)s.n,
                               hipMemcpyDeviceToHost, s.stream));
    double sum[8];
    HIP_TRY(hipMemcpyAsync(sum, s.cmp_sum, sizeof(double) * 6, hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    if (sum6)
        for (int k = 0; k < 6; ++k) sum6[k] = sum[k];
    return 0;
}

int tmps_run(tmps_system* h, int nsteps, double dt, int run_pre, int use_cundall, double gamma_lin,
             double gamma_ang) {
    return run(*sys_view(h), nsteps, dt, run_pre, use_cundall, gamma_lin, gamma_ang);
}

int tmps_prepare_run(tmps_system* h, int nsteps, double dt, int use_cundall, double gamma_lin,
                     double gamma_ang) {
    return prepare_run(*sys_view(h), nsteps, dt, use_cundall, gamma_lin, gamma_ang);
}

int tmps_get_counters(tmps_system* h, tmps_counters* out) {
    System& s = *sys_view(h);
    int rc = ctl_pull_view(s);
    if (rc) return rc;
    out->timestep = s.timestep;
    out->has_built = s.has_built;
    out->build_count = s.build_count;
    out->last_build_step = s.last_build_step;
    out->current_step = s.current_step;
    out->setup_done = s.setup_done;
    out->npairs = s.npairs;
    out->nslots = s.nslots;
    out->ntouch = s.hctl->ntouch;
    out->rebuilds = s.rebuilds;
    out->servo_step_count = s.hctl->servo.step_count;
    out->servo_nsteps = s.hctl->servo.nsteps;
    return 0;
}

int tmps_set_counters(tmps_system* h, const tmps_counters* in) {
    System& s = *sys_view(h);
    s.timestep = in->timestep;
    s.has_built = in->has_built;
    s.build_count = in->build_count;
    s.last_build_step = in->last_build_step;
    s.current_step = in->current_step;
    s.setup_done = in->setup_done;
    int rc = ctl_edit(s);
    if (rc) return rc;
    s.hctl->servo.step_count = in->servo_step_count;
    s.hctl->servo.nsteps = in->servo_nsteps;
    s.hctl->last_build_step = s.last_build_step;
    return ctl_commit(s);
}

int tmps_last_pressure(tmps_system* h, double out6[6]) {
    System& s = *sys_view(h);
    int rc = ctl_pull_view(s);
    if (rc) return rc;
    for (int k = 0; k < 6; ++k) out6[k] = s.hctl->tensor[k];
    return 0;
}

int tmps_set_profiling(tmps_system* h, int enabled) {
    System& s = *sys_view(h);
    s.profiling = enabled != 0;
    for (int c = 0; c < TMPS_PROF_CLASSES; ++c) {
        s.prof_ms[c] = 0.0;
        s.prof_cnt[c] = 0;
    }
    return 0;
}

int tmps_set_launch_mode(tmps_system* h, int mode) {
    System& s = *sys_view(h);
    if (mode != TMPS_LAUNCH_GRAPH && mode != TMPS_LAUNCH_DIRECT)
        return tmps_fail(TMPS_ERR_ARG, "launch mode must be TMPS_LAUNCH_GRAPH or TMPS_LAUNCH_DIRECT");
    s.launch_mode = mode;
    return 0;
}

int tmps_kernel_times(tmps_system* h, double out_ms[TMPS_PROF_CLASSES],
                      int counts[TMPS_PROF_CLASSES]) {
    System& s = *sys(h);
    for (int c = 0; c < TMPS_PROF_CLASSES; ++c) {
        out_ms[c] = s.prof_cnt[c] ? s.prof_ms[c] / s.prof_cnt[c] : 0.0;
        if (counts) counts[c] = s.prof_cnt[c];
    }
    return 0;
}

}  // extern "C"
