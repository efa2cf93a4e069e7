// Ensembles: many independent FP32-split systems (periodic packings of
// uneven size, each with its own box, servo, contact coefficients and pair
// style) advanced together. No reference counterpart: the reference runs one
// system per Simulation.
//
// Each member keeps its own buffers and control block. A captured segment
// advances every running member with one launch per step and pair style
// (k_batch_step_split over a block map of (member, local block)) and one
// close launch with a block per member (k_batch_close). Each member carries
// its own progress on the device (the buffer set its next step reads, a
// pending half-kick, the step that ends its run), so a member whose skin
// check asks for a rebuild stops there while the others never wait. After
// the segment the steps of all such members are redone together, each with
// its list rebuild: the host work of the single-system redo_step per member,
// and each of its kernels launched once for all of them over block maps
// (redo_batched). The members then rejoin the next segment. A member's
// arithmetic, bins, list, tree and servo are those of its single-system run,
// so its results are bitwise the ones tmps_run gives it alone.
//
// Launch configuration (segment length, order of the blocks in the map)
// changes no result; tmps_batch_run picks it by timing the first segments of
// the run itself (the autotuner).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "tmps_internal.h"

namespace {

__global__ void k_batch_status(const CloseEntry* table, int n, int4* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Ctl* c = table[k].ctl;
    out[k] = make_int4(c->timestep, c->abort, c->error, c->bdone);
}

// Every member's control block into one array (one block per member), read
// back with the status: the host mirror a member's redo starts from.
__global__ void k_batch_ctl_gather(const CloseEntry* table, Ctl* out) {
    static_assert(sizeof(Ctl) % 4 == 0, "the control block is copied in 4-byte words");
    const unsigned* src = reinterpret_cast<const unsigned*>(table[blockIdx.x].ctl);
    unsigned* dst = reinterpret_cast<unsigned*>(out + blockIdx.x);
    for (int w = threadIdx.x; w < (int)(sizeof(Ctl) / 4); w += blockDim.x) dst[w] = src[w];
}

// A redone member rejoins: its next step reads set 0 after a half-kick.
__global__ void k_batch_rejoin(Ctl* ctl, int done) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->bpar = 0;
        ctl->bkick = 1;
        ctl->bdone = done;
    }
}

struct Config {
    int seg_len;
    int order;  // 0: members in sequence, 1: members grouped per XCD
};

const Config kConfigs[] = {{8, 0}, {8, 1}, {16, 0}, {16, 1}, {32, 0}, {32, 1}};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);


}  // namespace

struct tmps_batch_impl {
    std::vector<System*> sys;
    std::vector<int> natoms;  // members' atom counts when the tables were built
    std::vector<int> styles;  // and their pair styles
    hipStream_t stream = nullptr;
    void* d_table = nullptr;
    std::vector<char> h_table;
    size_t entry_size = 0;
    CloseEntry* d_close = nullptr;
    // block maps per order: every member's blocks, and per pair style
    int2* d_all[2] = {nullptr, nullptr};
    int2* d_style[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int nb_all = 0, nb_style[3] = {0, 0, 0};
    int max_parts = 1;
    int4* d_status = nullptr;
    int4* h_status = nullptr;
    Ctl* d_ctls = nullptr;  // members' control blocks after the last segment
    Ctl* h_ctls = nullptr;
    hipGraphExec_t exec[kNumConfigs] = {};
    double exec_dt[kNumConfigs] = {};
    int config = -1;  // chosen launch configuration (-1: not tuned yet)
    double config_rate[kNumConfigs] = {};  // measured atom-steps per second
    int rebuilds = 0;
    // the batched redo: entries and block maps (pinned staging, one device
    // buffer), the members' list flags, and the last run's redo passes
    std::vector<RedoEntry> redo_entries;
    char* h_redo = nullptr;
    void* d_redo = nullptr;
    size_t redo_cap = 0;
    int4* d_redo_flags = nullptr;
    int4* h_redo_flags = nullptr;
    int redo_passes = 0;
    double redo_ms = 0.0;
};

namespace {

using Batch = tmps_batch_impl;

void free_batch(Batch& B) {
    if (B.stream) (void)hipStreamSynchronize(B.stream);
    for (auto& e : B.exec)
        if (e) (void)hipGraphExecDestroy(e);
    if (B.d_table) (void)hipFree(B.d_table);
    if (B.d_close) (void)hipFree(B.d_close);
    for (int o = 0; o < 2; ++o) {
        if (B.d_all[o]) (void)hipFree(B.d_all[o]);
        for (int k = 0; k < 3; ++k)
            if (B.d_style[o][k]) (void)hipFree(B.d_style[o][k]);
    }
    if (B.d_status) (void)hipFree(B.d_status);
    if (B.h_status) (void)hipHostFree(B.h_status);
    if (B.d_ctls) (void)hipFree(B.d_ctls);
    if (B.h_ctls) (void)hipHostFree(B.h_ctls);
    if (B.d_redo) (void)hipFree(B.d_redo);
    if (B.h_redo) (void)hipHostFree(B.h_redo);
    if (B.d_redo_flags) (void)hipFree(B.d_redo_flags);
    if (B.h_redo_flags) (void)hipHostFree(B.h_redo_flags);
    if (B.stream) (void)hipStreamDestroy(B.stream);
}

// Block maps: order 0 lists members one after another; order 1 first deals
// whole members to 8 bins (one per XCD, largest first onto the lightest bin)
// and then interleaves the bins, so that workgroup w (which the hardware
// hands to XCD w mod 8) mostly finds its member's other blocks, and their
// neighbours' lines, in the same L2.
void block_map(const std::vector<System*>& sys, const std::vector<int>& members, int order,
               std::vector<int2>& out) {
    out.clear();
    if (order == 0) {
        for (int m : members) {
            const int nb = nblocks_for(sys[m]->n);
            for (int b = 0; b < nb; ++b) out.push_back(make_int2(m, b));
        }
        return;
    }
    std::vector<int> ord(members);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return sys[a]->n > sys[b]->n; });
    std::vector<std::vector<int2>> bins(8);
    std::vector<long long> load(8, 0);
    for (int m : ord) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        const int nb = nblocks_for(sys[m]->n);
        for (int b = 0; b < nb; ++b) bins[x].push_back(make_int2(m, b));
        load[x] += nb;
    }
    std::vector<size_t> pos(8, 0);
    size_t total = 0;
    for (auto& b : bins) total += b.size();
    for (size_t w = 0; w < total; ++w) {
        int x = (int)(w % 8);
        for (int tries = 0; tries < 8 && pos[x] >= bins[x].size(); ++tries) x = (x + 1) % 8;
        out.push_back(bins[x][pos[x]++]);
    }
}

int upload(Batch& B, const std::vector<int2>& v, int2** d) {
    if (*d) {
        HIP_TRY(hipFree(*d));
        *d = nullptr;
    }
    if (v.empty()) return 0;
    HIP_TRY(hipMalloc((void**)d, v.size() * sizeof(int2)));
    HIP_TRY(hipMemcpyAsync(*d, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice, B.stream));
    return 0;
}

// (Re)writes member k's table entry on the host (its buffers can move at a
// rebuild); upload_table sends the table to the device.
int fill_entry(Batch& B, int k) {
    return batch_split_entry(*B.sys[k], B.h_table.data() + (size_t)k * B.entry_size);
}

int upload_table(Batch& B) {
    HIP_TRY(hipMemcpyAsync(B.d_table, B.h_table.data(), B.h_table.size(), hipMemcpyHostToDevice,
                           B.stream));
    return 0;
}

int style_of(const System& s) { return std::min(std::max(s.pair.style, 0), 2); }

// The close table: a member's servo can be switched between runs.
int write_close_table(Batch& B) {
    const int ninst = (int)B.sys.size();
    std::vector<CloseEntry> ce(ninst);
    for (int k = 0; k < ninst; ++k) {
        System& s = *B.sys[k];
        ce[k].ctl = s.ctl;
        ce[k].partial = s.partial;
        ce[k].nparts = nblocks_for(std::max(s.n, 1));
        ce[k].servo_on = (s.have_servo && s.servo.enabled) ? 1 : 0;
        ce[k].natoms = s.n;
        ce[k].pad = 0;
    }
    HIP_TRY(copy_sync(B.stream, B.d_close, ce.data(), ninst * sizeof(CloseEntry), hipMemcpyHostToDevice));
    return 0;
}

int build(Batch& B) {
    const int ninst = (int)B.sys.size();
    B.entry_size = batch_split_entry_size();
    B.h_table.assign((size_t)ninst * B.entry_size, 0);
    HIP_TRY(hipMalloc(&B.d_table, B.h_table.size()));
    HIP_TRY(hipMalloc((void**)&B.d_close, ninst * sizeof(CloseEntry)));
    HIP_TRY(hipMalloc((void**)&B.d_status, ninst * sizeof(int4)));
    HIP_TRY(hipHostMalloc((void**)&B.h_status, ninst * sizeof(int4), hipHostMallocDefault));
    HIP_TRY(hipMalloc((void**)&B.d_ctls, ninst * sizeof(Ctl)));
    HIP_TRY(hipHostMalloc((void**)&B.h_ctls, ninst * sizeof(Ctl), hipHostMallocDefault));
    HIP_TRY(hipMalloc((void**)&B.d_redo_flags, ninst * sizeof(int4)));
    HIP_TRY(hipHostMalloc((void**)&B.h_redo_flags, ninst * sizeof(int4), hipHostMallocDefault));
    std::vector<int> all, by_style[3];
    for (int k = 0; k < ninst; ++k) {
        System& s = *B.sys[k];
        B.max_parts = std::max(B.max_parts, nblocks_for(std::max(s.n, 1)));
        all.push_back(k);
        by_style[style_of(s)].push_back(k);
        B.styles.push_back(style_of(s));
    }
    std::vector<int2> map;
    for (int o = 0; o < 2; ++o) {
        block_map(B.sys, all, o, map);
        B.nb_all = (int)map.size();
        int rc = upload(B, map, &B.d_all[o]);
        if (rc) return rc;
        for (int st = 0; st < 3; ++st) {
            block_map(B.sys, by_style[st], o, map);
            B.nb_style[st] = (int)map.size();
            rc = upload(B, map, &B.d_style[o][st]);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipStreamSynchronize(B.stream));
    return 0;
}

// One segment: the pending half-kicks, `len` steps of every running member,
// the tail of the members that finished, and the members' status.
int enqueue_segment(Batch& B, const Config& c, double dt) {
    const float fdt = (float)dt;
    int rc = launch_batch_split_kick(B.d_table, B.d_all[c.order], B.nb_all, fdt, B.stream);
    for (int k = 0; k < c.seg_len && !rc; ++k) {
        for (int st = 0; st < 3 && !rc; ++st)
            rc = launch_batch_split_step(st, B.d_table, B.d_style[c.order][st], B.nb_style[st], fdt,
                                         B.stream);
        if (!rc)
            rc = launch_batch_close(B.d_close, (int)B.sys.size(), B.max_parts, dt, B.stream);
    }
    if (!rc) rc = launch_batch_split_tail(B.d_table, B.d_all[c.order], B.nb_all, B.stream);
    if (rc) return rc;
    const int ninst = (int)B.sys.size();
    hipLaunchKernelGGL(k_batch_status, dim3((ninst + 255) / 256), dim3(256), 0, B.stream, B.d_close,
                       ninst, B.d_status);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_batch_ctl_gather, dim3(ninst), dim3(256), 0, B.stream, B.d_close,
                       B.d_ctls);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Captures configuration ci's segment (once per time step size).
int prepare_segment(Batch& B, int ci, double dt) {
    const Config& c = kConfigs[ci];
    if (!B.exec[ci] || B.exec_dt[ci] != dt) {
        if (B.exec[ci]) {
            HIP_TRY(hipStreamSynchronize(B.stream));
            (void)hipGraphExecDestroy(B.exec[ci]);
            B.exec[ci] = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_segment(B, c, dt);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(B.stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return tmps_fail_hip(e, "hipStreamEndCapture", __LINE__);
        e = hipGraphInstantiate(&B.exec[ci], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return tmps_fail_hip(e, "hipGraphInstantiate", __LINE__);
        B.exec_dt[ci] = dt;
    }
    return 0;
}

int launch_segment(Batch& B, int ci, double dt) {
    int rc = prepare_segment(B, ci, dt);
    if (rc) return rc;
    HIP_TRY(hipGraphLaunch(B.exec[ci], B.stream));
    HIP_TRY(hipMemcpyAsync(B.h_status, B.d_status, B.sys.size() * sizeof(int4),
                           hipMemcpyDeviceToHost, B.stream));
    HIP_TRY(hipMemcpyAsync(B.h_ctls, B.d_ctls, B.sys.size() * sizeof(Ctl), hipMemcpyDeviceToHost,
                           B.stream));
    return 0;
}

// Starts member s's run of nsteps (as tmps_run does before its loop).
int start_member(System& s, int nsteps, double dt, int run_pre) {
    if (!s.have_pair) return tmps_fail(TMPS_ERR_ARG, "every ensemble member needs a pair style");
    int rc = 0;
    if (run_pre || !s.setup_done) {
        rc = tmps_api::init_simulation(s, dt, 0, 0.0, 0.0);
        if (rc) return rc;
    }
    rc = tmps_api::ensure_split(s);
    if (!rc) rc = tmps_api::ctl_pull(s);
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
    H.bpar = 0;
    H.bend = s.timestep + nsteps;
    H.bkick = 1;
    H.bdone = nsteps == 0 ? 1 : 0;
    s.rebuilds = 0;
    rc = tmps_api::ctl_push(s);
    if (!rc) rc = launch_split_consts(s, 0);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s.stream));
    return 0;
}

// Redoes the members in `which` together: each member's host work is
// redo_step's (tmps_api::redo_batch_begin / redo_batch_list), and each
// kernel of the redo runs once for all of them over block maps. Two host
// round trips in all: the list builds' flags, and the caller's final wait.
int redo_batched(Batch& B, const std::vector<int>& which, double dt) {
    const int m = (int)which.size();
    if (m == 0) return 0;
    std::vector<RedoEntry>& E = B.redo_entries;
    E.resize(m);
    std::vector<int2> atoms, cells, cell_tiles, atom_tiles, copy, save, style[3];
    int max_parts = 1;
    for (int q = 0; q < m; ++q) {
        System& s = *B.sys[which[q]];
        *s.hctl = B.h_ctls[which[q]];
        int rc = tmps_api::check_run_error(s);
        if (rc) return rc;
        s.timestep = s.hctl->timestep;
        const int par = s.hctl->bpar;
        rc = tmps_api::redo_batch_begin(s, E[q]);
        if (rc) return rc;
        const RedoEntry& e = E[q];
        const int nb = nblocks_for(e.n);
        for (int b = 0; b < nb; ++b) atoms.push_back(make_int2(q, b));
        for (int b = 0; b < nb; ++b) style[style_of(s)].push_back(make_int2(q, b));
        for (int b = 0; b < nblocks_for(e.ncell + 1); ++b) cells.push_back(make_int2(q, b));
        for (int b = 0; b < (e.ncell + TMPS_SCAN_TILE - 1) / TMPS_SCAN_TILE; ++b)
            cell_tiles.push_back(make_int2(q, b));
        for (int b = 0; b < (e.n + TMPS_SCAN_TILE - 1) / TMPS_SCAN_TILE; ++b)
            atom_tiles.push_back(make_int2(q, b));
        if (par == 1)
            for (int b = 0; b < e.copy_blocks; ++b) copy.push_back(make_int2(q, b));
        for (int b = 0; b < e.save_blocks; ++b) save.push_back(make_int2(q, b));
        max_parts = std::max(max_parts, nb);
    }
    // one staging buffer: the entries, then the maps
    const size_t entries_bytes = (size_t)m * sizeof(RedoEntry);
    std::vector<const std::vector<int2>*> lists = {&atoms, &cells, &cell_tiles, &atom_tiles,
                                                   &copy, &save, &style[0], &style[1], &style[2]};
    size_t bytes = entries_bytes;
    for (const auto* l : lists) bytes += l->size() * sizeof(int2);
    if (bytes > B.redo_cap) {
        HIP_TRY(hipStreamSynchronize(B.stream));
        if (B.d_redo) HIP_TRY(hipFree(B.d_redo));
        if (B.h_redo) HIP_TRY(hipHostFree(B.h_redo));
        B.d_redo = nullptr;
        B.h_redo = nullptr;
        B.redo_cap = 0;
        const size_t cap = bytes + bytes / 2 + 4096;
        HIP_TRY(hipMalloc(&B.d_redo, cap));
        HIP_TRY(hipHostMalloc((void**)&B.h_redo, cap, hipHostMallocDefault));
        B.redo_cap = cap;
    }
    std::memcpy(B.h_redo, E.data(), entries_bytes);
    const int2* dmap[9];
    size_t off = entries_bytes;
    for (int k = 0; k < 9; ++k) {
        const size_t nbytes = lists[k]->size() * sizeof(int2);
        if (nbytes) std::memcpy(B.h_redo + off, lists[k]->data(), nbytes);
        dmap[k] = reinterpret_cast<const int2*>(static_cast<char*>(B.d_redo) + off);
        off += nbytes;
    }
    RedoMaps M{};
    M.entries = m;
    M.max_parts = max_parts;
    M.atoms = dmap[0]; M.n_atoms = (int)atoms.size();
    M.cells = dmap[1]; M.n_cells = (int)cells.size();
    M.cell_tiles = dmap[2]; M.n_cell_tiles = (int)cell_tiles.size();
    M.atom_tiles = dmap[3]; M.n_atom_tiles = (int)atom_tiles.size();
    M.copy = dmap[4]; M.n_copy = (int)copy.size();
    M.save = dmap[5]; M.n_save = (int)save.size();
    for (int st = 0; st < 3; ++st) {
        M.style[st] = dmap[6 + st];
        M.n_style[st] = (int)style[st].size();
    }
    const RedoEntry* T = static_cast<const RedoEntry*>(B.d_redo);
    const hipStream_t st = B.stream;
    HIP_TRY(hipMemcpyAsync(B.d_redo, B.h_redo, bytes, hipMemcpyHostToDevice, st));
    int rc = launch_redo_resume(T, M, st);
    if (!rc) rc = launch_redo_begin_split(T, M, (float)dt, st);
    if (!rc) rc = launch_redo_save(T, M, st);
    if (!rc) rc = launch_redo_pbc(T, M, st);
    if (!rc) rc = launch_redo_bins(T, M, st);
    if (!rc) rc = launch_redo_rows(T, M, B.d_redo_flags, st);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(B.h_redo_flags, B.d_redo_flags, m * sizeof(int4), hipMemcpyDeviceToHost,
                           st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int q = 0; q < m; ++q) {
        rc = tmps_api::redo_batch_list(*B.sys[which[q]], B.h_redo_flags[q], E[q]);
        if (rc) return rc;
    }
    std::memcpy(B.h_redo, E.data(), entries_bytes);
    HIP_TRY(hipMemcpyAsync(B.d_redo, B.h_redo, entries_bytes, hipMemcpyHostToDevice, st));
    rc = launch_redo_list(T, M, st);
    if (!rc) rc = launch_redo_built(T, M, st);
    if (!rc) rc = launch_redo_encode_split(T, M, st);
    if (!rc) rc = launch_redo_restore(T, M, st);
    if (!rc) rc = launch_redo_force_split(T, M, (float)dt, st);
    if (!rc) rc = launch_redo_end(T, M, dt, st);
    if (!rc) rc = launch_redo_copy_vw_split(T, M, st);
    if (!rc) rc = launch_redo_rejoin(T, M, st);
    return rc;
}

int finish_member(System& s) {
    int rc = tmps_api::ctl_pull(s);
    if (rc) return rc;
    Ctl& H = *s.hctl;
    s.timestep = H.timestep;
    if (s.has_built) s.current_step = s.timestep;
    s.servo_step_count = H.servo.step_count;
    s.servo_nsteps = H.servo.nsteps;
    H.bpar = 0;
    H.bkick = 0;
    H.bdone = 0;
    H.remap.step = -1;
    return tmps_api::ctl_push(s);
}

int run(Batch& B, int nsteps, double dt, int run_pre, int autotune) {
    const int ninst = (int)B.sys.size();
    for (System* s : B.sys) {
        int rc = start_member(*s, nsteps, dt, run_pre);
        if (rc) return rc;
    }
    // tables: member buffers may have moved since the last run
    for (int k = 0; k < ninst; ++k) {
        int rc = fill_entry(B, k);
        if (rc) return rc;
    }
    {
        int rc = upload_table(B);
        if (rc) return rc;
    }
    int rc0 = write_close_table(B);
    if (rc0) return rc0;
    HIP_TRY(hipStreamSynchronize(B.stream));
    B.rebuilds = 0;
    B.redo_passes = 0;
    B.redo_ms = 0.0;
    // the autotuner: the first segments of the run try each configuration
    // once; its rate is the atom-steps that segment completed per second of
    // its launch (host redo work excluded), and the fastest runs the rest
    int probe = (autotune || B.config < 0) ? 0 : kNumConfigs;
    double best_rate = -1.0;
    int best = B.config >= 0 ? B.config : 2;
    std::vector<int> prev(ninst), redo;
    for (int k = 0; k < ninst; ++k) prev[k] = B.sys[k]->timestep;
    bool running = nsteps > 0;
    while (running) {
        const int ci = probe < kNumConfigs ? probe : best;
        int rc = prepare_segment(B, ci, dt);
        if (rc) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        rc = launch_segment(B, ci, dt);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(B.stream));
        const double sec =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        running = false;
        double atom_steps = 0.0;
        redo.clear();
        for (int k = 0; k < ninst; ++k) {
            const int4 st = B.h_status[k];
            atom_steps += (double)B.sys[k]->n * (st.x - prev[k]);
            prev[k] = st.x;
            if (st.y || st.z)
                redo.push_back(k);
            else if (!st.w)
                running = true;
        }
        const auto r0 = std::chrono::steady_clock::now();
        rc = redo_batched(B, redo, dt);
        if (rc) return rc;
        B.rebuilds += (int)redo.size();
        for (int k : redo) {
            // the rebuild may have moved the member's list buffers
            rc = fill_entry(B, k);
            if (rc) return rc;
            prev[k] = B.sys[k]->timestep;
            if (!B.sys[k]->hctl->bdone) running = true;
        }
        if (!redo.empty()) {
            rc = upload_table(B);
            if (rc) return rc;
            HIP_TRY(hipStreamSynchronize(B.stream));
            B.redo_passes += 1;
            B.redo_ms +=
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0)
                    .count();
        }
        if (probe < kNumConfigs) {
            const double rate = sec > 0.0 ? atom_steps / sec : 0.0;
            B.config_rate[ci] = rate;
            if (rate > best_rate) {
                best_rate = rate;
                best = ci;
            }
            if (++probe == kNumConfigs) B.config = best;
        }
    }
    if (B.config < 0 && best_rate > 0.0) B.config = best;
    for (System* s : B.sys) {
        int rc = finish_member(*s);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace

extern "C" {

int tmps_batch_create(tmps_system* const* systems, int n, tmps_batch** out) {
    if (!out) return tmps_fail(TMPS_ERR_ARG, "null output handle");
    *out = nullptr;
    if (n < 1 || !systems) return tmps_fail(TMPS_ERR_ARG, "an ensemble needs at least one system");
    Batch* B = new Batch();
    for (int k = 0; k < n; ++k) {
        System* s = reinterpret_cast<System*>(systems[k]);
        if (!s) {
            delete B;
            return tmps_fail(TMPS_ERR_ARG, "null ensemble member");
        }
        if (s->precision != TMPS_FP32_SPLIT) {
            delete B;
            return tmps_fail(TMPS_ERR_UNSUPPORTED, "an ensemble runs FP32-split systems");
        }
        if (s->n < 1 || !s->have_pair) {
            delete B;
            return tmps_fail(TMPS_ERR_ARG, "every ensemble member needs atoms and a pair style");
        }
        for (int q = 0; q < k; ++q)
            if (B->sys[q] == s) {
                delete B;
                return tmps_fail(TMPS_ERR_ARG, "a system appears twice in the ensemble");
            }
        B->sys.push_back(s);
        B->natoms.push_back(s->n);
    }
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return tmps_fail(TMPS_ERR_HIP, "hipStreamCreate failed");
    }
    int rc = build(*B);
    if (rc) {
        free_batch(*B);
        delete B;
        return rc;
    }
    *out = reinterpret_cast<tmps_batch*>(B);
    return 0;
}

int tmps_batch_destroy(tmps_batch* b) {
    if (!b) return 0;
    Batch* B = reinterpret_cast<Batch*>(b);
    free_batch(*B);
    delete B;
    return 0;
}

int tmps_batch_run(tmps_batch* b, int nsteps, double dt, int run_pre, int autotune) {
    if (!b) return tmps_fail(TMPS_ERR_ARG, "null ensemble");
    if (nsteps < 0) return tmps_fail(TMPS_ERR_ARG, "negative step count");
    Batch& B = *reinterpret_cast<Batch*>(b);
    for (System* m : B.sys) m->hctl_fresh = false;
    for (size_t k = 0; k < B.sys.size(); ++k)
        if (B.sys[k]->n != B.natoms[k] || !B.sys[k]->have_pair ||
            style_of(*B.sys[k]) != B.styles[k])
            return tmps_fail(TMPS_ERR_ARG,
                             "an ensemble member's atom count or pair style changed since the "
                             "ensemble was created; create it again");
    return run(B, nsteps, dt, run_pre, autotune);
}

int tmps_batch_tuning(tmps_batch* b, int* seg_len, int* block_order, double rates[6],
                      int* rebuilds, int* redo_passes, double* redo_ms) {
    if (!b) return tmps_fail(TMPS_ERR_ARG, "null ensemble");
    Batch& B = *reinterpret_cast<Batch*>(b);
    const int c = B.config >= 0 ? B.config : 2;
    if (seg_len) *seg_len = kConfigs[c].seg_len;
    if (block_order) *block_order = kConfigs[c].order;
    if (rates)
        for (int k = 0; k < kNumConfigs; ++k) rates[k] = B.config_rate[k];
    if (rebuilds) *rebuilds = B.rebuilds;
    if (redo_passes) *redo_passes = B.redo_passes;
    if (redo_ms) *redo_ms = B.redo_ms;
    return 0;
}

}  // extern "C"
