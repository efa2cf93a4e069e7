/*
 * tmps.h -- C ABI of libtmps_hip.so, the MI355X (gfx950) implementation of
 * the taichimps per-step DEM path.
 *
 * Every entry point returns an int status (TMPS_OK == 0, negative on error;
 * tmps_last_error() gives the message). Host arrays are borrowed for the
 * duration of the call; device memory is owned by the library. One caller
 * thread per system; all work runs on the system's own HIP stream.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to src/taichimps/ of the reference repository).
 */
#ifndef TMPS_H
#define TMPS_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define TMPS_OK 0
#define TMPS_ERR_ARG -1          /* ValueError in the reference */
#define TMPS_ERR_HIP -2          /* a HIP runtime call failed */
#define TMPS_ERR_OVERFLOW -3     /* neighbor.py:426-431 RuntimeError */
#define TMPS_ERR_BOX -4          /* deform_pressure.py:1167-1171 RuntimeError */
#define TMPS_ERR_CAPACITY -5     /* atom.py:74-77 ValueError */
#define TMPS_ERR_UNSUPPORTED -6  /* a setting this build does not implement */

/* ---- precision of the per-atom state --------------------------------- */
#define TMPS_FP64 0        /* all arithmetic in FP64 (the reference's f64 path) */
#define TMPS_FP32_SPLIT 1  /* FP64 cell origin + FP32 offset; FP32 contact and
                              integration; FP64 virial, pressure and servo */

/* ---- per-atom fields (AtomSystem, atom.py:36-53; NeighborList.x0) ---- */
enum tmps_field {
    TMPS_X = 0,      /* double[n][3] */
    TMPS_V = 1,      /* double[n][3] */
    TMPS_OMEGA = 2,  /* double[n][3] */
    TMPS_F = 3,      /* double[n][3] */
    TMPS_TORQUE = 4, /* double[n][3] */
    TMPS_RADIUS = 5, /* double[n]    */
    TMPS_RMASS = 6,  /* double[n]    */
    TMPS_TAG = 7,    /* int[n]       */
    TMPS_VIRIAL = 8, /* double[n][6] */
    TMPS_X0 = 9      /* double[n][3], read-only */
};

/* Pair-style coefficients after the reference's host-side derivations:
 * gran/hooke/history (GRANULAR/hooke_history.py:48-68), gran/hertz/history
 * (GRANULAR/hertz_history.py:36-55) and pair granular
 * (GRANULAR/granular.py:122-269). */
#define TMPS_PAIR_HOOKE_HISTORY 0
#define TMPS_PAIR_HERTZ_HISTORY 1
#define TMPS_PAIR_GRANULAR 2
typedef struct {
    int style;
    int limit_damping;
    double kn, kt, gamman, gammat, xmu; /* legacy styles; gammat 0 if dampflag 0 */
    int normal_model;     /* granular.py:50-52  hooke, hertz, hertz/material */
    int damping_model;    /* granular.py:54-59  none .. coeff_restitution */
    int tangential_model; /* granular.py:61-64  none .. mindlin */
    double g_kn, g_damp, g_kt, g_xt, g_mu;
} tmps_pair_params;

/* fix deform/pressure settings for the x/y/z entries
 * (EXTRA_FIX/deform_pressure.py:151-238, AxisSet :90-132). Style codes are
 * DEVICE_STYLE_CODE (:76-82), couple codes COUPLE_CODE (:83). */
typedef struct {
    int enabled;
    int nevery;
    int couple;    /* 0 none, 1 xyz, 2 xy, 3 yz, 4 xz */
    int remap;     /* 0 x, 1 v, 2 none */
    int normalize; /* normalize/pressure */
    double max_rate;
    int style[3];
    double rate[3], vel[3], scale[3], flo[3], fhi[3], dlo[3], dhi[3];
    double amp[3], tper[3], ptarget[3], pgain[3];
} tmps_servo_params;

/* Bookkeeping shared with the Python NeighborList / Simulation objects. */
typedef struct {
    int timestep;        /* Simulation.timestep */
    int has_built;       /* NeighborList.has_built */
    int build_count;     /* NeighborList.build_count */
    int last_build_step; /* NeighborList.last_build_step */
    int current_step;    /* NeighborList.current_step */
    int setup_done;      /* Simulation.setup_done */
    int npairs;          /* half-list pairs (NeighborList.npairs) */
    int nslots;          /* full-list slots (2 * npairs) */
    int ntouch;          /* touching full-list slots after the last force pass */
    int rebuilds;        /* rebuilds performed by the last tmps_run */
    int servo_step_count;
    int servo_nsteps;
} tmps_counters;

typedef struct tmps_system tmps_system;

/* ---- library --------------------------------------------------------- */
const char* tmps_version(void);
const char* tmps_last_error(void);
int tmps_device_name(char* buf, int len);

/* ---- system (AtomSystem.__init__ atom.py:30-53 together with the field
 *      allocations of NeighborList neighbor.py:23-120 and ContactHistory
 *      contact_history.py:37-80) ------------------------------------------ */
int tmps_create(int max_atoms, int max_neighbors_per_atom, int precision, tmps_system** out);
int tmps_destroy(tmps_system* s);
int tmps_set_nlocal(tmps_system* s, int nlocal);          /* AtomSystem.nlocal */
int tmps_set_field(tmps_system* s, int field, const void* host, int count);  /* Field.from_numpy */
int tmps_get_field(tmps_system* s, int field, void* host, int count);        /* Field.to_numpy */
/* The same two through device memory: `dev` is a caller's device buffer
 * (e.g. a torch tensor's data_ptr()), borrowed for the call, in
 * tmps_get_field's layout. Both are ordered on the system's stream and return
 * without waiting: order the caller's work after it (tmps_sync, or events on
 * a stream given by tmps_set_stream). TMPS_RADIUS and TMPS_RMASS are set
 * through host memory (they update host-side state), and TMPS_RADIUS is read
 * from the host's exact copy. No reference counterpart (the reference moves
 * fields through numpy, Field.from_numpy / to_numpy). */
int tmps_set_field_device(tmps_system* s, int field, const void* dev, int count);
int tmps_get_field_device(tmps_system* s, int field, void* dev, int count);
int tmps_sync(tmps_system* s);                                               /* ti.sync */

/* ---- Domain (domain.py) ---------------------------------------------- */
/* Domain._sync / set_box_and_h_rate, domain.py:235-307 */
int tmps_set_box(tmps_system* s, const double lo[3], const double hi[3], const int periodic[3],
                 const double h_rate[6], int vremap);
/* Domain.pull, domain.py:104-128 */
int tmps_get_box(tmps_system* s, double lo[3], double hi[3], double h_rate[6], int* vremap);
/* Domain.pbc, domain.py:573-584 */
int tmps_pbc(tmps_system* s);

/* ---- AtomSystem.clear_forces, atom.py:208-220 --------------------------- */
int tmps_clear_forces(tmps_system* s);

/* ---- NeighborList (neighbor.py) ---------------------------------------- */
/* skin and neigh_modify delay/every/check (neighbor.py:88-94), max_grid_cells
 * (:115) and max_neighbors_per_atom (:35-38, the half-list overflow bound) */
int tmps_neigh_config(tmps_system* s, double skin, int delay, int every, int check,
                      int max_grid_cells, int max_neighbors_per_atom);
/* check_displacement, neighbor.py:214-245: *moved = 1 if any |x-x0|^2 > skin^2/4 */
int tmps_neigh_check(tmps_system* s, int* moved);
/* NeighborList.decide, neighbor.py:371-403 */
int tmps_neigh_decide(tmps_system* s, int timestep, int have_timestep, int* rebuild);
/* NeighborList.build, neighbor.py:414-435 (grid, bins, list, x0, counters) */
int tmps_neigh_build(tmps_system* s);
/* grid_dim / cell_size, neighbor.py:185-189; npairs (half list) */
int tmps_neigh_info(tmps_system* s, int gdim[3], double csize[3], int* npairs);
/* pair_i / pair_j / ContactHistory.partner / shear in the reference's flat
 * half-list layout (neighbor.py:284-303, contact_history.py:53-59) */
int tmps_neigh_get_pairs(tmps_system* s, int* pair_i, int* pair_j, int* partner, double* shear,
                         int cap);
/* num_neighbors (half list) per atom, neighbor.py:45 */
int tmps_neigh_get_counts(tmps_system* s, int* num_neighbors, int n);
/* cell_of / cell_start / cell_particles, neighbor.py:116-120 */
int tmps_neigh_get_bins(tmps_system* s, int* cell_of, int* cell_start, int* cell_particles);

/* ---- ContactHistory (contact_history.py) ------------------------------- */
int tmps_history_reset(tmps_system* s);    /* reset, :82-89 */
int tmps_history_save(tmps_system* s);     /* save_state, :141-145 */
int tmps_history_restore(tmps_system* s);  /* restore_state, :147-154 */
/* seed the current list's history in the flat half-list layout
 * (partner tag or -1, shear), as ContactHistory.partner/shear.from_numpy */
int tmps_history_set_pairs(tmps_system* s, const int* partner, const double* shear, int npairs);
/* ContactHistory.load_restart, contact_history.py:176-214: seed the saved
 * history from a restart file, per atom (n == nlocal) a strictly ascending row
 * of partner indices nbr[start[i] .. start[i+1]) and their shear (3 doubles
 * each, in the lower-index-first convention). The next save is skipped (the
 * reference's setup_pending) and the following list build's restore carries
 * the seeded contacts onto the new list. */
int tmps_history_seed(tmps_system* s, int n, const int* start, const int* nbr,
                      const double* shear);

/* ---- pair styles (GRANULAR/base.py:25-38) ------------------------------- */
int tmps_pair_set(tmps_system* s, const tmps_pair_params* p);
/* GranularPair.compute: forces, torques and per-atom virial of every contact,
 * history update when shearupdate != 0. accumulate != 0 adds to the current
 * f/torque/virial (the reference's atomic_add after clear_forces). */
int tmps_pair_compute(tmps_system* s, double dt, int shearupdate, int accumulate);

/* ---- fixes ---------------------------------------------------------------- */
int tmps_nve_initial(tmps_system* s, double dt); /* FixNVESphere, nve_sphere.py:98-105 */
int tmps_nve_final(tmps_system* s, double dt);   /* nve_sphere.py:107-113 */
/* FixDampingCundall.post_force, damping_cundall.py:55-63 */
int tmps_cundall(tmps_system* s, double gamma_lin, double gamma_ang);
/* FixDeformPressure: settings, setup (:392-445), end_of_step (:1113-1219).
 * tensor_out (nullable) receives the pressure tensor the servo used. */
int tmps_servo_set(tmps_system* s, const tmps_servo_params* p);
int tmps_servo_setup(tmps_system* s, int nsteps_total, double dt);
int tmps_servo_end_of_step(tmps_system* s, double dt, double tensor_out[6]);

/* ---- fix wall/gran (EXTRA_FIX/wall_gran.py) and fix wall/gran/region
 *      (EXTRA_FIX/wall_gran_region.py) ------------------------------------
 * A wall in contact with the atoms of a system (FP64, or FP32-split with the
 * contact evaluated in FP64 on decoded positions), as a partner of zero
 * radius and infinite mass: a flat, axis-aligned plane (FixWallGran's
 * settings, wall_gran.py:57-111) or the curved side of a cylinder region
 * (FixWallGranRegion, wall_gran_region.py:45-112), after their host
 * derivations (gammat = 0 when dampflag is 0). A wall owns its per-atom shear
 * history (max_atoms of the system, zeroed at creation). */
#define TMPS_WALL_PLANE 0
#define TMPS_WALL_CYLINDER 1
typedef struct {
    int axis;          /* plane normal / cylinder axis: 0, 1 or 2 */
    int side;          /* plane: -1 wall below the atoms (normal +axis), +1 above;
                          cylinder: -1 atoms inside, +1 outside */
    int use_history;   /* the history styles (1) or the history-free ones (0) */
    int limit_damping;
    int shear_axis;    /* plane: in-plane wall velocity axis, -1 for none */
    double kn, kt, gamman, gammat, xmu;
    int geometry;      /* TMPS_WALL_PLANE or TMPS_WALL_CYLINDER */
    int hertz;         /* cylinder: hertz normal force and viscoelastic damping */
    int has_bounds;    /* cylinder: only atoms with axis_lo <= x[axis] <= axis_hi */
    int pad;
    double c1, c2, radius, axis_lo, axis_hi;  /* cylinder cross-section centre, radius */
} tmps_wall_params;
typedef struct tmps_wall tmps_wall;
int tmps_wall_create(tmps_system* s, const tmps_wall_params* p, tmps_wall** out);
int tmps_wall_destroy(tmps_wall* w);  /* also valid after its system is destroyed */
/* FixWallGran.reset_history, wall_gran.py:113-118 */
int tmps_wall_reset(tmps_wall* w);
/* FixWallGran.post_force_kernel, wall_gran.py:120-211, or
 * FixWallGranRegion.post_force_kernel, wall_gran_region.py:114-271: adds the
 * wall's force and torque to every atom's. A plane takes this step's wall
 * position and its normal and in-plane velocities (post_force, :213-236,
 * computes them); a cylinder ignores coord, vwall_n and vwall_s. */
int tmps_wall_post_force(tmps_wall* w, double dt, double coord, double vwall_n, double vwall_s,
                         int history_update);
/* FixWallGran.shear / touch: per-atom shear (3n, nullable) and touch flags
 * (n, nullable) */
int tmps_wall_get_history(tmps_wall* w, double* shear, int* touch, int n);
/* The walls a device-resident run (tmps_run) applies inside its one-kernel
 * step, as Simulation.run_gpu runs the fixes' post_force between the pair
 * force and final_integrate (simulation.py:236-247; wall_gran.py:213-236):
 * at most 2, before_cundall[k] set when wall k comes before fix
 * damping/cundall among the fixes, coord[k] / vwall_s[k] a plane's position
 * and in-plane velocity for the whole run (a wiggling plane is not fused),
 * dt the run's timestep. Setup calls them with history_update off. The walls
 * must outlive the setting; n = 0 clears it. */
int tmps_run_set_walls(tmps_system* s, int n, tmps_wall* const* walls, const int* before_cundall,
                       const double* coord, const double* vwall_s, double dt);

/* ---- computes (computes/thermo.py) --------------------------------------- */
/* compute_virial_kernel, thermo.py:62-109: sum of the per-atom virial
 * (+ m v_a v_b when keflag) as a pairwise tree over atoms */
int tmps_virial_reduce(tmps_system* s, int keflag, double out6[6]);
/* compute_ke_kernel, thermo.py:46-60 */
int tmps_ke(tmps_system* s, double* ke_trans, double* ke_rot);

/* ---- per-contact computes (computes/contact_atom.py, fabric.py,
 *      stress_atom.py) --------------------------------------------------------
 * Evaluated on the current positions and neighbor list (FP32-split systems
 * decode their positions to FP64 first). Sums over atoms or contacts are
 * pairwise trees, so they are deterministic. */
/* ComputeContactAtom.compute, contact_atom.py:19-64: per atom, the list
 * partners overlapping it (|d|^2 < (ri + rj)^2 and |d|^2 > 1e-18), counts
 * n ints (nullable); total (nullable): their sum, twice the contacts */
int tmps_contact_atom(tmps_system* s, int* counts, long long* total);
/* ComputeFabric.compute, fabric.py:26-97: phi = (sum over contacts of
 * n n^T) / Nc as [xx yy zz xy xz yz] (zeros without contacts), and Nc
 * (nullable) */
int tmps_fabric(tmps_system* s, double phi[6], int* ncontacts);
/* No counterpart in the reference: tmps_fabric that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same results as
 * tmps_fabric, same kernels; their scratch (the block partials and the sum)
 * is a fixed file-scope device array of 256 KiB + 64 B per GPU, held until
 * the process ends, so it takes at most 1,048,576 atoms (TMPS_ERR_CAPACITY
 * beyond). Calls that overlap on one GPU share that array and race. */
int tmps_fabric_noalloc(tmps_system* s, double phi[6], int* ncontacts);
/* ComputeStressAtom.compute / compute_macro, stress_atom.py:30-109: per-atom
 * stress -virial (minus m v_a v_b when keflag) into stress (n x 6,
 * nullable), and its sum over atoms into sum6 (nullable; compute_macro
 * divides it by the box volume) */
int tmps_stress_atom(tmps_system* s, int keflag, double* stress, double sum6[6]);

/* ---- Simulation (simulation.py) ------------------------------------------- */
/* Simulation.run(steps), simulation.py:320-332, with an nve/sphere
 * integrator, the configured pair style, optional damping/cundall
 * (cundall != 0) and optional deform/pressure servo (tmps_servo_set).
 * Runs device-resident: steps are replayed from captured HIP graphs and the
 * host synchronises once per segment (or on a neighbor rebuild). */
int tmps_run(tmps_system* s, int nsteps, double dt, int run_pre, int cundall, double gamma_lin,
             double gamma_ang);
/* The segment graphs a tmps_run of `nsteps` from the current step would
 * launch (with no rebuild request), captured now without running anything,
 * so that run pays no capture. Same arguments as tmps_run; call it after the
 * run's setup (a first tmps_run). No reference counterpart: Taichi compiles
 * its kernels on first use (simulation.py:320-332 runs them). */
int tmps_prepare_run(tmps_system* s, int nsteps, double dt, int cundall, double gamma_lin,
                     double gamma_ang);
int tmps_get_counters(tmps_system* s, tmps_counters* out);
int tmps_set_counters(tmps_system* s, const tmps_counters* in);
/* pressure tensor of the last servo update (virial+kinetic over volume) */
int tmps_last_pressure(tmps_system* s, double out6[6]);

/* ---- adjoint ---------------------------------------------------------------- */
/* No reference counterpart: the reference computes no gradients. A
 * hand-written reverse of the FP64 step (gran/hooke/history or
 * gran/hertz/history contacts, nve/sphere, deform/pressure pressure styles with
 * remap x and nevery 1) for calibration against measured stress.
 *
 * tmps_adjoint_forward runs Verlet::setup and `nsteps` steps as tmps_run does
 * (the fix setup must have been done by the caller), checkpointing the state
 * every `every` steps. It returns the final pressure tensor (xx yy zz xy xz yz,
 * virial plus kinetic over volume, as the servo saw it) and the coordination
 * number (touching contacts per atom, nullable).
 *
 * tmps_adjoint_backward takes dL/d(final pressure tensor) and returns
 * dL/d(kn, kt, gamman, gammat, xmu), dL/d(radius) per atom at fixed density
 * (n doubles) and, each nullable, dL/d(initial x, v, omega) (3n doubles each).
 * It replays each segment from its checkpoint and leaves the system in the
 * state the forward run ended in. */
int tmps_adjoint_forward(tmps_system* s, int nsteps, int every, double dt, double tensor_out[6],
                         double* coordination);
int tmps_adjoint_backward(tmps_system* s, const double dL_dtensor[6], double grad_coeffs[5],
                          double* grad_radius, double* grad_x, double* grad_v, double* grad_omega);

/* ---- ensembles ---------------------------------------------------------------- */
/* No reference counterpart: the reference runs one system per Simulation.
 * An ensemble advances many independent FP32-split systems (each with its own
 * box, atoms, pair style and coefficients, servo) together: per step one
 * launch per pair style covers every member's atoms and one launch closes
 * every member's step, captured as HIP graph segments. A member whose skin
 * check asks for a rebuild stops; after the segment the steps of all such
 * members are redone with their list rebuilds, each kernel launched once for
 * all of them, and they rejoin; the others do not wait. Every member's
 * results are bitwise those tmps_run gives it alone.
 *
 * tmps_batch_create takes `n` distinct systems, each FP32-split, with atoms,
 * a pair style (hooke/history, hertz/history or granular) and, optionally,
 * deform/pressure; no fix damping/cundall. The systems stay owned by the
 * caller and must outlive the ensemble; their atom counts must not change.
 * tmps_batch_run is tmps_run(nsteps, dt, run_pre) on every member. With
 * `autotune` (and always on the first run) its first segments time each
 * launch configuration (segment length 8/16/32, members' blocks in sequence
 * or grouped per XCD); tmps_batch_tuning reports the chosen configuration,
 * each configuration's measured rate (atom-steps/s, 6 entries), the number of
 * member steps the last run redid after a rebuild request, the number of
 * redo passes (one per segment with such requests) and their total wall ms. */
typedef struct tmps_batch tmps_batch;
int tmps_batch_create(tmps_system* const* systems, int n, tmps_batch** out);
int tmps_batch_destroy(tmps_batch* b);
int tmps_batch_run(tmps_batch* b, int nsteps, double dt, int run_pre, int autotune);
int tmps_batch_tuning(tmps_batch* b, int* seg_len, int* block_order, double rates[6],
                      int* rebuilds, int* redo_passes, double* redo_ms);

/* ---- stream ---------------------------------------------------------------- */
/* The HIP stream (a hipStream_t) the system's work runs on: a caller's
 * stream, borrowed (it must outlive its use), or NULL for the system's own
 * stream. Work already queued on the previous stream completes first. Replaces
 * the one implicit stream of the reference's Taichi kernels (every kernel of
 * simulation.py:205-332 runs on it); a caller's stream lets the path join a
 * caller's ordering (PyTorch's current stream, HIP events around a run). */
int tmps_set_stream(tmps_system* s, void* stream);
int tmps_get_stream(tmps_system* s, void** stream);

/* ---- measurement ----------------------------------------------------------- */
/* How tmps_run issues a segment's kernels: replayed from a captured HIP graph
 * (the default) or launched one by one on the system's stream. No reference
 * counterpart; the reference leaves kernel dispatch to Taichi. */
#define TMPS_LAUNCH_GRAPH 0
#define TMPS_LAUNCH_DIRECT 1
int tmps_set_launch_mode(tmps_system* s, int mode);
/* Average duration (ms) of each kernel class of tmps_run while profiling is
 * enabled (direct launches, HIP events on the system's stream):
 * [0] a segment's first half-kick, [1] the step pass kernel (its own start to
 * its own stop), [2] a segment's tail, [3] the step's close kernel (tree and
 * servo, measured from the pass's stop, so it includes the launch gap). */
#define TMPS_PROF_CLASSES 4
int tmps_set_profiling(tmps_system* s, int enabled);
int tmps_kernel_times(tmps_system* s, double out_ms[TMPS_PROF_CLASSES],
                      int counts[TMPS_PROF_CLASSES]);

#ifdef __cplusplus
}
#endif

#endif /* TMPS_H */
