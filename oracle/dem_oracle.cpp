// CPU FP64 restatement of the taichimps per-step DEM path -- TEST INFRASTRUCTURE.
//
// This file is the parity oracle. Only tests, the smoke check and the CPU
// baseline leg of the benchmark may load it, and only as the checker; the
// product path (taichimps_amd) never calls it.
//
// Every function restates one reference function, cited as file:line relative
// to the reference repository (src/taichimps/...). Build flags are fixed at
// -O2 -ffp-contract=off so that every expression rounds exactly as written.
//
// Deterministic choices where the reference leaves the order open (its Taichi
// kernels accumulate with float atomics and fill bins with atomic slots):
//   * bins list their members in ascending atom index;
//   * each atom's half-list row (j > i) is sorted ascending by j, so the pair
//     list is (i ascending, j ascending) and every per-atom sum of f, torque
//     and virial runs over partners in ascending index order;
//   * the global virial(+kinetic) tensor is a pairwise sum: per-atom leaves,
//     zero-padded to a power of two, added in adjacent pairs level by level.
//     (The reference's own order, computes/thermo.py:62-109, is a strided
//     chunk sum; orc_virial(mode=1) keeps that order for comparison.)
//
// Parity pins for each function are listed in the header of tests/test_oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

// liboracle_omp.so (ORC_OMP, -fopenmp) is the same restatement with its
// per-atom and per-pair loops on host threads and the pair tally's updates of
// shared atoms made atomic, as the reference's Taichi kernels do: the CPU
// baseline of the benchmark. Its sums depend on thread timing, so only the
// serial liboracle.so checks parity.
#ifdef ORC_OMP
#include <omp.h>
#define ORC_PAR _Pragma("omp parallel for schedule(static)")
#define ORC_ATOMIC _Pragma("omp atomic")
#else
#define ORC_PAR
#define ORC_ATOMIC
#endif

namespace {

enum { STYLE_HOOKE_HISTORY = 0, STYLE_HERTZ_HISTORY = 1, STYLE_GRANULAR = 2 };
// granular.py:50-64
enum { NORMAL_HOOKE = 0, NORMAL_HERTZ = 1, NORMAL_HERTZ_MATERIAL = 2 };
enum { DAMP_NONE = 0, DAMP_VELOCITY = 1, DAMP_MASS_VELOCITY = 2, DAMP_VISCOELASTIC = 3,
       DAMP_TSUJI = 4, DAMP_COEFF_RESTITUTION = 5 };
enum { TANG_NONE = 0, TANG_LINEAR_NOHISTORY = 1, TANG_LINEAR_HISTORY = 2, TANG_MINDLIN = 3 };
// deform_pressure.py:76-77
enum { S_NONE = 0, S_ERATE = 1, S_TRATE = 2, S_VEL = 3, S_WIGGLE = 4, S_FINAL = 5,
       S_DELTA = 6, S_SCALE = 7, S_PRESSURE = 8, S_PMEAN = 9 };

const double GRAN_EPSILON = 1e-10;  // granular.py:87

}  // namespace

extern "C" {

// Mirrors the pair-style constructor arguments (hooke_history.py:48-68,
// hertz_history.py:36-55, granular.py:122-205) after their derivations.
typedef struct {
    int style;
    int limit_damping;
    double kn, kt, gamman, gammat, xmu;   // gammat already 0 when dampflag == 0
    int normal_model, damping_model, tangential_model;
    double g_kn, g_damp, g_kt, g_xt, g_mu;
} orc_pair_params;

// Mirrors FixDeformPressure settings for the x/y/z entries
// (deform_pressure.py:151-238 and AxisSet :90-132).
typedef struct {
    int enabled;
    int nevery;
    int couple;      // 0 none, 1 xyz, 2 xy, 3 yz, 4 xz  (COUPLE_CODE :83)
    int remap;       // 0 x, 1 v, 2 none
    int normalize;
    double max_rate;
    int style[3];
    double rate[3], vel[3], scale[3], flo[3], fhi[3], dlo[3], dhi[3];
    double amp[3], tper[3], ptarget[3], pgain[3];
} orc_servo_params;

}  // extern "C"

namespace {

// Mirrors FixWallGran's constructor settings (wall_gran.py:57-111) and, for
// geometry 1, FixWallGranRegion's cylinder (wall_gran_region.py:45-112).
extern "C" {
typedef struct {
    int axis, side, use_history, limit_damping, shear_axis;
    double kn, kt, gamman, gammat, xmu;  // gammat already 0 when dampflag is 0
    int geometry, hertz, has_bounds, pad;
    double c1, c2, radius, axis_lo, axis_hi;
} orc_wall_params;
}

// One flat wall: settings, its motion and its per-atom shear history.
struct OrcWall {
    orc_wall_params p{};
    double coord = 0.0, shear_vel = 0.0, amp = 0.0, period = 0.0, elapsed = 0.0;
    std::vector<double> shear;  // [n][3]
    std::vector<int> touch;
};

struct Oracle {
    int n = 0;
    int max_nb = 64;
    // AtomSystem fields, atom.py:36-53 (x, v, omega, f, torque AoS [n][3])
    std::vector<double> x, v, omega, f, torque, radius, rmass, virial;
    std::vector<int> tag;

    // Domain device state, domain.py:55-68 (orthogonal box)
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, prd[3] = {1, 1, 1};
    int per[3] = {1, 1, 1};
    double h_rate[6] = {0, 0, 0, 0, 0, 0};
    int vremap = 0;

    // NeighborList, neighbor.py:23-120
    double skin = 0.001;
    int delay = 0, every = 1, check = 1;
    int max_grid_cells = 200000;
    int has_built = 0, build_count = 0, last_build_step = -1, current_step = 0;
    int gdim[3] = {1, 1, 1};
    double csize[3] = {1, 1, 1};
    std::vector<int> cell_of, cell_start, cell_particles;
    std::vector<int> num_neighbors, neighbors;  // [n][max_nb]
    std::vector<int> pair_start, pair_i, pair_j;
    int npairs = 0;
    std::vector<double> x0;

    // ContactHistory, contact_history.py:47-78
    std::vector<int> partner;
    std::vector<double> shear;
    std::vector<int> save_count, save_tag;
    std::vector<double> save_shear;

    orc_pair_params pair{};
    int have_pair = 0;

    // FixDampingCundall, damping_cundall.py:15-25
    int cundall = 0;
    double gamma_lin = 0.0, gamma_ang = 0.0;

    // FixDeformPressure runtime state
    orc_servo_params servo{};
    double lo_start[3] = {0, 0, 0}, hi_start[3] = {0, 0, 0}, cum[3] = {0, 0, 0};
    int step_count = 0, nsteps = 0, nsteps_total = 0;

    int timestep = 0;
    int setup_done = 0;
    int rebuilds_in_run = 0;
    double last_tensor[6] = {0, 0, 0, 0, 0, 0};

    // FixWallGran fixes, applied after damping/cundall in post_force
    std::vector<OrcWall> walls;
};

// ------------------------------------------------------- sin, exp
// The servo's wiggle and trate styles call sin and exp (deform_pressure.py:
// 646, 653; math.sin / math.exp on the host path, ti.sin / ti.exp on the
// device). Host and device libraries may round those differently in the last
// bit, so the restatement evaluates them by one fixed method that the device
// build follows too: Cody-Waite reduction (pi/2 in three parts, ln 2 in two,
// leading parts of 33 and 32 bits) and Taylor polynomials (degree 17/18 on
// |r| <= pi/4, degree 15 on |r| <= ln(2)/2) in Horner form. Within two ulps
// of the libm functions (tests/test_oracle_servo.py).

double det_sin_poly(double r) {
    const double z = r * r;
    double p = -1.0 / 355687428096000.0;        // -1/17!
    p = p * z + 1.0 / 1307674368000.0;          //  1/15!
    p = p * z - 1.0 / 6227020800.0;             // -1/13!
    p = p * z + 1.0 / 39916800.0;               //  1/11!
    p = p * z - 1.0 / 362880.0;                 // -1/9!
    p = p * z + 1.0 / 5040.0;                   //  1/7!
    p = p * z - 1.0 / 120.0;                    // -1/5!
    p = p * z + 1.0 / 6.0;                      //  1/3!
    return r - (r * z) * p;
}

double det_cos_poly(double r) {
    const double z = r * r;
    double p = 1.0 / 6402373705728000.0;        //  1/18!
    p = p * z - 1.0 / 20922789888000.0;         // -1/16!
    p = p * z + 1.0 / 87178291200.0;            //  1/14!
    p = p * z - 1.0 / 479001600.0;              // -1/12!
    p = p * z + 1.0 / 3628800.0;                //  1/10!
    p = p * z - 1.0 / 40320.0;                  // -1/8!
    p = p * z + 1.0 / 720.0;                    //  1/6!
    p = p * z - 1.0 / 24.0;                     // -1/4!
    p = p * z + 0.5;
    return 1.0 - z * p;
}

double det_sin(double x) {
    const double pio2_1 = 1.5707963267341256, pio2_2 = 6.077100506303966e-11,
                 pio2_3 = 2.0222662487959506e-21;
    const double k = std::rint(x * (2.0 / M_PI));
    const double r = ((x - k * pio2_1) - k * pio2_2) - k * pio2_3;
    switch ((long long)k & 3) {
        case 0: return det_sin_poly(r);
        case 1: return det_cos_poly(r);
        case 2: return -det_sin_poly(r);
        default: return -det_cos_poly(r);
    }
}

double det_exp(double x) {
    const double ln2_1 = 0.6931471806019545, ln2_2 = -4.2009150726810846e-11;
    const double k = std::rint(x * (1.0 / M_LN2));
    const double r = (x - k * ln2_1) - k * ln2_2;
    // 1/n! for n = 15 down to 2, then 1, 1
    const double c[14] = {1.0 / 1307674368000.0, 1.0 / 87178291200.0, 1.0 / 6227020800.0,
                          1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                          1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0,
                          1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5};
    double p = c[0];
    for (int m = 1; m < 14; ++m) p = p * r + c[m];
    p = p * r + 1.0;
    p = p * r + 1.0;
    return std::ldexp(p, (int)k);
}

// ---------------------------------------------------------------- Domain

// Domain.minimum_image, orthogonal branch, domain.py:481-486
inline void minimum_image(const Oracle& s, double d[3]) {
    for (int k = 0; k < 3; ++k)
        if (s.per[k] == 1) d[k] -= s.prd[k] * std::round(d[k] / s.prd[k]);
}

// Domain.minimum_image_and_vshift + image_offsets + apply_image_offsets,
// orthogonal branch, domain.py:388-462
inline void minimum_image_and_vshift(const Oracle& s, double d[3], double dv[3]) {
    double nimg[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k)
        if (s.per[k] == 1) nimg[k] = std::round(d[k] / s.prd[k]);
    dv[0] = dv[1] = dv[2] = 0.0;
    if (s.vremap != 0) {
        const double* h = s.h_rate;
        dv[0] = nimg[0] * h[0] + nimg[1] * h[5] + nimg[2] * h[4];
        dv[1] = nimg[1] * h[1] + nimg[2] * h[3];
        dv[2] = nimg[2] * h[2];
    }
    for (int k = 0; k < 3; ++k) d[k] -= nimg[k] * s.prd[k];
}

// Domain.pbc_wrap orthogonal branch, domain.py:513-528, and
// _pbc_vremap_kernel / image_count / vremap_delta, domain.py:502-571
void pbc(Oracle& s) {
    for (int i = 0; i < s.n; ++i) {
        double* xi = &s.x[3 * i];
        if (s.vremap) {
            double nimg[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k)
                if (s.per[k] == 1) nimg[k] = std::floor((xi[k] - s.lo[k]) / s.prd[k]);
            if (nimg[0] != 0.0 || nimg[1] != 0.0 || nimg[2] != 0.0) {
                const double* h = s.h_rate;
                double dvx = -(nimg[0] * h[0] + nimg[1] * h[5] + nimg[2] * h[4]);
                double dvy = -(nimg[1] * h[1] + nimg[2] * h[3]);
                double dvz = -(nimg[2] * h[2]);
                s.v[3 * i + 0] += dvx;
                s.v[3 * i + 1] += dvy;
                s.v[3 * i + 2] += dvz;
            }
        }
        for (int k = 0; k < 3; ++k) {
            if (s.per[k] == 1) {
                double sl = (xi[k] - s.lo[k]) / s.prd[k];
                xi[k] = s.lo[k] + (sl - std::floor(sl)) * s.prd[k];
            }
        }
    }
}

// ---------------------------------------------------------- NeighborList

// NeighborList.setup_grid, neighbor.py:143-189 (orthogonal widths :122-141)
void setup_grid(Oracle& s, double max_cutoff) {
    double cut = max_cutoff + s.skin;
    if (cut <= 0) cut = 1.0;
    // perpendicular_widths: vol / |b x c| etc. with zero tilt
    const double xprd = s.prd[0], yprd = s.prd[1], zprd = s.prd[2];
    const double vol = xprd * yprd * zprd;
    double c0 = yprd * zprd - 0.0 * 0.0;   // |b x c| = b1*c2 - b2*c1
    double c1 = 0.0 * xprd - 0.0 * zprd;   // |c x a| component
    double c2 = xprd * yprd - 0.0 * 0.0;   // |a x b| = a0*b1 - a1*b0
    (void)c1;
    double w0 = vol / std::sqrt(c0 * c0);
    // |c x a|: c = (0,0,zprd), a = (xprd,0,0): c x a = (0, zprd*xprd, 0)
    double ca = zprd * xprd - 0.0 * 0.0;
    double w1 = vol / std::sqrt(ca * ca);
    double w2 = vol / std::sqrt(c2 * c2);
    int gx = std::max(1, (int)std::floor(w0 / cut));
    int gy = std::max(1, (int)std::floor(w1 / cut));
    int gz = std::max(1, (int)std::floor(w2 / cut));
    if (s.per[0] == 1 && gx == 2) gx = 1;
    if (s.per[1] == 1 && gy == 2) gy = 1;
    if (s.per[2] == 1 && gz == 2) gz = 1;
    long long total = (long long)gx * gy * gz;
    if (total > s.max_grid_cells) {
        double factor = std::pow((double)s.max_grid_cells / (double)total, 1.0 / 3.0);
        gx = std::max(1, (int)(gx * factor));
        gy = std::max(1, (int)(gy * factor));
        gz = std::max(1, (int)(gz * factor));
        // deviation from neighbor.py:178-184: the 2-bin collapse of a periodic
        // axis is applied again after the scaling, which can bring an axis
        // back to 2 bins (every candidate there would otherwise be listed
        // twice)
        if (s.per[0] == 1 && gx == 2) gx = 1;
        if (s.per[1] == 1 && gy == 2) gy = 1;
        if (s.per[2] == 1 && gz == 2) gz = 1;
    }
    s.gdim[0] = gx; s.gdim[1] = gy; s.gdim[2] = gz;
    s.csize[0] = 1.0 / gx; s.csize[1] = 1.0 / gy; s.csize[2] = 1.0 / gz;
}

// NeighborList.get_cell_coord, neighbor.py:191-206 (lamda_of, domain.py:556-558,
// to_lamda :360-375 with zero tilt)
inline void cell_coord(const Oracle& s, const double* p, int c[3]) {
    double d0 = p[0] - s.lo[0], d1 = p[1] - s.lo[1], d2 = p[2] - s.lo[2];
    double lam[3];
    lam[0] = d0 / s.prd[0] + 0.0 * d1 + 0.0 * d2;
    lam[1] = d1 / s.prd[1] + 0.0 * d2;
    lam[2] = d2 / s.prd[2];
    for (int k = 0; k < 3; ++k) {
        int ck = (int)std::floor(lam[k] / s.csize[k]);
        c[k] = std::max(0, std::min(ck, s.gdim[k] - 1));
    }
}

// NeighborList.build_grid, neighbor.py:247-282 (serial: ascending order in bins)
void build_grid(Oracle& s) {
    const int gx = s.gdim[0], gy = s.gdim[1], gz = s.gdim[2];
    const int ncell = gx * gy * gz;
    std::vector<int> count(ncell, 0);
    s.cell_of.assign(s.n, 0);
    for (int i = 0; i < s.n; ++i) {
        int c[3];
        cell_coord(s, &s.x[3 * i], c);
        int idx = c[0] + c[1] * gx + c[2] * gx * gy;
        s.cell_of[i] = idx;
        count[idx]++;
    }
    s.cell_start.assign(ncell + 1, 0);
    for (int c = 0; c < ncell; ++c) s.cell_start[c + 1] = s.cell_start[c] + count[c];
    std::vector<int> fill(ncell, 0);
    s.cell_particles.assign(s.n, 0);
    for (int i = 0; i < s.n; ++i) {
        int c = s.cell_of[i];
        s.cell_particles[s.cell_start[c] + fill[c]++] = i;
    }
}

// NeighborList.build_neighbor_list, neighbor.py:305-369; rows sorted ascending.
int build_neighbor_list(Oracle& s) {
    const int gx = s.gdim[0], gy = s.gdim[1], gz = s.gdim[2];
    const int gxy = gx * gy;
    const int px = s.per[0], py = s.per[1], pz = s.per[2];
    int overflow = 0;
    s.num_neighbors.assign(s.n, 0);
    s.neighbors.assign((size_t)s.n * s.max_nb, 0);
#ifdef ORC_OMP
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
#endif
    for (int i = 0; i < s.n; ++i) {
        const double* pi = &s.x[3 * i];
        const double ri = s.radius[i];
        int c[3];
        cell_coord(s, pi, c);
        int count = 0;
        int* row = &s.neighbors[(size_t)i * s.max_nb];
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    int nx = c[0] + dx, ny = c[1] + dy, nz = c[2] + dz;
                    bool skip = (px == 1 && gx == 1 && dx != 0) ||
                                (py == 1 && gy == 1 && dy != 0) ||
                                (pz == 1 && gz == 1 && dz != 0);
                    if (px == 1) nx = (nx % gx + gx) % gx;
                    if (py == 1) ny = (ny % gy + gy) % gy;
                    if (pz == 1) nz = (nz % gz + gz) % gz;
                    if (skip || nx < 0 || nx >= gx || ny < 0 || ny >= gy || nz < 0 || nz >= gz)
                        continue;
                    int cidx = nx + ny * gx + nz * gxy;
                    for (int p = s.cell_start[cidx]; p < s.cell_start[cidx + 1]; ++p) {
                        int j = s.cell_particles[p];
                        if (j > i) {
                            const double* pj = &s.x[3 * j];
                            double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                            minimum_image(s, d);
                            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                            double rad_sum = ri + s.radius[j] + s.skin;
                            if (rsq < rad_sum * rad_sum) {
                                if (count < s.max_nb) row[count++] = j;
                                else overflow |= 1;
                            }
                        }
                    }
                }
        std::sort(row, row + count);
        s.num_neighbors[i] = count;
    }
    return overflow;
}

// NeighborList.build_pair_list, neighbor.py:284-303
void build_pair_list(Oracle& s) {
    s.pair_start.assign(s.n + 1, 0);
    for (int i = 0; i < s.n; ++i) s.pair_start[i + 1] = s.pair_start[i] + s.num_neighbors[i];
    s.npairs = s.pair_start[s.n];
    s.pair_i.assign(s.npairs, 0);
    s.pair_j.assign(s.npairs, 0);
    for (int i = 0; i < s.n; ++i)
        for (int k = 0; k < s.num_neighbors[i]; ++k) {
            s.pair_i[s.pair_start[i] + k] = i;
            s.pair_j[s.pair_start[i] + k] = s.neighbors[(size_t)i * s.max_nb + k];
        }
}

// NeighborList.build, neighbor.py:414-435. Returns 1 on overflow.
int nlist_build(Oracle& s) {
    if (s.n == 0) return 0;
    double max_r = s.radius[0];
    for (int i = 1; i < s.n; ++i) max_r = std::max(max_r, s.radius[i]);
    setup_grid(s, 2.0 * max_r);
    build_grid(s);
    int overflow = build_neighbor_list(s);
    build_pair_list(s);
    if (overflow) return 1;
    s.x0 = s.x;  // store_x0 :208-212
    s.has_built = 1;
    s.build_count += 1;
    s.last_build_step = s.current_step;
    return 0;
}

// NeighborList.check_displacement + decide, neighbor.py:214-245,371-403
bool nlist_decide(Oracle& s, int timestep) {
    if (!s.has_built) return true;
    s.current_step = timestep;
    int since = s.current_step - s.last_build_step;
    if (since < s.delay) return false;
    if (since % s.every != 0) return false;
    if (!s.check) return true;
    if (s.n == 0) return false;
    const double trig = 0.25 * s.skin * s.skin;
    int need = 0;
#ifdef ORC_OMP
#pragma omp parallel for schedule(static) reduction(| : need)
#endif
    for (int i = 0; i < s.n; ++i) {
        double d0 = s.x[3 * i] - s.x0[3 * i];
        double d1 = s.x[3 * i + 1] - s.x0[3 * i + 1];
        double d2 = s.x[3 * i + 2] - s.x0[3 * i + 2];
        if (d0 * d0 + d1 * d1 + d2 * d2 > trig) need |= 1;
    }
    return need != 0;
}

// --------------------------------------------------------- ContactHistory

void history_reset(Oracle& s, size_t cap) {
    s.partner.assign(cap, -1);
    s.shear.assign(cap * 3, 0.0);
}

// save_state_kernel, contact_history.py:91-110 (serial order)
void history_save(Oracle& s) {
    s.save_count.assign(s.n, 0);
    s.save_tag.assign((size_t)s.n * s.max_nb, 0);
    s.save_shear.assign((size_t)s.n * s.max_nb * 3, 0.0);
    for (int nc = 0; nc < s.npairs; ++nc) {
        int ptag = s.partner[nc];
        if (ptag >= 0) {
            int i = s.pair_i[nc];
            int m = s.save_count[i]++;
            s.save_tag[(size_t)i * s.max_nb + m] = ptag;
            for (int k = 0; k < 3; ++k)
                s.save_shear[((size_t)i * s.max_nb + m) * 3 + k] = s.shear[(size_t)nc * 3 + k];
        }
    }
}

// restore_state_kernel, contact_history.py:112-139
void history_restore(Oracle& s) {
    if ((int)s.partner.size() < s.npairs) history_reset(s, s.npairs);
    for (int nc = 0; nc < s.npairs; ++nc) {
        int i = s.pair_i[nc];
        int jtag = s.tag[s.pair_j[nc]];
        int found = -1;
        for (int m = 0; m < s.save_count[i]; ++m)
            if (s.save_tag[(size_t)i * s.max_nb + m] == jtag) { found = m; break; }
        if (found >= 0) {
            s.partner[nc] = jtag;
            for (int k = 0; k < 3; ++k)
                s.shear[(size_t)nc * 3 + k] = s.save_shear[((size_t)i * s.max_nb + found) * 3 + k];
        } else {
            s.partner[nc] = -1;
            for (int k = 0; k < 3; ++k) s.shear[(size_t)nc * 3 + k] = 0.0;
        }
    }
}

// ----------------------------------------------------------- pair styles

inline void cross(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
inline double norm3(const double a[3]) { return std::sqrt(dot(a, a)); }

// Accumulate a pair's force F on i (and -F on j), the torques and the
// pairwise virial (Pair::ev_tally_xyz form).
inline void tally(Oracle& s, int i, int j, const double dpos[3], const double F[3],
                  const double ti[3], const double tj[3]) {
    for (int k = 0; k < 3; ++k) {
        ORC_ATOMIC s.f[3 * i + k] += F[k];
        ORC_ATOMIC s.f[3 * j + k] += -F[k];
        ORC_ATOMIC s.torque[3 * i + k] += ti[k];
        ORC_ATOMIC s.torque[3 * j + k] += tj[k];
    }
    double vir[6] = {0.5 * (dpos[0] * F[0]), 0.5 * (dpos[1] * F[1]), 0.5 * (dpos[2] * F[2]),
                     0.5 * (dpos[0] * F[1]), 0.5 * (dpos[0] * F[2]), 0.5 * (dpos[1] * F[2])};
    for (int k = 0; k < 6; ++k) {
        ORC_ATOMIC s.virial[6 * i + k] += vir[k];
        ORC_ATOMIC s.virial[6 * j + k] += vir[k];
    }
}

// GranHookeHistory.compute_kernel (hooke_history.py:70-195) and
// GranHertzHistory.compute_kernel (hertz_history.py:57-171).
void pair_legacy_history(Oracle& s, double dt, int shearupdate, bool hertz) {
    const orc_pair_params& P = s.pair;
    ORC_PAR
    for (int nc = 0; nc < s.npairs; ++nc) {
        const int i = s.pair_i[nc], j = s.pair_j[nc];
        const double* xi = &s.x[3 * i];
        const double* xj = &s.x[3 * j];
        const double* vi = &s.v[3 * i];
        const double* vj = &s.v[3 * j];
        const double* oi = &s.omega[3 * i];
        const double* oj = &s.omega[3 * j];
        const double ri = s.radius[i], rj = s.radius[j];
        const double mi = s.rmass[i], mj = s.rmass[j];
        double dpos[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        double dvj[3];
        minimum_image_and_vshift(s, dpos, dvj);
        const double rsq = dot(dpos, dpos);
        const double radsum = ri + rj;
        double* sh = &s.shear[(size_t)nc * 3];
        if (rsq < radsum * radsum && rsq > 1e-28) {
            const double r = std::sqrt(rsq);
            const double delta = radsum - r;
            double polyhertz = 1.0;
            if (hertz) {
                const double reff = (ri * rj) / (ri + rj);
                polyhertz = std::sqrt(reff * delta);
            }
            const double meff = (mi * mj) / (mi + mj);
            const double rsqinv = 1.0 / rsq;
            const double rinv = 1.0 / r;
            double vrel[3];
            for (int k = 0; k < 3; ++k) vrel[k] = vi[k] - (vj[k] + dvj[k]);
            const double vnnr = dot(vrel, dpos);
            double vt[3], wr[3], vtr[3], dxw[3];
            for (int k = 0; k < 3; ++k) {
                double vn = dpos[k] * (vnnr * rsqinv);
                vt[k] = vrel[k] - vn;
                wr[k] = (ri * oi[k] + rj * oj[k]) * rinv;
            }
            cross(dpos, wr, dxw);
            for (int k = 0; k < 3; ++k) vtr[k] = vt[k] + dxw[k];
            const double damp = meff * P.gamman * vnnr * rsqinv;
            double ccel;
            if (hertz) ccel = (P.kn * delta * rinv - damp) * polyhertz;
            else ccel = P.kn * delta * rinv - damp;
            if (P.limit_damping == 1 && ccel < 0.0) ccel = 0.0;

            const int jtag = s.tag[j];
            double shear[3] = {sh[0], sh[1], sh[2]};
            if (s.partner[nc] != jtag) {
                shear[0] = shear[1] = shear[2] = 0.0;
                s.partner[nc] = jtag;
            }
            const double gammat_eff = meff * P.gammat;
            if (shearupdate != 0)
                for (int k = 0; k < 3; ++k) shear[k] = shear[k] + vtr[k] * dt;
            const double shrmag = norm3(shear);
            if (shearupdate != 0) {
                const double rsht = dot(shear, dpos) * rsqinv;
                for (int k = 0; k < 3; ++k) shear[k] = shear[k] - rsht * dpos[k];
            }
            double fs[3];
            if (hertz) {
                for (int k = 0; k < 3; ++k)
                    fs[k] = -polyhertz * (P.kt * shear[k] + gammat_eff * vtr[k]);
            } else {
                for (int k = 0; k < 3; ++k) fs[k] = -(P.kt * shear[k] + gammat_eff * vtr[k]);
            }
            const double fs_mag = norm3(fs);
            const double fn_coulomb = P.xmu * std::fabs(ccel * r);
            if (fs_mag > fn_coulomb) {
                if (shrmag != 0.0) {
                    const double ratio = fn_coulomb / fs_mag;
                    for (int k = 0; k < 3; ++k) {
                        const double damp_corr = gammat_eff * vtr[k] / P.kt;
                        shear[k] = ratio * (shear[k] + damp_corr) - damp_corr;
                        fs[k] = ratio * fs[k];
                    }
                } else {
                    fs[0] = fs[1] = fs[2] = 0.0;
                }
            }
            sh[0] = shear[0]; sh[1] = shear[1]; sh[2] = shear[2];
            double F[3];
            for (int k = 0; k < 3; ++k) F[k] = dpos[k] * ccel + fs[k];
            double c[3], tor[3], ti[3], tj[3];
            cross(dpos, fs, c);
            for (int k = 0; k < 3; ++k) {
                tor[k] = rinv * c[k];
                ti[k] = -ri * tor[k];
                tj[k] = -rj * tor[k];
            }
            tally(s, i, j, dpos, F, ti, tj);
        } else {
            s.partner[nc] = -1;
            sh[0] = sh[1] = sh[2] = 0.0;
        }
    }
}

// PairGranular._rotate_rescale, granular.py:271-280
inline void rotate_rescale(double h[3], const double n[3]) {
    const double shrmag = norm3(h);
    const double hn = dot(h, n);
    double out[3];
    for (int k = 0; k < 3; ++k) out[k] = h[k] - hn * n[k];
    const double prjmag = norm3(out);
    double scale = 0.0;
    if (prjmag > 0.0) scale = shrmag / prjmag;
    for (int k = 0; k < 3; ++k) h[k] = out[k] * scale;
}

// PairGranular.compute_kernel, granular.py:282-430
void pair_granular(Oracle& s, double dt, int history_update) {
    const orc_pair_params& P = s.pair;
    const bool uses_history =
        P.tangential_model == TANG_LINEAR_HISTORY || P.tangential_model == TANG_MINDLIN;
    ORC_PAR
    for (int nc = 0; nc < s.npairs; ++nc) {
        const int i = s.pair_i[nc], j = s.pair_j[nc];
        const double* xi = &s.x[3 * i];
        const double* xj = &s.x[3 * j];
        const double* vi = &s.v[3 * i];
        const double* vj = &s.v[3 * j];
        const double* oi = &s.omega[3 * i];
        const double* oj = &s.omega[3 * j];
        const double ri = s.radius[i], rj = s.radius[j];
        const double mi = s.rmass[i], mj = s.rmass[j];
        const double radsum = ri + rj;
        double dpos[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        double dvj[3];
        minimum_image_and_vshift(s, dpos, dvj);
        const double rsq = dot(dpos, dpos);
        double* hs = &s.shear[(size_t)nc * 3];
        if (rsq < radsum * radsum && rsq > 1e-28) {
            const double r = std::sqrt(rsq);
            const double rinv = 1.0 / r;
            double nrm[3] = {dpos[0] * rinv, dpos[1] * rinv, dpos[2] * rinv};
            const double delta = radsum - r;
            const double reff = (ri * rj) / radsum;
            const double meff = (mi * mj) / (mi + mj);
            const double a = std::sqrt(delta * reff);
            double vr[3];
            for (int k = 0; k < 3; ++k) vr[k] = vi[k] - (vj[k] + dvj[k]);
            const double vnnr = dot(vr, nrm);
            double vt[3], wr[3], wxn[3], vtr[3];
            for (int k = 0; k < 3; ++k) {
                vt[k] = vr[k] - vnnr * nrm[k];
                wr[k] = ri * oi[k] + rj * oj[k];
            }
            cross(wr, nrm, wxn);
            for (int k = 0; k < 3; ++k) vtr[k] = vt[k] - wxn[k];
            const double vrel = norm3(vtr);

            double fne;
            if (P.normal_model != NORMAL_HOOKE) fne = P.g_kn * a * delta;
            else fne = P.g_kn * delta;

            double damp_prefactor = 0.0;
            switch (P.damping_model) {
                case DAMP_VELOCITY: damp_prefactor = P.g_damp; break;
                case DAMP_MASS_VELOCITY: damp_prefactor = P.g_damp * meff; break;
                case DAMP_VISCOELASTIC: damp_prefactor = P.g_damp * meff * a; break;
                case DAMP_TSUJI:
                case DAMP_COEFF_RESTITUTION: {
                    double arg = 0.0;
                    if (delta > 0.0) arg = std::max(0.0, meff * fne / delta);
                    damp_prefactor = P.g_damp * std::sqrt(arg);
                } break;
                default: break;
            }
            double fntot = fne - damp_prefactor * vnnr;
            if (P.limit_damping == 1) fntot = std::max(fntot, 0.0);
            const double fscrit = P.g_mu * std::fabs(fntot);

            double fs[3] = {0.0, 0.0, 0.0};
            if (P.tangential_model == TANG_LINEAR_NOHISTORY) {
                const double damp_t = P.g_xt * damp_prefactor;
                double ft = 0.0;
                if (vrel != 0.0) ft = std::min(fscrit, damp_t * vrel) / vrel;
                for (int k = 0; k < 3; ++k) fs[k] = -ft * vtr[k];
            } else if (uses_history) {
                const int jtag = s.tag[j];
                double hist[3] = {hs[0], hs[1], hs[2]};
                if (s.partner[nc] != jtag) {
                    hist[0] = hist[1] = hist[2] = 0.0;
                    s.partner[nc] = jtag;
                }
                double k_eff = P.g_kt;
                if (P.tangential_model == TANG_MINDLIN) k_eff = P.g_kt * a;
                const double damp_t = P.g_xt * damp_prefactor;
                if (history_update != 0) {
                    const double rsht = dot(hist, nrm);
                    if (std::fabs(rsht) * k_eff > GRAN_EPSILON * fscrit) rotate_rescale(hist, nrm);
                    for (int k = 0; k < 3; ++k) hist[k] = hist[k] + vtr[k] * dt;
                }
                double fdamp[3];
                for (int k = 0; k < 3; ++k) {
                    fdamp[k] = -damp_t * vtr[k];
                    fs[k] = -k_eff * hist[k] + fdamp[k];
                }
                const double magfs = norm3(fs);
                if (magfs > fscrit) {
                    if (norm3(hist) != 0.0) {
                        for (int k = 0; k < 3; ++k) fs[k] = fs[k] * (fscrit / magfs);
                        for (int k = 0; k < 3; ++k) hist[k] = -(fs[k] - fdamp[k]) / k_eff;
                    } else {
                        fs[0] = fs[1] = fs[2] = 0.0;
                    }
                }
                hs[0] = hist[0]; hs[1] = hist[1]; hs[2] = hist[2];
            }
            double F[3];
            for (int k = 0; k < 3; ++k) F[k] = fntot * nrm[k] + fs[k];
            double c[3], ti[3], tj[3];
            cross(nrm, fs, c);
            for (int k = 0; k < 3; ++k) {
                const double tor = -c[k];
                ti[k] = (ri - 0.5 * delta) * tor;
                tj[k] = (rj - 0.5 * delta) * tor;
            }
            tally(s, i, j, dpos, F, ti, tj);
        } else if (uses_history) {
            s.partner[nc] = -1;
            hs[0] = hs[1] = hs[2] = 0.0;
        }
    }
}

void pair_compute(Oracle& s, double dt, int shearupdate) {
    if (s.n == 0 || !s.have_pair) return;
    if (s.pair.style == STYLE_GRANULAR) pair_granular(s, dt, shearupdate);
    else pair_legacy_history(s, dt, shearupdate, s.pair.style == STYLE_HERTZ_HISTORY);
}

// AtomSystem.clear_forces, atom.py:201-220
void clear_forces(Oracle& s) {
    std::fill(s.f.begin(), s.f.end(), 0.0);
    std::fill(s.torque.begin(), s.torque.end(), 0.0);
    std::fill(s.virial.begin(), s.virial.end(), 0.0);
}

// ------------------------------------------------------------- integrator

// FixNVESphere.initial_integrate_kernel, nve_sphere.py:43-76
void nve_initial(Oracle& s, double dt) {
    const double dtv = dt, dtf = 0.5 * dt;
    ORC_PAR
    for (int i = 0; i < s.n; ++i) {
        const double m = s.rmass[i], r = s.radius[i];
        const double inertia = 0.4 * m * r * r;
        for (int k = 0; k < 3; ++k) s.v[3 * i + k] += dtf * (s.f[3 * i + k] / m);
        for (int k = 0; k < 3; ++k) s.x[3 * i + k] = s.x[3 * i + k] + dtv * s.v[3 * i + k];
        for (int k = 0; k < 3; ++k) s.omega[3 * i + k] += dtf * (s.torque[3 * i + k] / inertia);
    }
}

// FixNVESphere.final_integrate_kernel, nve_sphere.py:78-96
void nve_final(Oracle& s, double dt) {
    const double dtf = 0.5 * dt;
    ORC_PAR
    for (int i = 0; i < s.n; ++i) {
        const double m = s.rmass[i], r = s.radius[i];
        const double inertia = 0.4 * m * r * r;
        for (int k = 0; k < 3; ++k) s.v[3 * i + k] += dtf * (s.f[3 * i + k] / m);
        for (int k = 0; k < 3; ++k) s.omega[3 * i + k] += dtf * (s.torque[3 * i + k] / inertia);
    }
}

// FixDampingCundall.post_force_kernel, damping_cundall.py:35-53
void cundall_post_force(Oracle& s) {
    for (int i = 0; i < s.n; ++i)
        for (int d = 0; d < 3; ++d) {
            double wl = s.f[3 * i + d] * s.v[3 * i + d];
            double sf = wl >= 0.0 ? 1.0 : -1.0;
            s.f[3 * i + d] *= 1.0 - s.gamma_lin * sf;
            double wa = s.torque[3 * i + d] * s.omega[3 * i + d];
            double st = wa >= 0.0 ? 1.0 : -1.0;
            s.torque[3 * i + d] *= 1.0 - s.gamma_ang * st;
        }
}

// ------------------------------------------------------------------ walls

// FixWallGran.post_force_kernel, wall_gran.py:120-211: the wall as a partner
// of zero radius and infinite mass (radsum = Reff = r_i, meff = m_i, moment
// arm r_i), hooke normal force with mass-velocity damping, and the
// linear_history_classic tangential force (or the history-free viscous one).
void wall_kernel(Oracle& s, OrcWall& w, double dt, double coord, double vwall_n,
                 double vwall_s, int history_update) {
    const orc_wall_params& P = w.p;
    const int axis = P.axis;
    double nv[3] = {0.0, 0.0, 0.0};
    nv[axis] = P.side == 1 ? -1.0 : 1.0;
    double vwall[3] = {0.0, 0.0, 0.0};
    vwall[axis] = vwall_n;
    if (P.shear_axis >= 0) vwall[P.shear_axis] = vwall_s;
    for (int i = 0; i < s.n; ++i) {
        const double ri = s.radius[i];
        const double xa = s.x[3 * i + axis];
        const double dist = P.side == -1 ? xa - coord : coord - xa;
        const double delta = ri - dist;
        if (delta <= 0.0) {
            if (P.use_history) {
                w.touch[i] = 0;
                for (int k = 0; k < 3; ++k) w.shear[3 * i + k] = 0.0;
            }
            continue;
        }
        const double meff = s.rmass[i];
        double vr[3];
        for (int k = 0; k < 3; ++k) vr[k] = s.v[3 * i + k] - vwall[k];
        const double vnnr = vr[0] * nv[0] + vr[1] * nv[1] + vr[2] * nv[2];
        double vt[3], wr[3], vtr[3];
        for (int k = 0; k < 3; ++k) vt[k] = vr[k] - vnnr * nv[k];
        for (int k = 0; k < 3; ++k) wr[k] = ri * s.omega[3 * i + k];
        const double wxn[3] = {wr[1] * nv[2] - wr[2] * nv[1], wr[2] * nv[0] - wr[0] * nv[2],
                               wr[0] * nv[1] - wr[1] * nv[0]};
        for (int k = 0; k < 3; ++k) vtr[k] = vt[k] - wxn[k];
        const double vrel = std::sqrt(vtr[0] * vtr[0] + vtr[1] * vtr[1] + vtr[2] * vtr[2]);
        const double damp_prefactor = meff * P.gamman;
        double fntot = P.kn * delta - damp_prefactor * vnnr;
        if (P.limit_damping) fntot = std::max(fntot, 0.0);
        const double fscrit = P.xmu * std::fabs(fntot);
        const double damp_t = meff * P.gammat;
        double fs[3] = {0.0, 0.0, 0.0};
        if (P.use_history) {
            double h[3] = {w.shear[3 * i], w.shear[3 * i + 1], w.shear[3 * i + 2]};
            if (w.touch[i] == 0) {
                h[0] = h[1] = h[2] = 0.0;
                w.touch[i] = 1;
            }
            if (history_update)
                for (int k = 0; k < 3; ++k) h[k] = h[k] + vtr[k] * dt;
            const double shrmag = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
            if (history_update) {
                const double hn = h[0] * nv[0] + h[1] * nv[1] + h[2] * nv[2];
                for (int k = 0; k < 3; ++k) h[k] = h[k] - hn * nv[k];
            }
            double fdamp[3];
            for (int k = 0; k < 3; ++k) fdamp[k] = -damp_t * vtr[k];
            for (int k = 0; k < 3; ++k) fs[k] = -P.kt * h[k] + fdamp[k];
            const double magfs = std::sqrt(fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]);
            if (magfs > fscrit) {
                if (shrmag != 0.0) {
                    const double sc = fscrit / magfs;
                    for (int k = 0; k < 3; ++k) fs[k] = fs[k] * sc;
                    for (int k = 0; k < 3; ++k) h[k] = -(fs[k] - fdamp[k]) / P.kt;
                } else {
                    fs[0] = fs[1] = fs[2] = 0.0;
                }
            }
            for (int k = 0; k < 3; ++k) w.shear[3 * i + k] = h[k];
        } else {
            double ft = 0.0;
            if (vrel != 0.0) ft = std::min(fscrit, damp_t * vrel) / vrel;
            for (int k = 0; k < 3; ++k) fs[k] = -ft * vtr[k];
        }
        for (int k = 0; k < 3; ++k) s.f[3 * i + k] = s.f[3 * i + k] + (fntot * nv[k] + fs[k]);
        const double nxf[3] = {nv[1] * fs[2] - nv[2] * fs[1], nv[2] * fs[0] - nv[0] * fs[2],
                               nv[0] * fs[1] - nv[1] * fs[0]};
        for (int k = 0; k < 3; ++k) s.torque[3 * i + k] = s.torque[3 * i + k] + -ri * nxf[k];
    }
}

// FixWallGranRegion.post_force_kernel, wall_gran_region.py:114-271: the
// curved side of a cylinder region (zero radius, infinite mass); hooke or
// hertz (contact radius sqrt(r delta) on stiffness and damping).
void wall_cyl_kernel(Oracle& s, OrcWall& w, double dt, int history_update) {
    const orc_wall_params& P = w.p;
    const int axis = P.axis;
    for (int i = 0; i < s.n; ++i) {
        const double* pos = &s.x[3 * i];
        const double r_i = s.radius[i];
        const double pos_axis = pos[axis];
        auto reset = [&]() {
            if (P.use_history) {
                w.touch[i] = 0;
                for (int k = 0; k < 3; ++k) w.shear[3 * i + k] = 0.0;
            }
        };
        if (P.has_bounds == 1 && (pos_axis < P.axis_lo || pos_axis > P.axis_hi)) {
            reset();
            continue;
        }
        double d1, d2;
        if (axis == 0) {
            d1 = pos[1] - P.c1;
            d2 = pos[2] - P.c2;
        } else if (axis == 1) {
            d1 = pos[0] - P.c1;
            d2 = pos[2] - P.c2;
        } else {
            d1 = pos[0] - P.c1;
            d2 = pos[1] - P.c2;
        }
        const double r_perp = std::sqrt(d1 * d1 + d2 * d2);
        if (!(r_perp > 1e-14)) {
            reset();
            continue;
        }
        const double inv_r_perp = 1.0 / r_perp;
        const double u1 = d1 * inv_r_perp, u2 = d2 * inv_r_perp;
        const double s1 = P.side == -1 ? -u1 : u1, s2 = P.side == -1 ? -u2 : u2;
        double nw[3];
        if (axis == 0) {
            nw[0] = 0.0; nw[1] = s1; nw[2] = s2;
        } else if (axis == 1) {
            nw[0] = s1; nw[1] = 0.0; nw[2] = s2;
        } else {
            nw[0] = s1; nw[1] = s2; nw[2] = 0.0;
        }
        const double delta =
            P.side == -1 ? r_i - (P.radius - r_perp) : r_i - (r_perp - P.radius);
        if (!(delta > 0.0)) {
            reset();
            continue;
        }
        const double mi = s.rmass[i];
        const double* vi = &s.v[3 * i];
        const double* oi = &s.omega[3 * i];
        const double oxn[3] = {oi[1] * nw[2] - oi[2] * nw[1], oi[2] * nw[0] - oi[0] * nw[2],
                               oi[0] * nw[1] - oi[1] * nw[0]};
        double vr[3], vrt[3];
        for (int k = 0; k < 3; ++k) vr[k] = vi[k] - r_i * oxn[k];
        const double vn = vr[0] * nw[0] + vr[1] * nw[1] + vr[2] * nw[2];
        for (int k = 0; k < 3; ++k) vrt[k] = vr[k] - vn * nw[k];
        const double vrel = std::sqrt(vrt[0] * vrt[0] + vrt[1] * vrt[1] + vrt[2] * vrt[2]);
        double poly = 1.0;
        if (P.hertz == 1) poly = std::sqrt(r_i * delta);
        const double fn_elastic = P.kn * delta * poly;
        const double damp_prefactor = mi * P.gamman * poly;
        double fntot = fn_elastic - damp_prefactor * vn;
        if (P.limit_damping) fntot = std::max(fntot, 0.0);
        const double fscrit = P.xmu * std::fabs(fntot);
        const double damp_t = mi * P.gammat * poly;
        double ft[3] = {0.0, 0.0, 0.0};
        if (P.use_history) {
            double h[3] = {w.shear[3 * i], w.shear[3 * i + 1], w.shear[3 * i + 2]};
            if (w.touch[i] == 0) {
                h[0] = h[1] = h[2] = 0.0;
                w.touch[i] = 1;
            }
            if (history_update)
                for (int k = 0; k < 3; ++k) h[k] = h[k] + vrt[k] * dt;
            const double shrmag = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
            if (history_update) {
                const double hn = h[0] * nw[0] + h[1] * nw[1] + h[2] * nw[2];
                for (int k = 0; k < 3; ++k) h[k] = h[k] - hn * nw[k];
            }
            double fdamp[3];
            for (int k = 0; k < 3; ++k) fdamp[k] = -damp_t * vrt[k];
            for (int k = 0; k < 3; ++k) ft[k] = -P.kt * poly * h[k] + fdamp[k];
            const double magfs = std::sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
            if (magfs > fscrit) {
                if (shrmag != 0.0) {
                    const double sc = fscrit / magfs;
                    for (int k = 0; k < 3; ++k) ft[k] = ft[k] * sc;
                    for (int k = 0; k < 3; ++k) h[k] = -(ft[k] - fdamp[k]) / (P.kt * poly);
                } else {
                    ft[0] = ft[1] = ft[2] = 0.0;
                }
            }
            for (int k = 0; k < 3; ++k) w.shear[3 * i + k] = h[k];
        } else {
            double fsc = 0.0;
            if (vrel != 0.0) fsc = std::min(fscrit, damp_t * vrel) / vrel;
            for (int k = 0; k < 3; ++k) ft[k] = -fsc * vrt[k];
        }
        for (int k = 0; k < 3; ++k) s.f[3 * i + k] = s.f[3 * i + k] + (fntot * nw[k] + ft[k]);
        const double nxf[3] = {nw[1] * ft[2] - nw[2] * ft[1], nw[2] * ft[0] - nw[0] * ft[2],
                               nw[0] * ft[1] - nw[1] * ft[0]};
        for (int k = 0; k < 3; ++k) s.torque[3 * i + k] = s.torque[3 * i + k] + -r_i * nxf[k];
    }
}

// FixWallGran.post_force, wall_gran.py:213-236: the wall's position and
// normal velocity under `wiggle`, then the contact kernel; the clock
// advances by dt per call.
void wall_post_force(Oracle& s, OrcWall& w, double dt, int history_update) {
    if (w.p.geometry == 1) {  // FixWallGranRegion.post_force, wall_gran_region.py:273-290
        wall_cyl_kernel(s, w, dt, history_update);
        return;
    }
    double coord = w.coord, vwall_n = 0.0;
    if (w.amp != 0.0) {
        const double omega_w = 2.0 * M_PI / w.period;
        const double arg = omega_w * w.elapsed;
        coord += w.amp * (1.0 - std::cos(arg));
        vwall_n = w.amp * omega_w * std::sin(arg);
    }
    wall_kernel(s, w, dt, coord, vwall_n, w.shear_vel, history_update);
    w.elapsed += dt;
}

// ----------------------------------------------------------------- virial

// Per-atom leaf of the virial(+kinetic) sum, compute_virial_kernel thermo.py:62-109
inline void virial_leaf(const Oracle& s, int i, int keflag, double out[6]) {
    for (int k = 0; k < 6; ++k) out[k] = s.virial[6 * i + k];
    if (keflag) {
        const double m = s.rmass[i];
        const double* vi = &s.v[3 * i];
        out[0] = out[0] + m * vi[0] * vi[0];
        out[1] = out[1] + m * vi[1] * vi[1];
        out[2] = out[2] + m * vi[2] * vi[2];
        out[3] = out[3] + m * vi[0] * vi[1];
        out[4] = out[4] + m * vi[0] * vi[2];
        out[5] = out[5] + m * vi[1] * vi[2];
    }
}

// mode 0: pairwise tree (the order the device reduction uses);
// mode 1: the reference's strided chunk order, thermo.py:28-30,98-109.
void virial_tensor(const Oracle& s, int keflag, int mode, double out[6]) {
    const int n = s.n;
    if (mode == 1) {
        int chunks = (int)std::min(1024.0, std::max(64.0, 2.0 * std::sqrt((double)std::max(n, 1))));
        std::vector<double> part((size_t)chunks * 6, 0.0);
        for (int c = 0; c < chunks; ++c) {
            double acc[6] = {0, 0, 0, 0, 0, 0};
            for (int i = c; i < n; i += chunks) {
                for (int k = 0; k < 6; ++k) acc[k] += s.virial[6 * i + k];
                if (keflag) {
                    const double m = s.rmass[i];
                    const double* vi = &s.v[3 * i];
                    acc[0] += m * vi[0] * vi[0];
                    acc[1] += m * vi[1] * vi[1];
                    acc[2] += m * vi[2] * vi[2];
                    acc[3] += m * vi[0] * vi[1];
                    acc[4] += m * vi[0] * vi[2];
                    acc[5] += m * vi[1] * vi[2];
                }
            }
            for (int k = 0; k < 6; ++k) part[(size_t)c * 6 + k] = acc[k];
        }
        for (int k = 0; k < 6; ++k) {
            double total = 0.0;
            for (int c = 0; c < chunks; ++c) total += part[(size_t)c * 6 + k];
            out[k] = total;
        }
        return;
    }
    size_t P = 1;
    while (P < (size_t)n) P <<= 1;
    std::vector<double> buf(P * 6, 0.0);
    ORC_PAR
    for (int i = 0; i < n; ++i) virial_leaf(s, i, keflag, &buf[(size_t)i * 6]);
    for (size_t w = 1; w < P; w <<= 1) {
        const long long pairs = (long long)(P / (2 * w));
        ORC_PAR
        for (long long q = 0; q < pairs; ++q) {
            const size_t i = (size_t)q * 2 * w;
            for (int k = 0; k < 6; ++k) buf[i * 6 + k] = buf[i * 6 + k] + buf[(i + w) * 6 + k];
        }
    }
    for (int k = 0; k < 6; ++k) out[k] = buf[k];
}

// Pairwise tree over per-atom leaves of NC components (zero-padded to a power
// of two, adjacent pairs combined level by level), as virial_tensor mode 0.
void pairwise_tree(std::vector<double>& buf, int n, int nc, double* out) {
    size_t P = 1;
    while (P < (size_t)n) P <<= 1;
    buf.resize(P * nc, 0.0);
    for (size_t w = 1; w < P; w <<= 1)
        for (size_t i = 0; i < P; i += 2 * w)
            for (int k = 0; k < nc; ++k) buf[i * nc + k] = buf[i * nc + k] + buf[(i + w) * nc + k];
    for (int k = 0; k < nc; ++k) out[k] = buf[k];
}

// ----------------------------------------------------- per-contact computes

// ComputeContactAtom.compute_kernel, contact_atom.py:19-52: every half-list
// pair (i, j > i) that overlaps at the current positions (rsq < radsum^2 and
// rsq > 1e-18) is tallied to both atoms.
void contact_atom(const Oracle& s, int* count) {
    for (int i = 0; i < s.n; ++i) count[i] = 0;
    if (!s.has_built) return;
    for (int i = 0; i < s.n; ++i) {
        const double* xi = &s.x[3 * i];
        const double ri = s.radius[i];
        for (int k = 0; k < s.num_neighbors[i]; ++k) {
            const int j = s.neighbors[(size_t)i * s.max_nb + k];
            double d[3] = {xi[0] - s.x[3 * j], xi[1] - s.x[3 * j + 1], xi[2] - s.x[3 * j + 2]};
            minimum_image(s, d);
            const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double radsum = ri + s.radius[j];
            if (rsq < radsum * radsum && rsq > 1e-18) {
                ++count[i];
                ++count[j];
            }
        }
    }
}

// ComputeFabric.compute_kernel, fabric.py:26-78: the sums of n_a n_b over the
// overlapping half-list pairs (n = d / |d|, d = x_i - x_j) and their number Nc,
// then phi = sum * (1 / Nc). The sums are per-atom leaves (ascending j) in a
// pairwise tree over atoms; the reference adds them with float atomics in an
// unspecified order.
void fabric(const Oracle& s, double phi[6], int* ncontacts) {
    std::vector<double> buf((size_t)std::max(s.n, 1) * 7, 0.0);
    if (s.has_built) {
        for (int i = 0; i < s.n; ++i) {
            const double* xi = &s.x[3 * i];
            const double ri = s.radius[i];
            double* leaf = &buf[(size_t)i * 7];
            for (int k = 0; k < s.num_neighbors[i]; ++k) {
                const int j = s.neighbors[(size_t)i * s.max_nb + k];
                double d[3] = {xi[0] - s.x[3 * j], xi[1] - s.x[3 * j + 1],
                               xi[2] - s.x[3 * j + 2]};
                minimum_image(s, d);
                const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double radsum = ri + s.radius[j];
                if (rsq < radsum * radsum && rsq > 1e-18) {
                    const double r = std::sqrt(rsq);
                    const double nx = d[0] / r, ny = d[1] / r, nz = d[2] / r;
                    leaf[0] = leaf[0] + nx * nx;
                    leaf[1] = leaf[1] + ny * ny;
                    leaf[2] = leaf[2] + nz * nz;
                    leaf[3] = leaf[3] + nx * ny;
                    leaf[4] = leaf[4] + nx * nz;
                    leaf[5] = leaf[5] + ny * nz;
                    leaf[6] = leaf[6] + 1.0;
                }
            }
        }
    }
    double sum[7];
    pairwise_tree(buf, std::max(s.n, 1), 7, sum);
    const int nc = (int)sum[6];
    *ncontacts = nc;
    if (nc > 0) {
        const double inv_nc = 1.0 / (double)nc;
        for (int k = 0; k < 6; ++k) phi[k] = sum[k] * inv_nc;
    } else {
        for (int k = 0; k < 6; ++k) phi[k] = 0.0;
    }
}

// ComputeStressAtom.compute_stress_kernel, stress_atom.py:30-52: s = -virial,
// minus m v_a v_b with keflag; the macro sum of s over atoms is a pairwise
// tree (the reference adds with float atomics).
void stress_atom(const Oracle& s, int keflag, double* stress, double sum6[6]) {
    std::vector<double> buf((size_t)std::max(s.n, 1) * 6, 0.0);
    for (int i = 0; i < s.n; ++i) {
        double* o = &buf[(size_t)i * 6];
        for (int k = 0; k < 6; ++k) o[k] = -s.virial[6 * i + k];
        if (keflag) {
            const double m = s.rmass[i];
            const double* vi = &s.v[3 * i];
            o[0] = o[0] - m * vi[0] * vi[0];
            o[1] = o[1] - m * vi[1] * vi[1];
            o[2] = o[2] - m * vi[2] * vi[2];
            o[3] = o[3] - m * vi[0] * vi[1];
            o[4] = o[4] - m * vi[0] * vi[2];
            o[5] = o[5] - m * vi[1] * vi[2];
        }
        if (stress)
            for (int k = 0; k < 6; ++k) stress[6 * i + k] = o[k];
    }
    pairwise_tree(buf, std::max(s.n, 1), 6, sum6);
}

// ------------------------------------------------------------- the servo

// FixDeformPressure._strain_h_rate, deform_pressure.py:447-472 (x/y/z entries)
void servo_strain_h_rate(const Oracle& s, double dt, double h[6]) {
    for (int k = 0; k < 6; ++k) h[k] = 0.0;
    const orc_servo_params& S = s.servo;
    double delt = s.nsteps_total * dt;
    for (int i = 0; i < 3; ++i) {
        double length = s.hi_start[i] - s.lo_start[i];
        int st = S.style[i];
        if (st == S_ERATE || st == S_TRATE) h[i] = S.rate[i] * length;
        else if (st == S_VEL) h[i] = S.vel[i];
        else if (st == S_WIGGLE) h[i] = 2.0 * M_PI / S.tper[i] * S.amp[i];
        else if ((st == S_FINAL || st == S_DELTA || st == S_SCALE) && delt > 0.0) {
            double stop;
            if (st == S_FINAL) stop = S.fhi[i] - S.flo[i];
            else if (st == S_DELTA) stop = length + S.dhi[i] - S.dlo[i];
            else stop = S.scale[i] * length;
            h[i] = (stop - length) / delt;
        }
    }
}

// FixDeformPressure.setup, deform_pressure.py:392-445
void servo_setup(Oracle& s, int nsteps_total, double dt) {
    s.nsteps = 0;
    s.nsteps_total = nsteps_total;
    for (int i = 0; i < 3; ++i) {
        s.lo_start[i] = s.lo[i];
        s.hi_start[i] = s.hi[i];
        s.cum[i] = 0.0;
    }
    double h[6];
    servo_strain_h_rate(s, dt, h);
    for (int k = 0; k < 6; ++k) s.h_rate[k] = h[k];
    s.vremap = (s.servo.remap == 1) ? 1 : 0;
}

// _limit_rate, deform_pressure.py:727-741
inline double limit_rate(const orc_servo_params& S, double rate, double ptarget) {
    if (S.normalize) {
        if (ptarget == 0.0) {
            if (rate != 0.0) rate = std::copysign(S.max_rate, rate);
        } else {
            rate /= std::fabs(ptarget);
        }
    }
    if (S.max_rate != 0.0 && std::fabs(rate) > S.max_rate) rate = std::copysign(S.max_rate, rate);
    return rate;
}

// FixDeformPressure.end_of_step host path (deform_pressure.py:1113-1219) with
// _current_pressure :706-725, _apply_strain :758-831, _apply_pressure :833-857,
// _normalize_and_flip :1003-1083 (zero tilt) and remap_positions :493-530.
// Returns 1 if a box dimension went non-positive (the RuntimeError at :1167-1171).
int servo_end_of_step(Oracle& s, double dt) {
    const orc_servo_params& S = s.servo;
    s.step_count += 1;
    s.nsteps += 1;
    if (s.step_count % S.nevery != 0) return 0;

    bool needs_pressure = false;
    for (int i = 0; i < 3; ++i)
        if (S.style[i] == S_PRESSURE || S.style[i] == S_PMEAN) needs_pressure = true;

    double p[3] = {0, 0, 0};
    if (needs_pressure) {
        double vol = s.prd[0] * s.prd[1] * s.prd[2];
        double t[6] = {0, 0, 0, 0, 0, 0};
        if (s.n > 0 && vol > 0.0) {
            double vt[6];
            virial_tensor(s, 1, 0, vt);
            for (int k = 0; k < 6; ++k) t[k] = vt[k] / vol;
        }
        for (int k = 0; k < 6; ++k) s.last_tensor[k] = t[k];
        double scalar = (t[0] + t[1] + t[2]) / 3.0;
        if (S.couple == 1) {
            p[0] = p[1] = p[2] = scalar;
        } else if (S.couple >= 2) {
            int a = 0, b = 1;
            if (S.couple == 3) { a = 1; b = 2; }
            if (S.couple == 4) { a = 0; b = 2; }
            double mean = (t[a] + t[b]) / 2.0;
            for (int i = 0; i < 3; ++i) p[i] = t[i];
            p[a] = mean;
            p[b] = mean;
        } else {
            for (int i = 0; i < 3; ++i) {
                if (S.style[i] == S_PRESSURE) p[i] = t[i];
                else if (S.style[i] == S_PMEAN) p[i] = scalar;
            }
        }
    }
    double old_prd[3] = {s.prd[0], s.prd[1], s.prd[2]};
    double old_lo[3] = {s.lo[0], s.lo[1], s.lo[2]};
    double lo_t[3], hi_t[3];
    const double elapsed = s.nsteps * dt;
    const double frac = s.nsteps_total > 0 ? (double)s.nsteps / (double)s.nsteps_total : 0.0;
    for (int i = 0; i < 3; ++i) {
        const double ls = s.lo_start[i], hs = s.hi_start[i];
        lo_t[i] = s.lo[i];
        hi_t[i] = s.hi[i];
        switch (S.style[i]) {
            case S_TRATE: {
                double mid = 0.5 * (ls + hs);
                double half = 0.5 * (hs - ls) * det_exp(S.rate[i] * elapsed);
                lo_t[i] = mid - half; hi_t[i] = mid + half;
            } break;
            case S_ERATE: {
                double shift = 0.5 * elapsed * S.rate[i] * (hs - ls);
                lo_t[i] = ls - shift; hi_t[i] = hs + shift;
            } break;
            case S_VEL: {
                double shift = 0.5 * elapsed * S.vel[i];
                lo_t[i] = ls - shift; hi_t[i] = hs + shift;
            } break;
            case S_WIGGLE: {
                double shift = 0.5 * S.amp[i] * det_sin(2.0 * M_PI * elapsed / S.tper[i]);
                lo_t[i] = ls - shift; hi_t[i] = hs + shift;
            } break;
            case S_FINAL:
            case S_DELTA:
            case S_SCALE: {
                double lo_stop, hi_stop;
                if (S.style[i] == S_FINAL) { lo_stop = S.flo[i]; hi_stop = S.fhi[i]; }
                else if (S.style[i] == S_DELTA) { lo_stop = ls + S.dlo[i]; hi_stop = hs + S.dhi[i]; }
                else {
                    double mid = 0.5 * (ls + hs);
                    double half = 0.5 * S.scale[i] * (hs - ls);
                    lo_stop = mid - half; hi_stop = mid + half;
                }
                lo_t[i] = ls + frac * (lo_stop - ls);
                hi_t[i] = hs + frac * (hi_stop - hs);
            } break;
            case S_PRESSURE:
            case S_PMEAN: {
                double rate = limit_rate(S, S.pgain[i] * (p[i] - S.ptarget[i]), S.ptarget[i]);
                s.cum[i] += dt * rate * old_prd[i];
                lo_t[i] = ls - 0.5 * s.cum[i];
                hi_t[i] = hs + 0.5 * s.cum[i];
            } break;
            default: break;
        }
    }
    double new_prd[3];
    for (int i = 0; i < 3; ++i) new_prd[i] = hi_t[i] - lo_t[i];
    for (int i = 0; i < 3; ++i)
        if (new_prd[i] <= 0.0) return 1;
    double hr[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) hr[i] = (new_prd[i] - old_prd[i]) / dt;
    hr[3] = (0.0 - 0.0) / dt; hr[4] = (0.0 - 0.0) / dt; hr[5] = (0.0 - 0.0) / dt;

    if (S.remap == 0 && s.n > 0) {
        // remap_positions with zero tilt: old_hinv_tilt = [-0/(x*y), (0*0-y*0)/(xyz), -0/(y*z)]
        const double xprd = old_prd[0], yprd = old_prd[1], zprd = old_prd[2];
        const double oh0 = -0.0 / (xprd * yprd);
        const double oh1 = (0.0 * 0.0 - yprd * 0.0) / (xprd * yprd * zprd);
        const double oh2 = -0.0 / (yprd * zprd);
        ORC_PAR
        for (int i = 0; i < s.n; ++i) {
            double* xi = &s.x[3 * i];
            double d0 = xi[0] - old_lo[0], d1 = xi[1] - old_lo[1], d2 = xi[2] - old_lo[2];
            double l0 = d0 / old_prd[0] + oh0 * d1 + oh1 * d2;
            double l1 = d1 / old_prd[1] + oh2 * d2;
            double l2 = d2 / old_prd[2];
            xi[0] = new_prd[0] * l0 + 0.0 * l1 + 0.0 * l2 + lo_t[0];
            xi[1] = new_prd[1] * l1 + 0.0 * l2 + lo_t[1];
            xi[2] = new_prd[2] * l2 + lo_t[2];
        }
    }
    // Domain.set_box_and_h_rate -> _sync_kernel, domain.py:286-307,188-233
    for (int i = 0; i < 3; ++i) {
        s.lo[i] = lo_t[i];
        s.hi[i] = hi_t[i];
        s.prd[i] = s.hi[i] - s.lo[i];
    }
    for (int k = 0; k < 6; ++k) s.h_rate[k] = hr[k];
    s.vremap = (S.remap == 1) ? 1 : 0;
    return 0;
}

// ---------------------------------------------------------- Simulation

// Simulation.reneighbor_if_needed, simulation.py:122-159. Returns <0 on error,
// 1 if rebuilt, 0 otherwise.
int reneighbor_if_needed(Oracle& s, bool force) {
    bool decided = nlist_decide(s, s.timestep);
    if (!decided && !force) return 0;
    if (s.has_built) history_save(s);   // save_state (npairs of the old list)
    pbc(s);
    int ovf = nlist_build(s);
    if (ovf) return -1;
    if (s.save_count.empty()) s.save_count.assign(s.n, 0);
    history_restore(s);
    s.rebuilds_in_run += 1;
    return 1;
}

// Simulation.init_simulation, simulation.py:170-203
int init_simulation(Oracle& s, double dt) {
    s.setup_done = 1;
    clear_forces(s);
    if (s.have_pair) {
        int rc = reneighbor_if_needed(s, true);
        if (rc < 0) return rc;
        pair_compute(s, dt, 0);
    }
    if (s.cundall) cundall_post_force(s);
    for (OrcWall& w : s.walls) wall_post_force(s, w, dt, 0);
    return 0;
}

// Simulation._sub_step_inner, simulation.py:205-247
int sub_step(Oracle& s, double dt) {
    s.timestep += 1;
    nve_initial(s, dt);
    {
        int rc = reneighbor_if_needed(s, false);
        if (rc < 0) return rc;
    }
    clear_forces(s);
    if (s.have_pair) pair_compute(s, dt, 1);
    if (s.cundall) cundall_post_force(s);
    for (OrcWall& w : s.walls) wall_post_force(s, w, dt, 1);
    nve_final(s, dt);
    if (s.servo.enabled) {
        if (servo_end_of_step(s, dt)) return -2;
    }
    return 0;
}

}  // namespace

// ================================================================ C API

extern "C" {

void* orc_create(int n, int max_nb) {
    Oracle* s = new Oracle();
    s->n = n;
    s->max_nb = max_nb;
    s->x.assign((size_t)3 * n, 0.0);
    s->v = s->x; s->omega = s->x; s->f = s->x; s->torque = s->x;
    s->radius.assign(n, 0.0);
    s->rmass.assign(n, 0.0);
    s->virial.assign((size_t)6 * n, 0.0);
    s->tag.assign(n, 0);
    for (int i = 0; i < n; ++i) s->tag[i] = i + 1;
    history_reset(*s, 0);
    return s;
}

void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

void orc_set_atoms(void* h, const double* x, const double* v, const double* omega,
                   const double* radius, const double* rmass, const int* tag) {
    Oracle& s = *static_cast<Oracle*>(h);
    std::memcpy(s.x.data(), x, sizeof(double) * 3 * s.n);
    std::memcpy(s.v.data(), v, sizeof(double) * 3 * s.n);
    std::memcpy(s.omega.data(), omega, sizeof(double) * 3 * s.n);
    std::memcpy(s.radius.data(), radius, sizeof(double) * s.n);
    std::memcpy(s.rmass.data(), rmass, sizeof(double) * s.n);
    std::memcpy(s.tag.data(), tag, sizeof(int) * s.n);
}

void orc_get_atoms(void* h, double* x, double* v, double* omega, double* f, double* torque,
                   double* virial) {
    Oracle& s = *static_cast<Oracle*>(h);
    if (x) std::memcpy(x, s.x.data(), sizeof(double) * 3 * s.n);
    if (v) std::memcpy(v, s.v.data(), sizeof(double) * 3 * s.n);
    if (omega) std::memcpy(omega, s.omega.data(), sizeof(double) * 3 * s.n);
    if (f) std::memcpy(f, s.f.data(), sizeof(double) * 3 * s.n);
    if (torque) std::memcpy(torque, s.torque.data(), sizeof(double) * 3 * s.n);
    if (virial) std::memcpy(virial, s.virial.data(), sizeof(double) * 6 * s.n);
}

void orc_set_forces(void* h, const double* f, const double* torque) {
    Oracle& s = *static_cast<Oracle*>(h);
    std::memcpy(s.f.data(), f, sizeof(double) * 3 * s.n);
    std::memcpy(s.torque.data(), torque, sizeof(double) * 3 * s.n);
}

void orc_set_box(void* h, const double* lo, const double* hi, const int* periodic) {
    Oracle& s = *static_cast<Oracle*>(h);
    for (int k = 0; k < 3; ++k) {
        s.lo[k] = lo[k];
        s.hi[k] = hi[k];
        s.prd[k] = s.hi[k] - s.lo[k];
        s.per[k] = periodic[k];
    }
}

void orc_set_h_rate(void* h, const double* h_rate, int vremap) {
    Oracle& s = *static_cast<Oracle*>(h);
    for (int k = 0; k < 6; ++k) s.h_rate[k] = h_rate[k];
    s.vremap = vremap;
}

void orc_get_box(void* h, double* lo, double* hi, double* h_rate) {
    Oracle& s = *static_cast<Oracle*>(h);
    for (int k = 0; k < 3; ++k) { lo[k] = s.lo[k]; hi[k] = s.hi[k]; }
    if (h_rate) for (int k = 0; k < 6; ++k) h_rate[k] = s.h_rate[k];
}

void orc_set_neigh(void* h, double skin, int delay, int every, int check, int max_grid_cells) {
    Oracle& s = *static_cast<Oracle*>(h);
    s.skin = skin;
    s.delay = delay;
    s.every = every;
    s.check = check;
    s.max_grid_cells = max_grid_cells;
}

void orc_set_pair(void* h, const orc_pair_params* p) {
    Oracle& s = *static_cast<Oracle*>(h);
    s.pair = *p;
    s.have_pair = 1;
}

void orc_set_servo(void* h, const orc_servo_params* p) {
    Oracle& s = *static_cast<Oracle*>(h);
    s.servo = *p;
    if (s.servo.nevery < 1) s.servo.nevery = 1;
}

void orc_set_cundall(void* h, int enabled, double gamma_lin, double gamma_ang) {
    Oracle& s = *static_cast<Oracle*>(h);
    s.cundall = enabled;
    s.gamma_lin = gamma_lin;
    s.gamma_ang = gamma_ang;
}

// Simulation.run(steps), simulation.py:320-332 (run_pre on). Returns 0, or
// -1 on neighbor overflow, -2 when the servo collapses the box.
int orc_run(void* h, int nsteps, double dt, int run_pre) {
    Oracle& s = *static_cast<Oracle*>(h);
    s.rebuilds_in_run = 0;
    if (s.servo.enabled) servo_setup(s, nsteps, dt);
    for (OrcWall& w : s.walls) w.elapsed = 0.0;  // FixWallGran.setup
    if (run_pre || !s.setup_done) {
        int rc = init_simulation(s, dt);
        if (rc) return rc;
    }
    for (int k = 0; k < nsteps; ++k) {
        int rc = sub_step(s, dt);
        if (rc) return rc;
    }
    return 0;
}

int orc_init_simulation(void* h, double dt) {
    return init_simulation(*static_cast<Oracle*>(h), dt);
}
int orc_step(void* h, double dt) { return sub_step(*static_cast<Oracle*>(h), dt); }
void orc_servo_setup(void* h, int nsteps_total, double dt) {
    servo_setup(*static_cast<Oracle*>(h), nsteps_total, dt);
}

// Single stages, for unit tests.
int orc_nlist_build(void* h) {
    Oracle& s = *static_cast<Oracle*>(h);
    int rc = nlist_build(s);
    if ((int)s.partner.size() < s.npairs) history_reset(s, s.npairs);
    return rc;
}
int orc_nlist_decide(void* h, int timestep) {
    return nlist_decide(*static_cast<Oracle*>(h), timestep) ? 1 : 0;
}
int orc_reneighbor(void* h, int force) {
    return reneighbor_if_needed(*static_cast<Oracle*>(h), force != 0);
}
void orc_pbc(void* h) { pbc(*static_cast<Oracle*>(h)); }
void orc_clear_forces(void* h) { clear_forces(*static_cast<Oracle*>(h)); }
void orc_pair_compute(void* h, double dt, int shearupdate) {
    pair_compute(*static_cast<Oracle*>(h), dt, shearupdate);
}
void orc_nve_initial(void* h, double dt) { nve_initial(*static_cast<Oracle*>(h), dt); }
void orc_nve_final(void* h, double dt) { nve_final(*static_cast<Oracle*>(h), dt); }
void orc_cundall(void* h) { cundall_post_force(*static_cast<Oracle*>(h)); }
void orc_virial(void* h, int keflag, int mode, double* out6) {
    virial_tensor(*static_cast<Oracle*>(h), keflag, mode, out6);
}
int orc_servo_end_of_step(void* h, double dt) {
    return servo_end_of_step(*static_cast<Oracle*>(h), dt);
}
// FixWallGran: add a wall (its index is returned), read its history, or run
// one post_force on the current forces.
int orc_add_wall(void* h, const orc_wall_params* p, double coord, double shear_vel, double amp,
                 double period) {
    Oracle& s = *static_cast<Oracle*>(h);
    OrcWall w;
    w.p = *p;
    w.coord = coord;
    w.shear_vel = shear_vel;
    w.amp = amp;
    w.period = period;
    w.shear.assign((size_t)3 * s.n, 0.0);
    w.touch.assign(s.n, 0);
    s.walls.push_back(w);
    return (int)s.walls.size() - 1;
}
void orc_wall_history(void* h, int idx, double* shear, int* touch) {
    Oracle& s = *static_cast<Oracle*>(h);
    const OrcWall& w = s.walls[idx];
    for (int i = 0; i < s.n; ++i) {
        for (int k = 0; k < 3; ++k) shear[3 * i + k] = w.shear[3 * i + k];
        touch[i] = w.touch[i];
    }
}
void orc_wall_post_force(void* h, int idx, double dt, int history_update) {
    Oracle& s = *static_cast<Oracle*>(h);
    wall_post_force(s, s.walls[idx], dt, history_update);
}
// Host threads the loops use (1 in the serial build).
double orc_det_sin(double x) { return det_sin(x); }
double orc_det_exp(double x) { return det_exp(x); }

int orc_num_threads(void) {
#ifdef ORC_OMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_contact_atom(void* h, int* count) { contact_atom(*static_cast<Oracle*>(h), count); }
void orc_fabric(void* h, double* phi6, int* ncontacts) {
    fabric(*static_cast<Oracle*>(h), phi6, ncontacts);
}
void orc_stress_atom(void* h, int keflag, double* stress, double* sum6) {
    stress_atom(*static_cast<Oracle*>(h), keflag, stress, sum6);
}

int orc_npairs(void* h) { return static_cast<Oracle*>(h)->npairs; }
void orc_get_pairs(void* h, int* pair_i, int* pair_j, int* partner, double* shear) {
    Oracle& s = *static_cast<Oracle*>(h);
    for (int nc = 0; nc < s.npairs; ++nc) {
        if (pair_i) pair_i[nc] = s.pair_i[nc];
        if (pair_j) pair_j[nc] = s.pair_j[nc];
        if (partner) partner[nc] = nc < (int)s.partner.size() ? s.partner[nc] : -1;
        if (shear)
            for (int k = 0; k < 3; ++k)
                shear[3 * nc + k] = nc < (int)s.partner.size() ? s.shear[(size_t)3 * nc + k] : 0.0;
    }
}
void orc_set_history(void* h, const int* partner, const double* shear) {
    Oracle& s = *static_cast<Oracle*>(h);
    history_reset(s, s.npairs);
    for (int nc = 0; nc < s.npairs; ++nc) {
        s.partner[nc] = partner[nc];
        for (int k = 0; k < 3; ++k) s.shear[(size_t)3 * nc + k] = shear[3 * nc + k];
    }
}
void orc_get_grid(void* h, int* gdim, double* csize, int* cell_of) {
    Oracle& s = *static_cast<Oracle*>(h);
    for (int k = 0; k < 3; ++k) { gdim[k] = s.gdim[k]; csize[k] = s.csize[k]; }
    if (cell_of) for (int i = 0; i < s.n; ++i) cell_of[i] = s.cell_of[i];
}
// counters: [timestep, build_count, last_build_step, npairs, rebuilds_in_run, has_built]
void orc_get_counters(void* h, int* out) {
    Oracle& s = *static_cast<Oracle*>(h);
    out[0] = s.timestep;
    out[1] = s.build_count;
    out[2] = s.last_build_step;
    out[3] = s.npairs;
    out[4] = s.rebuilds_in_run;
    out[5] = s.has_built;
}
void orc_get_tensor(void* h, double* out6) {
    Oracle& s = *static_cast<Oracle*>(h);
    for (int k = 0; k < 6; ++k) out6[k] = s.last_tensor[k];
}

}  // extern "C"
