// Tangent-linear restatement of one FP64 DEM run -- TEST INFRASTRUCTURE.
//
// Only tests and the benchmark's checks load it, and only as a checker of the
// device adjoint (taichimps_amd/csrc/adjoint.hip). It restates the run the
// adjoint differentiates -- Simulation.run(nsteps) with gran/hooke/history or
// gran/hertz/history, fix nve/sphere and the legacy isotropic fix
// deform/pressure (pressure on all axes, couple xyz, remap x, nevery 1) in a
// fully periodic orthogonal box -- over a scalar type T:
//
//   * T = double: the plain FP64 run, which must equal the oracle's run of the
//     same system bit for bit (dem_oracle.cpp; tests/test_tlm.py checks it);
//   * T = Dual: every value carries, beside its double (computed by the very
//     same operations, so it is the double run's value bit for bit), its
//     derivatives along ND input directions -- forward-mode differentiation,
//     i.e. the tangent-linear model of the discrete run. Every branch (contact
//     onset, Coulomb slip, skin trigger, rebuilds, periodic images) is decided
//     on the values, so the derivative is that of the trajectory with the
//     base run's branch sequence, which is what the adjoint computes too. A
//     change of the collision sequence cannot enter it, and no step size is
//     involved, unlike finite differences.
//
// The directions: dkn, dkt, dgamman, dgammat, dxmu (unit seeds) and one
// direction in the space of the radii (at fixed density: dm = 3 m / r dr, as
// the adjoint's radius gradient assumes). The outputs are the final pressure
// tensor (the servo's last virial+kinetic over volume) and its derivatives.
//
// Each function cites the reference lines it follows (paths relative to
// src/taichimps/ of the reference repository); the expression order is the
// oracle's, which follows the reference's.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int ND = 6;

struct Dual {
    double v;
    double d[ND];
    Dual() : v(0.0) { std::memset(d, 0, sizeof d); }
    Dual(double x) : v(x) { std::memset(d, 0, sizeof d); }  // NOLINT: constants
};

inline Dual operator+(const Dual& a, const Dual& b) {
    Dual r(a.v + b.v);
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
    Dual r(a.v - b.v);
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
inline Dual operator-(const Dual& a) {
    Dual r(-a.v);
    for (int k = 0; k < ND; ++k) r.d[k] = -a.d[k];
    return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
    Dual r(a.v * b.v);
    for (int k = 0; k < ND; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
    Dual r(a.v / b.v);
    for (int k = 0; k < ND; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
inline Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
inline Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }
inline Dual& operator*=(Dual& a, const Dual& b) { return a = a * b; }

inline Dual sqrt(const Dual& a) {
    Dual r(std::sqrt(a.v));
    for (int k = 0; k < ND; ++k) r.d[k] = r.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
    return r;
}
inline Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
// piecewise constant: no derivative
inline Dual floor(const Dual& a) { return Dual(std::floor(a.v)); }
inline Dual round(const Dual& a) { return Dual(std::round(a.v)); }

inline double val(double x) { return x; }
inline double val(const Dual& x) { return x.v; }
using std::fabs;
using std::floor;
using std::round;
using std::sqrt;

// Pair coefficients and run settings (as doubles; the Dual run seeds them).
struct TlmParams {
    int hertz;
    int limit_damping;
    double kn, kt, gamman, gammat, xmu;
    double skin, dt, ptarget, pgain;
};

// One pair's contribution, tallied after all pairs are evaluated.
template <class T>
struct PairOut {
    int touch;
    T F[3], ti[3], tj[3], vr[6];
};

template <class T>
struct Run {
    int n = 0;
    std::vector<T> x, v, om, f, tq, vir, radius, rmass;
    std::vector<int> tag;
    std::vector<double> x0;  // positions at the last build (the skin check)
    T lo[3], hi[3], prd[3];
    T kn, kt, gamman, gammat, xmu;
    TlmParams P{};
    // neighbor list (half list, rows ascending) and contact history per pair
    std::vector<int> pi, pj, partner;
    std::vector<T> shear;
    int has_built = 0, build_count = 0, last_build_step = -1, current_step = 0, timestep = 0;
    // servo
    T lo_start[3], hi_start[3], cum[3];
    int nsteps = 0;
    T tensor[6];
    // work buffers kept across steps
    std::vector<T> tree;
    std::vector<PairOut<T>> pout;
};

// Domain.minimum_image (orthogonal, fully periodic), domain.py:481-486, on
// the values only: the list build's test.
inline void min_image_val(const double prd[3], double d[3]) {
    for (int k = 0; k < 3; ++k) d[k] -= prd[k] * std::round(d[k] / prd[k]);
}

// Domain.pbc_wrap, orthogonal branch, domain.py:513-528
template <class T>
void pbc(Run<T>& s) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.n; ++i)
        for (int k = 0; k < 3; ++k) {
            T sl = (s.x[3 * i + k] - s.lo[k]) / s.prd[k];
            s.x[3 * i + k] = s.lo[k] + (sl - floor(sl)) * s.prd[k];
        }
}

// NeighborList.build (neighbor.py:414-435): the half list j > i within
// r_i + r_j + skin (neighbor.py:305-369), rows ascending. The set does not
// depend on the bins (bins of at least one cutoff, each visited once), so
// this restatement bins on its own grid.
template <class T>
void build_list(Run<T>& s) {
    double prd[3], lo[3];
    for (int k = 0; k < 3; ++k) {
        prd[k] = val(s.prd[k]);
        lo[k] = val(s.lo[k]);
    }
    double max_r = val(s.radius[0]);
    for (int i = 1; i < s.n; ++i) max_r = std::max(max_r, val(s.radius[i]));
    const double cut = 2.0 * max_r + s.P.skin;
    int g[3];
    for (int k = 0; k < 3; ++k) g[k] = std::max(1, (int)std::floor(prd[k] / cut));
    const int ncell = g[0] * g[1] * g[2];
    std::vector<int> cell(s.n), start(ncell + 1, 0), members(s.n);
    for (int i = 0; i < s.n; ++i) {
        int c[3];
        for (int k = 0; k < 3; ++k) {
            const double lam = (val(s.x[3 * i + k]) - lo[k]) / prd[k];
            c[k] = std::max(0, std::min((int)std::floor(lam * g[k]), g[k] - 1));
        }
        cell[i] = c[0] + g[0] * (c[1] + g[1] * c[2]);
        start[cell[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) start[c + 1] += start[c];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < s.n; ++i) members[fill[cell[i]]++] = i;
    }
    s.pi.clear();
    s.pj.clear();
    std::vector<int> row;
    for (int i = 0; i < s.n; ++i) {
        const int ci = cell[i];
        const int c0 = ci % g[0], c1 = (ci / g[0]) % g[1], c2 = ci / (g[0] * g[1]);
        int visited[27], nv = 0;
        row.clear();
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int a = ((c0 + dx) % g[0] + g[0]) % g[0];
                    const int b = ((c1 + dy) % g[1] + g[1]) % g[1];
                    const int e = ((c2 + dz) % g[2] + g[2]) % g[2];
                    const int cc = a + g[0] * (b + g[1] * e);
                    if (std::find(visited, visited + nv, cc) != visited + nv) continue;
                    visited[nv++] = cc;
                    for (int p = start[cc]; p < start[cc + 1]; ++p) {
                        const int j = members[p];
                        if (j <= i) continue;
                        double d[3];
                        for (int k = 0; k < 3; ++k) d[k] = val(s.x[3 * i + k]) - val(s.x[3 * j + k]);
                        min_image_val(prd, d);
                        const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        const double rs = val(s.radius[i]) + val(s.radius[j]) + s.P.skin;
                        if (rsq < rs * rs) row.push_back(j);
                    }
                }
        std::sort(row.begin(), row.end());
        for (int j : row) {
            s.pi.push_back(i);
            s.pj.push_back(j);
        }
    }
    for (int k = 0; k < 3 * s.n; ++k) s.x0[k] = val(s.x[k]);  // store_x0, :208-212
    s.has_built = 1;
    s.build_count += 1;
    s.last_build_step = s.current_step;
}

// NeighborList.decide + check_displacement (neighbor.py:214-245,371-403),
// delay 0 every 1 check yes: any |x - x0|^2 > skin^2 / 4.
template <class T>
bool decide(Run<T>& s) {
    if (!s.has_built) return true;
    s.current_step = s.timestep;
    const double trig = 0.25 * s.P.skin * s.P.skin;
    for (int i = 0; i < s.n; ++i) {
        const double d0 = val(s.x[3 * i]) - s.x0[3 * i];
        const double d1 = val(s.x[3 * i + 1]) - s.x0[3 * i + 1];
        const double d2 = val(s.x[3 * i + 2]) - s.x0[3 * i + 2];
        if (d0 * d0 + d1 * d1 + d2 * d2 > trig) return true;
    }
    return false;
}

// Simulation.reneighbor_if_needed (simulation.py:122-159) with the history
// carried over by partner tag (contact_history.py:91-139).
template <class T>
void reneighbor(Run<T>& s, bool force) {
    if (!decide(s) && !force) return;
    // save_state: each live pair under its first atom, keyed by partner tag
    std::vector<std::vector<std::pair<int, int>>> saved(s.n);  // (tag, old pair)
    std::vector<T> old_shear = s.shear;
    if (s.has_built)
        for (size_t nc = 0; nc < s.pi.size(); ++nc)
            if (s.partner[nc] >= 0) saved[s.pi[nc]].push_back({s.partner[nc], (int)nc});
    pbc(s);
    build_list(s);
    const size_t np = s.pi.size();
    s.partner.assign(np, -1);
    s.shear.assign(3 * np, T(0.0));
    for (size_t nc = 0; nc < np; ++nc) {
        const int jtag = s.tag[s.pj[nc]];
        for (const auto& e : saved[s.pi[nc]])
            if (e.first == jtag) {
                s.partner[nc] = jtag;
                for (int k = 0; k < 3; ++k) s.shear[3 * nc + k] = old_shear[3 * (size_t)e.second + k];
                break;
            }
    }
}

// GranHookeHistory.compute_kernel (hooke_history.py:70-195) and
// GranHertzHistory.compute_kernel (hertz_history.py:57-171), with the
// pairwise virial of Pair::ev_tally_xyz.
// The contacts are evaluated in parallel (each pair owns its history slot)
// and tallied serially in list order, so every per-atom sum runs in the
// oracle's order whatever the thread count.
template <class T>
void pair_compute(Run<T>& s, int shearupdate) {
    const bool hertz = s.P.hertz != 0;
    const double dt = s.P.dt;
    const long long np = (long long)s.pi.size();
    std::vector<PairOut<T>>& out = s.pout;
    if (out.size() < (size_t)np) out.resize((size_t)np);
#pragma omp parallel for schedule(dynamic, 1024)
    for (long long nc = 0; nc < np; ++nc) {
        const int i = s.pi[nc], j = s.pj[nc];
        const T ri = s.radius[i], rj = s.radius[j];
        const T mi = s.rmass[i], mj = s.rmass[j];
        T dpos[3];
        for (int k = 0; k < 3; ++k) dpos[k] = s.x[3 * i + k] - s.x[3 * j + k];
        for (int k = 0; k < 3; ++k) {
            const T nimg = round(dpos[k] / s.prd[k]);
            dpos[k] = dpos[k] - nimg * s.prd[k];
        }
        const T rsq = dpos[0] * dpos[0] + dpos[1] * dpos[1] + dpos[2] * dpos[2];
        const T radsum = ri + rj;
        T* sh = &s.shear[3 * nc];
        if (val(rsq) < val(radsum * radsum) && val(rsq) > 1e-28) {
            const T r = sqrt(rsq);
            const T delta = radsum - r;
            T polyhertz = 1.0;
            if (hertz) {
                const T reff = (ri * rj) / (ri + rj);
                polyhertz = sqrt(reff * delta);
            }
            const T meff = (mi * mj) / (mi + mj);
            const T rsqinv = T(1.0) / rsq;
            const T rinv = T(1.0) / r;
            T vrel[3];
            for (int k = 0; k < 3; ++k) vrel[k] = s.v[3 * i + k] - (s.v[3 * j + k] + T(0.0));
            const T vnnr = vrel[0] * dpos[0] + vrel[1] * dpos[1] + vrel[2] * dpos[2];
            T vt[3], wr[3], vtr[3], dxw[3];
            for (int k = 0; k < 3; ++k) {
                const T vn = dpos[k] * (vnnr * rsqinv);
                vt[k] = vrel[k] - vn;
                wr[k] = (ri * s.om[3 * i + k] + rj * s.om[3 * j + k]) * rinv;
            }
            dxw[0] = dpos[1] * wr[2] - dpos[2] * wr[1];
            dxw[1] = dpos[2] * wr[0] - dpos[0] * wr[2];
            dxw[2] = dpos[0] * wr[1] - dpos[1] * wr[0];
            for (int k = 0; k < 3; ++k) vtr[k] = vt[k] + dxw[k];
            const T damp = meff * s.gamman * vnnr * rsqinv;
            T ccel;
            if (hertz) ccel = (s.kn * delta * rinv - damp) * polyhertz;
            else ccel = s.kn * delta * rinv - damp;
            if (s.P.limit_damping == 1 && val(ccel) < 0.0) ccel = 0.0;

            const int jtag = s.tag[j];
            T shear[3] = {sh[0], sh[1], sh[2]};
            if (s.partner[nc] != jtag) {
                shear[0] = shear[1] = shear[2] = 0.0;
                s.partner[nc] = jtag;
            }
            const T gammat_eff = meff * s.gammat;
            if (shearupdate != 0)
                for (int k = 0; k < 3; ++k) shear[k] = shear[k] + vtr[k] * T(dt);
            const T shrmag = sqrt(shear[0] * shear[0] + shear[1] * shear[1] + shear[2] * shear[2]);
            if (shearupdate != 0) {
                const T rsht = (shear[0] * dpos[0] + shear[1] * dpos[1] + shear[2] * dpos[2]) * rsqinv;
                for (int k = 0; k < 3; ++k) shear[k] = shear[k] - rsht * dpos[k];
            }
            T fs[3];
            if (hertz) {
                for (int k = 0; k < 3; ++k)
                    fs[k] = -polyhertz * (s.kt * shear[k] + gammat_eff * vtr[k]);
            } else {
                for (int k = 0; k < 3; ++k) fs[k] = -(s.kt * shear[k] + gammat_eff * vtr[k]);
            }
            const T fs_mag = sqrt(fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]);
            const T fn_coulomb = s.xmu * fabs(ccel * r);
            if (val(fs_mag) > val(fn_coulomb)) {
                if (val(shrmag) != 0.0) {
                    const T ratio = fn_coulomb / fs_mag;
                    for (int k = 0; k < 3; ++k) {
                        const T damp_corr = gammat_eff * vtr[k] / s.kt;
                        shear[k] = ratio * (shear[k] + damp_corr) - damp_corr;
                        fs[k] = ratio * fs[k];
                    }
                } else {
                    fs[0] = fs[1] = fs[2] = 0.0;
                }
            }
            sh[0] = shear[0];
            sh[1] = shear[1];
            sh[2] = shear[2];
            T F[3];
            for (int k = 0; k < 3; ++k) F[k] = dpos[k] * ccel + fs[k];
            T c[3], tor[3], ti[3], tj[3];
            c[0] = dpos[1] * fs[2] - dpos[2] * fs[1];
            c[1] = dpos[2] * fs[0] - dpos[0] * fs[2];
            c[2] = dpos[0] * fs[1] - dpos[1] * fs[0];
            for (int k = 0; k < 3; ++k) {
                tor[k] = rinv * c[k];
                ti[k] = -ri * tor[k];
                tj[k] = -rj * tor[k];
            }
            PairOut<T>& o = out[(size_t)nc];
            o.touch = 1;
            for (int k = 0; k < 3; ++k) {
                o.F[k] = F[k];
                o.ti[k] = ti[k];
                o.tj[k] = tj[k];
            }
            o.vr[0] = T(0.5) * (dpos[0] * F[0]);
            o.vr[1] = T(0.5) * (dpos[1] * F[1]);
            o.vr[2] = T(0.5) * (dpos[2] * F[2]);
            o.vr[3] = T(0.5) * (dpos[0] * F[1]);
            o.vr[4] = T(0.5) * (dpos[0] * F[2]);
            o.vr[5] = T(0.5) * (dpos[1] * F[2]);
        } else {
            out[(size_t)nc].touch = 0;
            s.partner[nc] = -1;
            sh[0] = sh[1] = sh[2] = 0.0;
        }
    }
    for (long long nc = 0; nc < np; ++nc) {
        const PairOut<T>& o = out[(size_t)nc];
        if (!o.touch) continue;
        const int i = s.pi[nc], j = s.pj[nc];
        for (int k = 0; k < 3; ++k) {
            s.f[3 * i + k] += o.F[k];
            s.f[3 * j + k] += -o.F[k];
            s.tq[3 * i + k] += o.ti[k];
            s.tq[3 * j + k] += o.tj[k];
        }
        for (int k = 0; k < 6; ++k) {
            s.vir[6 * i + k] += o.vr[k];
            s.vir[6 * j + k] += o.vr[k];
        }
    }
}

template <class T>
void clear_forces(Run<T>& s) {
    std::fill(s.f.begin(), s.f.end(), T(0.0));
    std::fill(s.tq.begin(), s.tq.end(), T(0.0));
    std::fill(s.vir.begin(), s.vir.end(), T(0.0));
}

// FixNVESphere.initial_integrate / final_integrate, nve_sphere.py:43-96
template <class T>
void nve_initial(Run<T>& s) {
    const T dtv = s.P.dt, dtf = 0.5 * s.P.dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.n; ++i) {
        const T m = s.rmass[i], r = s.radius[i];
        const T inertia = T(0.4) * m * r * r;
        for (int k = 0; k < 3; ++k) s.v[3 * i + k] += dtf * (s.f[3 * i + k] / m);
        for (int k = 0; k < 3; ++k) s.x[3 * i + k] = s.x[3 * i + k] + dtv * s.v[3 * i + k];
        for (int k = 0; k < 3; ++k) s.om[3 * i + k] += dtf * (s.tq[3 * i + k] / inertia);
    }
}

template <class T>
void nve_final(Run<T>& s) {
    const T dtf = 0.5 * s.P.dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.n; ++i) {
        const T m = s.rmass[i], r = s.radius[i];
        const T inertia = T(0.4) * m * r * r;
        for (int k = 0; k < 3; ++k) s.v[3 * i + k] += dtf * (s.f[3 * i + k] / m);
        for (int k = 0; k < 3; ++k) s.om[3 * i + k] += dtf * (s.tq[3 * i + k] / inertia);
    }
}

// The virial(+kinetic) sum as the pairwise tree over per-atom leaves,
// zero-padded to a power of two (computes/thermo.py:62-109 leaves).
template <class T>
void virial_tensor(Run<T>& s, T out[6]) {
    size_t P = 1;
    while (P < (size_t)s.n) P <<= 1;
    std::vector<T>& buf = s.tree;
    if (buf.size() != P * 6) buf.assign(P * 6, T(0.0));
    for (size_t q = (size_t)s.n * 6; q < P * 6; ++q) buf[q] = T(0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.n; ++i) {
        T* o = &buf[(size_t)i * 6];
        for (int k = 0; k < 6; ++k) o[k] = s.vir[6 * i + k];
        const T m = s.rmass[i];
        const T* vi = &s.v[3 * i];
        o[0] = o[0] + m * vi[0] * vi[0];
        o[1] = o[1] + m * vi[1] * vi[1];
        o[2] = o[2] + m * vi[2] * vi[2];
        o[3] = o[3] + m * vi[0] * vi[1];
        o[4] = o[4] + m * vi[0] * vi[2];
        o[5] = o[5] + m * vi[1] * vi[2];
    }
    for (size_t w = 1; w < P; w <<= 1) {
        const long long pairs = (long long)(P / (2 * w));
#pragma omp parallel for schedule(static) if (pairs > 4096)
        for (long long q = 0; q < pairs; ++q) {
            const size_t i = (size_t)q * 2 * w;
            for (int k = 0; k < 6; ++k) buf[i * 6 + k] = buf[i * 6 + k] + buf[(i + w) * 6 + k];
        }
    }
    for (int k = 0; k < 6; ++k) out[k] = buf[k];
}

// FixDeformPressure.end_of_step host path for the legacy isotropic form
// (deform_pressure.py:1113-1219, _apply_pressure :833-857, couple xyz, no
// max/rate, no normalize) and remap_positions (:493-530, zero tilt).
template <class T>
int servo_end_of_step(Run<T>& s) {
    s.nsteps += 1;
    const T vol = s.prd[0] * s.prd[1] * s.prd[2];
    T vt[6], t[6];
    virial_tensor(s, vt);
    for (int k = 0; k < 6; ++k) t[k] = vt[k] / vol;
    for (int k = 0; k < 6; ++k) s.tensor[k] = t[k];
    const T scalar = (t[0] + t[1] + t[2]) / T(3.0);
    T old_prd[3], old_lo[3], lo_t[3], hi_t[3], new_prd[3];
    for (int i = 0; i < 3; ++i) {
        old_prd[i] = s.prd[i];
        old_lo[i] = s.lo[i];
    }
    const double dt = s.P.dt;
    for (int i = 0; i < 3; ++i) {
        const T rate = T(s.P.pgain) * (scalar - T(s.P.ptarget));
        s.cum[i] += T(dt) * rate * old_prd[i];
        lo_t[i] = s.lo_start[i] - T(0.5) * s.cum[i];
        hi_t[i] = s.hi_start[i] + T(0.5) * s.cum[i];
    }
    for (int i = 0; i < 3; ++i) new_prd[i] = hi_t[i] - lo_t[i];
    for (int i = 0; i < 3; ++i)
        if (val(new_prd[i]) <= 0.0) return 1;
    const T xprd = old_prd[0], yprd = old_prd[1], zprd = old_prd[2];
    const T oh0 = T(-0.0) / (xprd * yprd);
    const T oh1 = (T(0.0) * T(0.0) - yprd * T(0.0)) / (xprd * yprd * zprd);
    const T oh2 = T(-0.0) / (yprd * zprd);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.n; ++i) {
        T* xi = &s.x[3 * i];
        const T d0 = xi[0] - old_lo[0], d1 = xi[1] - old_lo[1], d2 = xi[2] - old_lo[2];
        const T l0 = d0 / old_prd[0] + oh0 * d1 + oh1 * d2;
        const T l1 = d1 / old_prd[1] + oh2 * d2;
        const T l2 = d2 / old_prd[2];
        xi[0] = new_prd[0] * l0 + T(0.0) * l1 + T(0.0) * l2 + lo_t[0];
        xi[1] = new_prd[1] * l1 + T(0.0) * l2 + lo_t[1];
        xi[2] = new_prd[2] * l2 + lo_t[2];
    }
    for (int i = 0; i < 3; ++i) {
        s.lo[i] = lo_t[i];
        s.hi[i] = hi_t[i];
        s.prd[i] = s.hi[i] - s.lo[i];
    }
    return 0;
}

// Simulation.run(nsteps) with run_pre (simulation.py:320-332): the fix's
// setup (deform_pressure.py:392-445), Verlet::setup (simulation.py:170-203:
// forced build, setup force without the history update) and the steps of
// _sub_step_inner (:205-247). Returns 0, -1 on a collapsed box.
template <class T>
int run(Run<T>& s, int nsteps) {
    s.nsteps = 0;
    for (int i = 0; i < 3; ++i) {
        s.lo_start[i] = s.lo[i];
        s.hi_start[i] = s.hi[i];
        s.cum[i] = 0.0;
    }
    clear_forces(s);
    reneighbor(s, true);
    pair_compute(s, 0);
    for (int step = 0; step < nsteps; ++step) {
        s.timestep += 1;
        nve_initial(s);
        reneighbor(s, false);
        clear_forces(s);
        pair_compute(s, 1);
        nve_final(s);
        if (servo_end_of_step(s)) return -1;
    }
    return 0;
}

template <class T>
void load(Run<T>& s, int n, const double* x, const double* v, const double* omega,
          const double* radius, const double* rmass, const int* tag, const double* lo,
          const double* hi, const TlmParams& P) {
    s.n = n;
    s.P = P;
    s.x.assign(x, x + 3 * (size_t)n);
    s.v.assign(v, v + 3 * (size_t)n);
    s.om.assign(omega, omega + 3 * (size_t)n);
    s.radius.assign(radius, radius + n);
    s.rmass.assign(rmass, rmass + n);
    s.tag.assign(tag, tag + n);
    s.f.assign(3 * (size_t)n, T(0.0));
    s.tq.assign(3 * (size_t)n, T(0.0));
    s.vir.assign(6 * (size_t)n, T(0.0));
    s.x0.assign(3 * (size_t)n, 0.0);
    for (int k = 0; k < 3; ++k) {
        s.lo[k] = lo[k];
        s.hi[k] = hi[k];
        s.prd[k] = s.hi[k] - s.lo[k];
    }
    s.kn = P.kn;
    s.kt = P.kt;
    s.gamman = P.gamman;
    s.gammat = P.gammat;
    s.xmu = P.xmu;
}

}  // namespace

extern "C" {

typedef struct {
    int hertz;          // 0 gran/hooke/history, 1 gran/hertz/history
    int limit_damping;
    double kn, kt, gamman, gammat, xmu;
    double skin, dt, ptarget, pgain;
} orc_tlm_params;

// The FP64 run (T = double): final positions and velocities (n x 3 each,
// nullable), the final pressure tensor and the neighbor build count.
int orc_tlm_forward(int n, const double* x, const double* v, const double* omega,
                    const double* radius, const double* rmass, const int* tag, const double* lo,
                    const double* hi, const orc_tlm_params* p, int nsteps, double* x_out,
                    double* v_out, double tensor[6], int* builds) {
    TlmParams P{p->hertz, p->limit_damping, p->kn, p->kt, p->gamman, p->gammat, p->xmu,
                p->skin, p->dt, p->ptarget, p->pgain};
    Run<double> s;
    load(s, n, x, v, omega, radius, rmass, tag, lo, hi, P);
    const int rc = run(s, nsteps);
    if (x_out) std::copy(s.x.begin(), s.x.end(), x_out);
    if (v_out) std::copy(s.v.begin(), s.v.end(), v_out);
    for (int k = 0; k < 6; ++k) tensor[k] = s.tensor[k];
    if (builds) *builds = s.build_count;
    return rc;
}

// The tangent-linear run (T = Dual): the final pressure tensor and its
// derivatives dtensor[6][6] (row: tensor component; column: d/dkn, d/dkt,
// d/dgamman, d/dgammat, d/dxmu, and the derivative along the radius
// direction dradius, n doubles, at fixed density; nullable for none).
int orc_tlm_tangent(int n, const double* x, const double* v, const double* omega,
                    const double* radius, const double* rmass, const int* tag, const double* lo,
                    const double* hi, const orc_tlm_params* p, int nsteps, const double* dradius,
                    double tensor[6], double dtensor[36], int* builds) {
    TlmParams P{p->hertz, p->limit_damping, p->kn, p->kt, p->gamman, p->gammat, p->xmu,
                p->skin, p->dt, p->ptarget, p->pgain};
    Run<Dual> s;
    load(s, n, x, v, omega, radius, rmass, tag, lo, hi, P);
    s.kn.d[0] = 1.0;
    s.kt.d[1] = 1.0;
    s.gamman.d[2] = 1.0;
    s.gammat.d[3] = 1.0;
    s.xmu.d[4] = 1.0;
    if (dradius)
        for (int i = 0; i < n; ++i) {
            s.radius[i].d[5] = dradius[i];
            s.rmass[i].d[5] = 3.0 * rmass[i] / radius[i] * dradius[i];
        }
    const int rc = run(s, nsteps);
    for (int k = 0; k < 6; ++k) {
        tensor[k] = s.tensor[k].v;
        for (int q = 0; q < ND; ++q) dtensor[6 * k + q] = s.tensor[k].d[q];
    }
    if (builds) *builds = s.build_count;
    return rc;
}

}  // extern "C"
