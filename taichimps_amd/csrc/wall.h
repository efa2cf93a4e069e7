// The wall contacts of fix wall/gran and fix wall/gran/region on one atom,
// shared by the per-stage wall kernels and the one-kernel steps. The wall
// is a partner of zero radius and infinite mass; every expression rounds as
// written (no contraction), whichever unit includes this, so a wall gives
// the same bits in a stage kernel and in a fused step. Each returns whether
// the atom touches the wall (an atom out of contact keeps its force as it
// is).
#pragma once

#include "tmps_internal.h"

namespace tmps {

// FixWallGran.post_force_kernel, wall_gran.py:120-211: hooke normal force with
// mass-velocity damping and the linear_history_classic tangential force, or
// the history-free viscous one. dF / dT receive the force and torque added to
// the atom (zero out of contact); with history, the atom's shear (xyz) and
// touch flag (w) are read through h_prev (null: a zero history, read only in
// contact) and the new ones returned in h_new (zero out of contact).
__device__ inline bool wall_plane_contact(const tmps_wall_params& P, double coord,
                                          double vwall_n, double vwall_s, const double xs[3],
                                          double ri, double meff, const double v[3],
                                          const double w[3], double dt, int history_update,
                                          const double4* h_prev, double4& h_new, double dF[3],
                                          double dT[3]) {
#pragma clang fp contract(off)
    dF[0] = dF[1] = dF[2] = 0.0;
    dT[0] = dT[1] = dT[2] = 0.0;
    h_new = make_double4(0.0, 0.0, 0.0, 0.0);
    const int axis = P.axis;
    double nv[3] = {0.0, 0.0, 0.0};
    nv[axis] = P.side == 1 ? -1.0 : 1.0;
    double vwall[3] = {0.0, 0.0, 0.0};
    vwall[axis] = vwall_n;
    if (P.shear_axis >= 0) vwall[P.shear_axis] = vwall_s;
    const double xa = xs[axis];
    const double dist = P.side == -1 ? xa - coord : coord - xa;
    const double delta = ri - dist;
    if (delta <= 0.0) return false;
    double vr[3], vt[3], wr[3], vtr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) vr[k] = v[k] - vwall[k];
    const double vnnr = vr[0] * nv[0] + vr[1] * nv[1] + vr[2] * nv[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) vt[k] = vr[k] - vnnr * nv[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) wr[k] = ri * w[k];
    const double wxn[3] = {wr[1] * nv[2] - wr[2] * nv[1], wr[2] * nv[0] - wr[0] * nv[2],
                           wr[0] * nv[1] - wr[1] * nv[0]};
#pragma unroll
    for (int k = 0; k < 3; ++k) vtr[k] = vt[k] - wxn[k];
    const double vrel = sqrt(vtr[0] * vtr[0] + vtr[1] * vtr[1] + vtr[2] * vtr[2]);
    const double damp_prefactor = meff * P.gamman;
    double fntot = P.kn * delta - damp_prefactor * vnnr;
    if (P.limit_damping) fntot = fntot < 0.0 ? 0.0 : fntot;  // std::max(fntot, 0.0)
    const double fscrit = P.xmu * fabs(fntot);
    const double damp_t = meff * P.gammat;
    double fs[3] = {0.0, 0.0, 0.0};
    if (P.use_history) {
        const double4 hh = h_prev ? *h_prev : make_double4(0.0, 0.0, 0.0, 0.0);
        double h[3] = {hh.x, hh.y, hh.z};
        if (hh.w == 0.0) h[0] = h[1] = h[2] = 0.0;
        if (history_update)
#pragma unroll
            for (int k = 0; k < 3; ++k) h[k] = h[k] + vtr[k] * dt;
        const double shrmag = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        if (history_update) {
            const double hn = h[0] * nv[0] + h[1] * nv[1] + h[2] * nv[2];
#pragma unroll
            for (int k = 0; k < 3; ++k) h[k] = h[k] - hn * nv[k];
        }
        double fdamp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) fdamp[k] = -damp_t * vtr[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) fs[k] = -P.kt * h[k] + fdamp[k];
        const double magfs = sqrt(fs[0] * fs[0] + fs[1] * fs[1] + fs[2] * fs[2]);
        if (magfs > fscrit) {
            if (shrmag != 0.0) {
                const double sc = fscrit / magfs;
#pragma unroll
                for (int k = 0; k < 3; ++k) fs[k] = fs[k] * sc;
#pragma unroll
                for (int k = 0; k < 3; ++k) h[k] = -(fs[k] - fdamp[k]) / P.kt;
            } else {
                fs[0] = fs[1] = fs[2] = 0.0;
            }
        }
        h_new = make_double4(h[0], h[1], h[2], 1.0);
    } else {
        double ft = 0.0;
        if (vrel != 0.0) {
            const double dv = damp_t * vrel;
            ft = (dv < fscrit ? dv : fscrit) / vrel;  // std::min(fscrit, dv)
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) fs[k] = -ft * vtr[k];
    }
    dF[0] = fntot * nv[0] + fs[0];
    dF[1] = fntot * nv[1] + fs[1];
    dF[2] = fntot * nv[2] + fs[2];
    dT[0] = -ri * (nv[1] * fs[2] - nv[2] * fs[1]);
    dT[1] = -ri * (nv[2] * fs[0] - nv[0] * fs[2]);
    dT[2] = -ri * (nv[0] * fs[1] - nv[1] * fs[0]);
    return true;
}

// FixWallGranRegion.post_force_kernel, wall_gran_region.py:114-271: the
// curved side of a cylinder region. Hooke (mass_velocity damping) or hertz
// (viscoelastic: the contact radius sqrt(r delta) scales stiffness and
// damping), linear_history_classic or history-free tangential force; atoms
// outside [axis_lo, axis_hi] (when bounded) or on the axis are out of contact.
// The history is read and returned as wall_plane_contact's.
__device__ inline bool wall_cyl_contact(const tmps_wall_params& P, const double pos[3],
                                        double r_i, double mi, const double vi[3],
                                        const double oi[3], double dt, int history_update,
                                        const double4* h_prev, double4& h_new, double dF[3],
                                        double dT[3]) {
#pragma clang fp contract(off)
    dF[0] = dF[1] = dF[2] = 0.0;
    dT[0] = dT[1] = dT[2] = 0.0;
    h_new = make_double4(0.0, 0.0, 0.0, 0.0);
    const int axis = P.axis;
    const double pos_axis = pos[axis];
    if (P.has_bounds && (pos_axis < P.axis_lo || pos_axis > P.axis_hi)) return false;
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const double d1 = pos[a1] - P.c1, d2 = pos[a2] - P.c2;
    const double r_perp = sqrt(d1 * d1 + d2 * d2);
    if (!(r_perp > 1e-14)) return false;
    const double inv_r_perp = 1.0 / r_perp;
    const double u1 = d1 * inv_r_perp, u2 = d2 * inv_r_perp;
    double nw[3] = {0.0, 0.0, 0.0};
    nw[a1] = P.side == -1 ? -u1 : u1;
    nw[a2] = P.side == -1 ? -u2 : u2;
    const double delta = P.side == -1 ? r_i - (P.radius - r_perp) : r_i - (r_perp - P.radius);
    if (!(delta > 0.0)) return false;
    const double oxn[3] = {oi[1] * nw[2] - oi[2] * nw[1], oi[2] * nw[0] - oi[0] * nw[2],
                           oi[0] * nw[1] - oi[1] * nw[0]};
    double vr[3], vrt[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) vr[k] = vi[k] - r_i * oxn[k];
    const double vn = vr[0] * nw[0] + vr[1] * nw[1] + vr[2] * nw[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) vrt[k] = vr[k] - vn * nw[k];
    const double vrel = sqrt(vrt[0] * vrt[0] + vrt[1] * vrt[1] + vrt[2] * vrt[2]);
    double poly = 1.0;
    if (P.hertz) poly = sqrt(r_i * delta);
    const double fn_elastic = P.kn * delta * poly;
    const double damp_prefactor = mi * P.gamman * poly;
    double fntot = fn_elastic - damp_prefactor * vn;
    if (P.limit_damping) fntot = fntot < 0.0 ? 0.0 : fntot;  // std::max(fntot, 0.0)
    const double fscrit = P.xmu * fabs(fntot);
    const double damp_t = mi * P.gammat * poly;
    double ft[3] = {0.0, 0.0, 0.0};
    if (P.use_history) {
        const double4 hh = h_prev ? *h_prev : make_double4(0.0, 0.0, 0.0, 0.0);
        double h[3] = {hh.x, hh.y, hh.z};
        if (hh.w == 0.0) h[0] = h[1] = h[2] = 0.0;
        if (history_update)
#pragma unroll
            for (int k = 0; k < 3; ++k) h[k] = h[k] + vrt[k] * dt;
        const double shrmag = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        if (history_update) {
            const double hn = h[0] * nw[0] + h[1] * nw[1] + h[2] * nw[2];
#pragma unroll
            for (int k = 0; k < 3; ++k) h[k] = h[k] - hn * nw[k];
        }
        double fdamp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) fdamp[k] = -damp_t * vrt[k];
        const double ktp = -P.kt * poly;
#pragma unroll
        for (int k = 0; k < 3; ++k) ft[k] = ktp * h[k] + fdamp[k];
        const double magfs = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
        if (magfs > fscrit) {
            if (shrmag != 0.0) {
                const double sc = fscrit / magfs;
                const double kp = P.kt * poly;
#pragma unroll
                for (int k = 0; k < 3; ++k) ft[k] = ft[k] * sc;
#pragma unroll
                for (int k = 0; k < 3; ++k) h[k] = -(ft[k] - fdamp[k]) / kp;
            } else {
                ft[0] = ft[1] = ft[2] = 0.0;
            }
        }
        h_new = make_double4(h[0], h[1], h[2], 1.0);
    } else {
        double fsc = 0.0;
        if (vrel != 0.0) {
            const double dv = damp_t * vrel;
            fsc = (dv < fscrit ? dv : fscrit) / vrel;  // std::min(fscrit, dv)
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) ft[k] = -fsc * vrt[k];
    }
    dF[0] = fntot * nw[0] + ft[0];
    dF[1] = fntot * nw[1] + ft[1];
    dF[2] = fntot * nw[2] + ft[2];
    dT[0] = -r_i * (nw[1] * ft[2] - nw[2] * ft[1]);
    dT[1] = -r_i * (nw[2] * ft[0] - nw[0] * ft[2]);
    dT[2] = -r_i * (nw[0] * ft[1] - nw[1] * ft[0]);
    return true;
}

// One wall of a step on one atom (geometry from its settings).
__device__ inline bool wall_contact(const WallSpec& S, const double xs[3], double ri,
                                    double meff, const double v[3], const double w[3], double dt,
                                    int history_update, const double4* h_prev, double4& h_new,
                                    double dF[3], double dT[3]) {
    if (S.p.geometry == TMPS_WALL_CYLINDER)
        return wall_cyl_contact(S.p, xs, ri, meff, v, w, dt, history_update, h_prev, h_new, dF,
                                dT);
    return wall_plane_contact(S.p, S.coord, S.vwall_n, S.vwall_s, xs, ri, meff, v, w, dt,
                              history_update, h_prev, h_new, dF, dT);
}

// Whether the atom touches the wall: the contact functions' own test, in the
// same arithmetic, without the force. A step runs it on every atom and the
// contact (whose FP64 temporaries would otherwise be live in every atom's
// path) only on those it admits.
__device__ inline bool wall_touches(const WallSpec& S, const double xs[3], double ri) {
#pragma clang fp contract(off)
    const tmps_wall_params& P = S.p;
    if (P.geometry != TMPS_WALL_CYLINDER) {
        const double xa = xs[P.axis];
        const double dist = P.side == -1 ? xa - S.coord : S.coord - xa;
        return !(ri - dist <= 0.0);
    }
    const double pos_axis = xs[P.axis];
    if (P.has_bounds && (pos_axis < P.axis_lo || pos_axis > P.axis_hi)) return false;
    const int a1 = P.axis == 0 ? 1 : 0, a2 = P.axis == 2 ? 1 : 2;
    const double d1 = xs[a1] - P.c1, d2 = xs[a2] - P.c2;
    const double r_perp = sqrt(d1 * d1 + d2 * d2);
    if (!(r_perp > 1e-14)) return false;
    const double delta = P.side == -1 ? ri - (P.radius - r_perp) : ri - (r_perp - P.radius);
    return delta > 0.0;
}

// Adds a fused step's walls that come before (pass 0) or after (pass 1)
// damping/cundall to the atom's force and torque, F / T in FP64 or FP32 (an
// FP32 sum takes the FP64 increment and rounds once, as the stage kernels'
// f[i] + increment does). The histories are dense (a zero entry for an atom
// out of contact), but only the entries flagged possibly nonzero in W.nz are
// read or rewritten: an atom out of contact whose out-set entry is already
// zero costs one byte read, not a 32-byte store per wall.
template <typename R>
__device__ inline void walls_apply(const WallRefs& W, int pass, const double xs[3], double ri,
                                   double meff, const double v[3], const double w[3], double dt,
                                   int i, R F[3], R T[3]) {
#pragma clang fp contract(off)
    if (W.nwall == 0) return;
    const unsigned nz0 = W.nz[i];
    unsigned nz = nz0;
    for (int q = 0; q < W.nwall; ++q) {
        const WallSpec& S = W.spec[q];
        if ((S.before_cundall ? 0 : 1) != pass) continue;
        const unsigned bin = wall_nz_bit(q, W.in), bout = wall_nz_bit(q, W.out);
        double dF[3], dT[3];
        double4 hn = make_double4(0.0, 0.0, 0.0, 0.0);
        const bool touched =
            wall_touches(S, xs, ri) &&
            wall_contact(S, xs, ri, meff, v, w, dt, 1, (nz & bin) ? W.h_in[q] + i : nullptr, hn,
                         dF, dT);
        if (S.p.use_history) {
            if (touched) {
                W.h_out[q][i] = hn;
                nz |= bout;
            } else if (nz & bout) {
                W.h_out[q][i] = make_double4(0.0, 0.0, 0.0, 0.0);
                nz &= ~bout;
            }
        }
        if (!touched) continue;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            F[k] = (R)((double)F[k] + dF[k]);
            T[k] = (R)((double)T[k] + dT[k]);
        }
    }
    if (nz != nz0) W.nz[i] = (uint8_t)nz;
}

}  // namespace tmps
