// Reverse mode (vector-Jacobian product) of the legacy history contact laws,
// gran/hooke/history (GRANULAR/hooke_history.py:107-191) and
// gran/hertz/history (GRANULAR/hertz_history.py:94-168), in FP64.
//
// legacy_history_vjp recomputes the forward quantities of
// legacy_history_contact<double, HERTZ> (no remap-v velocity shift) and then
// walks its statements backwards: given the adjoints of the outputs (force on
// a, torques on a and b, the updated shear, the pairwise virial) it adds the
// adjoints of every input (separation, velocities, spins, radii, masses, the
// stored shear) and of the five coefficients kn, kt, gamman, gammat, xmu.
// The Coulomb-slip branch is differentiated on the branch the forward took,
// including the derivative of the rescale ratio fn / |fs|; the contact test
// and the branch choices themselves are piecewise constant.
#pragma once

#include "contact.h"

namespace tmps {

// c = a x b: a_bar += b x c_bar, b_bar += c_bar x a
__device__ __forceinline__ void cross3_vjp(const double a[3], const double b[3], const double cb[3],
                                           double ab[3], double bb[3]) {
    double t[3];
    cross3(b, cb, t);
    for (int k = 0; k < 3; ++k) ab[k] += t[k];
    cross3(cb, a, t);
    for (int k = 0; k < 3; ++k) bb[k] += t[k];
}

struct PairVjpIn {
    double Fb[3];    // adjoint of the force on a (b receives -F)
    double tab[3];   // adjoint of the torque on a
    double tbb[3];   // adjoint of the torque on b
    double shb[3];   // adjoint of the updated shear
    double virb[6];  // adjoint of the pairwise virial (xx yy zz xy xz yz)
};

struct PairVjpOut {
    double db[3];             // adjoint of the separation d = x_a - x_b (minimum image)
    double vab[3], vbb[3];    // velocities
    double oab[3], obb[3];    // spins
    double rab, rbb;          // radii
    double mab, mbb;          // masses
    double shob[3];           // the stored shear (zero unless touching before)
    double pb[5];             // kn, kt, gamman, gammat, xmu
};

template <bool HERTZ>
__device__ inline void legacy_history_vjp(const double d[3], double ra, double rb, double ma,
                                          double mb, const double va[3], const double vb[3],
                                          const double oa[3], const double ob[3], double kn,
                                          double kt, double gamman, double gammat, double xmu,
                                          int limit_damping, double dt, int shearupdate,
                                          bool touching_before, const double hist[3],
                                          const PairVjpIn& in, PairVjpOut& out) {
    // ---------------- forward, as legacy_history_contact
    const double rsq = dot3(d, d);
    const double r = sqrt(rsq);
    const double radsum = ra + rb;
    const double delta = radsum - r;
    double reff = 0.0, ph = 1.0;
    if (HERTZ) {
        reff = ra * rb / (ra + rb);
        ph = sqrt(reff * delta);
    }
    const double meff = ma * mb / (ma + mb);
    const double rsqinv = 1.0 / rsq;
    const double rinv = 1.0 / r;
    double vrel[3];
    for (int k = 0; k < 3; ++k) vrel[k] = va[k] - vb[k];
    const double vnnr = dot3(vrel, d);
    const double sv = vnnr * rsqinv;
    double vt[3], sw[3], wr[3], dxw[3], vtr[3];
    for (int k = 0; k < 3; ++k) {
        vt[k] = vrel[k] - d[k] * sv;
        sw[k] = ra * oa[k] + rb * ob[k];
        wr[k] = sw[k] * rinv;
    }
    cross3(d, wr, dxw);
    for (int k = 0; k < 3; ++k) vtr[k] = vt[k] + dxw[k];
    const double damp = meff * gamman * vnnr * rsqinv;
    const double g = kn * delta * rinv - damp;
    double ccel = HERTZ ? g * ph : g;
    bool clamped = false;
    if (limit_damping == 1 && ccel < 0.0) {
        ccel = 0.0;
        clamped = true;
    }
    double sh0[3];
    for (int k = 0; k < 3; ++k) sh0[k] = touching_before ? hist[k] : 0.0;
    const double gte = meff * gammat;
    double sh1[3];
    for (int k = 0; k < 3; ++k) sh1[k] = shearupdate ? sh0[k] + vtr[k] * dt : sh0[k];
    const double shrmag = sqrt(dot3(sh1, sh1));
    double rsht = 0.0, sh2[3];
    if (shearupdate) {
        rsht = dot3(sh1, d) * rsqinv;
        for (int k = 0; k < 3; ++k) sh2[k] = sh1[k] - rsht * d[k];
    } else {
        for (int k = 0; k < 3; ++k) sh2[k] = sh1[k];
    }
    double u[3], fs0[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = kt * sh2[k] + gte * vtr[k];
        fs0[k] = -ph * u[k];
    }
    const double fsmag = sqrt(dot3(fs0, fs0));
    const double q = ccel * r;
    const double fn = xmu * fabs(q);
    const bool slip = fsmag > fn;
    const bool rescale = slip && shrmag != 0.0;
    double ratio = 1.0, dc[3] = {0.0, 0.0, 0.0}, fs[3];
    if (rescale) {
        ratio = fn / fsmag;
        for (int k = 0; k < 3; ++k) {
            dc[k] = gte * vtr[k] / kt;
            fs[k] = ratio * fs0[k];
        }
    } else if (slip) {
        fs[0] = fs[1] = fs[2] = 0.0;
    } else {
        for (int k = 0; k < 3; ++k) fs[k] = fs0[k];
    }
    double F[3];
    for (int k = 0; k < 3; ++k) F[k] = d[k] * ccel + fs[k];
    double c[3], tor[3];
    cross3(d, fs, c);
    for (int k = 0; k < 3; ++k) tor[k] = rinv * c[k];

    // ---------------- reverse
    double Fb[3] = {in.Fb[0], in.Fb[1], in.Fb[2]};
    double db[3] = {0.0, 0.0, 0.0};
    // vir = 0.5 d (x) F
    const double* vb6 = in.virb;
    Fb[0] += 0.5 * (vb6[0] * d[0]);
    Fb[1] += 0.5 * (vb6[1] * d[1] + vb6[3] * d[0]);
    Fb[2] += 0.5 * (vb6[2] * d[2] + vb6[4] * d[0] + vb6[5] * d[1]);
    db[0] += 0.5 * (vb6[0] * F[0] + vb6[3] * F[1] + vb6[4] * F[2]);
    db[1] += 0.5 * (vb6[1] * F[1] + vb6[5] * F[2]);
    db[2] += 0.5 * (vb6[2] * F[2]);
    // ta = -ra tor, tb = -rb tor
    double torb[3];
    for (int k = 0; k < 3; ++k) torb[k] = -ra * in.tab[k] - rb * in.tbb[k];
    out.rab += -dot3(in.tab, tor);
    out.rbb += -dot3(in.tbb, tor);
    double cb[3];
    for (int k = 0; k < 3; ++k) cb[k] = rinv * torb[k];
    double rinvb = dot3(torb, c);
    // c = d x fs; F = d ccel + fs
    double fsb[3] = {0.0, 0.0, 0.0};
    cross3_vjp(d, fs, cb, db, fsb);
    for (int k = 0; k < 3; ++k) {
        db[k] += ccel * Fb[k];
        fsb[k] += Fb[k];
    }
    double ccelb = dot3(d, Fb);
    // the slip branch
    double fs0b[3] = {0.0, 0.0, 0.0}, sh2b[3], gteb = 0.0, vtrb[3] = {0.0, 0.0, 0.0};
    double ktb = 0.0, fnb = 0.0, fsmagb = 0.0;
    if (rescale) {
        double sd[3];
        for (int k = 0; k < 3; ++k) sd[k] = sh2[k] + dc[k];
        const double ratiob = dot3(fsb, fs0) + dot3(in.shb, sd);
        double dcb[3];
        for (int k = 0; k < 3; ++k) {
            fs0b[k] = ratio * fsb[k];
            sh2b[k] = ratio * in.shb[k];
            dcb[k] = (ratio - 1.0) * in.shb[k];
        }
        // dc = gte vtr / kt
        gteb += dot3(dcb, vtr) / kt;
        for (int k = 0; k < 3; ++k) vtrb[k] += gte * dcb[k] / kt;
        ktb += -dot3(dcb, dc) / kt;
        // ratio = fn / fsmag
        fnb += ratiob / fsmag;
        fsmagb += -ratiob * ratio / fsmag;
    } else {
        for (int k = 0; k < 3; ++k) sh2b[k] = in.shb[k];
        if (!slip)
            for (int k = 0; k < 3; ++k) fs0b[k] = fsb[k];
    }
    // fn = xmu |ccel r|
    double rb_ = 0.0;
    if (fnb != 0.0) {
        const double sq = q > 0.0 ? 1.0 : (q < 0.0 ? -1.0 : 0.0);
        out.pb[4] += fnb * fabs(q);
        ccelb += fnb * xmu * sq * r;
        rb_ += fnb * xmu * sq * ccel;
    }
    // fsmag = |fs0|
    if (fsmagb != 0.0 && fsmag > 0.0)
        for (int k = 0; k < 3; ++k) fs0b[k] += fsmagb * fs0[k] / fsmag;
    // fs0 = -ph (kt sh2 + gte vtr)
    double phb = -dot3(fs0b, u);
    double ub[3];
    for (int k = 0; k < 3; ++k) ub[k] = -ph * fs0b[k];
    ktb += dot3(ub, sh2);
    gteb += dot3(ub, vtr);
    for (int k = 0; k < 3; ++k) {
        sh2b[k] += kt * ub[k];
        vtrb[k] += gte * ub[k];
    }
    // sh2 = sh1 - rsht d, rsht = (sh1 . d) rsqinv
    double sh1b[3] = {sh2b[0], sh2b[1], sh2b[2]};
    double rsqinvb = 0.0;
    if (shearupdate) {
        const double rshtb = -dot3(sh2b, d);
        for (int k = 0; k < 3; ++k) {
            db[k] += -rsht * sh2b[k];
            sh1b[k] += rshtb * rsqinv * d[k];
            db[k] += rshtb * rsqinv * sh1[k];
        }
        rsqinvb += rshtb * dot3(sh1, d);
    }
    // sh1 = sh0 + vtr dt
    if (shearupdate)
        for (int k = 0; k < 3; ++k) vtrb[k] += dt * sh1b[k];
    if (touching_before)
        for (int k = 0; k < 3; ++k) out.shob[k] += sh1b[k];
    // gte = meff gammat
    double meffb = gteb * gammat;
    out.pb[3] += gteb * meff;
    // ccel = g ph (hertz) or g, clamped at zero under limit_damping
    double gb = 0.0;
    if (!clamped) {
        gb = HERTZ ? ccelb * ph : ccelb;
        if (HERTZ) phb += ccelb * g;
    }
    out.pb[0] += gb * delta * rinv;
    double deltab = gb * kn * rinv;
    rinvb += gb * kn * delta;
    const double dampb = -gb;
    // damp = meff gamman vnnr rsqinv
    meffb += dampb * gamman * vnnr * rsqinv;
    out.pb[2] += dampb * meff * vnnr * rsqinv;
    double vnnrb = dampb * meff * gamman * rsqinv;
    rsqinvb += dampb * meff * gamman * vnnr;
    // vtr = vt + d x wr
    double wrb[3] = {0.0, 0.0, 0.0};
    cross3_vjp(d, wr, vtrb, db, wrb);
    // wr = (ra oa + rb ob) rinv
    for (int k = 0; k < 3; ++k) {
        const double swb = rinv * wrb[k];
        rinvb += wrb[k] * sw[k];
        out.rab += swb * oa[k];
        out.oab[k] += ra * swb;
        out.rbb += swb * ob[k];
        out.obb[k] += rb * swb;
    }
    // vt = vrel - d sv, sv = vnnr rsqinv
    double vrelb[3];
    double svb = 0.0;
    for (int k = 0; k < 3; ++k) {
        vrelb[k] = vtrb[k];
        db[k] += -sv * vtrb[k];
        svb += -d[k] * vtrb[k];
    }
    vnnrb += svb * rsqinv;
    rsqinvb += svb * vnnr;
    // vnnr = vrel . d
    for (int k = 0; k < 3; ++k) {
        vrelb[k] += vnnrb * d[k];
        db[k] += vnnrb * vrel[k];
    }
    for (int k = 0; k < 3; ++k) {
        out.vab[k] += vrelb[k];
        out.vbb[k] -= vrelb[k];
    }
    // meff = ma mb / (ma + mb)
    {
        const double s = 1.0 / (ma + mb);
        out.mab += meffb * (mb * s) * (mb * s);
        out.mbb += meffb * (ma * s) * (ma * s);
    }
    // ph = sqrt(reff delta), reff = ra rb / (ra + rb)
    if (HERTZ && ph > 0.0) {
        const double reffb = phb * delta / (2.0 * ph);
        deltab += phb * reff / (2.0 * ph);
        const double s = 1.0 / (ra + rb);
        out.rab += reffb * (rb * s) * (rb * s);
        out.rbb += reffb * (ra * s) * (ra * s);
    }
    // delta = ra + rb - r
    out.rab += deltab;
    out.rbb += deltab;
    rb_ += -deltab;
    // rinv = 1/r, rsqinv = 1/rsq, r = sqrt(rsq), rsq = d . d
    rb_ += -rinvb * rinv * rinv;
    double rsqb = -rsqinvb * rsqinv * rsqinv;
    rsqb += rb_ / (2.0 * r);
    for (int k = 0; k < 3; ++k) db[k] += 2.0 * rsqb * d[k];
    out.pb[1] += ktb;
    for (int k = 0; k < 3; ++k) out.db[k] += db[k];
}

}  // namespace tmps
