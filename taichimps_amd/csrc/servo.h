// The deform/pressure servo's per-step update, shared by the FP64 and
// FP32-split step kernels. It runs in FP64 in both, with every expression
// rounded as written (no contraction), so the box the two precisions see
// follows the same arithmetic as the CPU oracle.
#pragma once

#include "detmath.h"
#include "tmps_internal.h"

namespace tmps {

// _limit_rate_device, deform_pressure.py:582-596
__device__ inline double limit_rate(const tmps_servo_params& S, double rate, double ptarget) {
#pragma clang fp contract(off)
    double out = rate;
    if (S.normalize == 1) {
        if (ptarget == 0.0) {
            if (out != 0.0) out = S.max_rate * (out > 0.0 ? 1.0 : -1.0);
        } else {
            out = out / fabs(ptarget);
        }
    }
    if (S.max_rate != 0.0 && fabs(out) > S.max_rate) out = S.max_rate * (out > 0.0 ? 1.0 : -1.0);
    return out;
}

// FixDeformPressure.end_of_step for x/y/z styles (deform_pressure.py:1113-1219
// host path, :598-702 device path): pressure tensor from the reduced
// virial(+kinetic) sum, couple, per-axis targets, remap record, box publish.
__device__ inline void servo_update(Ctl* ctl, const double vs[6], double dt, int natoms) {
#pragma clang fp contract(off)
    // work on register copies: the control block is global memory, and a
    // chain of dependent loads and stores on it costs microseconds
    DevServo SV = ctl->servo;
    const tmps_servo_params& S = SV.p;
    DevBox B = ctl->box;
    SV.step_count += 1;
    SV.nsteps += 1;
    ctl->remap.step = -1;
    if (SV.step_count % S.nevery != 0) {
        ctl->servo.step_count = SV.step_count;
        ctl->servo.nsteps = SV.nsteps;
        return;
    }

    bool needs_pressure = false;
    for (int i = 0; i < 3; ++i)
        if (S.style[i] == 8 || S.style[i] == 9) needs_pressure = true;
    double p[3] = {0.0, 0.0, 0.0};
    if (needs_pressure) {
        const double vol = B.prd[0] * B.prd[1] * B.prd[2];
        double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (natoms > 0 && vol > 0.0)
            for (int k = 0; k < 6; ++k) t[k] = vs[k] / vol;
        for (int k = 0; k < 6; ++k) ctl->tensor[k] = t[k];
        const double scalar = (t[0] + t[1] + t[2]) / 3.0;
        if (S.couple == 1) {
            p[0] = p[1] = p[2] = scalar;
        } else if (S.couple >= 2) {
            int a = 0, b = 1;
            if (S.couple == 3) { a = 1; b = 2; }
            if (S.couple == 4) { a = 0; b = 2; }
            const double mean = (t[a] + t[b]) / 2.0;
            for (int i = 0; i < 3; ++i) p[i] = t[i];
            p[a] = mean;
            p[b] = mean;
        } else {
            for (int i = 0; i < 3; ++i) {
                if (S.style[i] == 8) p[i] = t[i];
                else if (S.style[i] == 9) p[i] = scalar;
            }
        }
    }
    double old_prd[3], old_lo[3], lo_t[3], hi_t[3];
    for (int i = 0; i < 3; ++i) {
        old_prd[i] = B.prd[i];
        old_lo[i] = B.lo[i];
        lo_t[i] = B.lo[i];
        hi_t[i] = B.hi[i];
    }
    const double elapsed = SV.nsteps * dt;
    const double frac = SV.nsteps_total > 0 ? (double)SV.nsteps / (double)SV.nsteps_total : 0.0;
    for (int i = 0; i < 3; ++i) {
        const double ls = SV.lo_start[i], hs = SV.hi_start[i];
        switch (S.style[i]) {
            case 2: {  // trate
                const double mid = 0.5 * (ls + hs);
                const double half = 0.5 * (hs - ls) * det_exp(S.rate[i] * elapsed);
                lo_t[i] = mid - half; hi_t[i] = mid + half;
            } break;
            case 1: {  // erate
                const double shift = 0.5 * elapsed * S.rate[i] * (hs - ls);
                lo_t[i] = ls - shift; hi_t[i] = hs + shift;
            } break;
            case 3: {  // vel
                const double shift = 0.5 * elapsed * S.vel[i];
                lo_t[i] = ls - shift; hi_t[i] = hs + shift;
            } break;
            case 4: {  // wiggle
                const double shift =
                    0.5 * S.amp[i] * det_sin(2.0 * 3.14159265358979323846 * elapsed / S.tper[i]);
                lo_t[i] = ls - shift; hi_t[i] = hs + shift;
            } break;
            case 5:
            case 6:
            case 7: {  // final, delta, scale
                double lo_stop, hi_stop;
                if (S.style[i] == 5) { lo_stop = S.flo[i]; hi_stop = S.fhi[i]; }
                else if (S.style[i] == 6) { lo_stop = ls + S.dlo[i]; hi_stop = hs + S.dhi[i]; }
                else {
                    const double mid = 0.5 * (ls + hs);
                    const double half = 0.5 * S.scale[i] * (hs - ls);
                    lo_stop = mid - half; hi_stop = mid + half;
                }
                lo_t[i] = ls + frac * (lo_stop - ls);
                hi_t[i] = hs + frac * (hi_stop - hs);
            } break;
            case 8:
            case 9: {  // pressure, pressure/mean
                const double rate = limit_rate(S, S.pgain[i] * (p[i] - S.ptarget[i]), S.ptarget[i]);
                SV.cum[i] += dt * rate * old_prd[i];
                lo_t[i] = ls - 0.5 * SV.cum[i];
                hi_t[i] = hs + 0.5 * SV.cum[i];
            } break;
            default: break;
        }
    }
    double new_prd[3];
    for (int i = 0; i < 3; ++i) new_prd[i] = hi_t[i] - lo_t[i];
    ctl->servo.step_count = SV.step_count;
    ctl->servo.nsteps = SV.nsteps;
    for (int i = 0; i < 3; ++i) ctl->servo.cum[i] = SV.cum[i];
    for (int i = 0; i < 3; ++i)
        if (new_prd[i] <= 0.0) {
            ctl->error = TMPS_ERR_BOX;
            ctl->abort = 1;
            return;
        }
    if (S.remap == 0) {
        RemapRec R;
        const double xprd = old_prd[0], yprd = old_prd[1], zprd = old_prd[2];
        R.oh[0] = -0.0 / (xprd * yprd);
        R.oh[1] = (0.0 * 0.0 - yprd * 0.0) / (xprd * yprd * zprd);
        R.oh[2] = -0.0 / (yprd * zprd);
        for (int i = 0; i < 3; ++i) {
            R.old_lo[i] = old_lo[i];
            R.old_prd[i] = old_prd[i];
            R.new_lo[i] = lo_t[i];
            R.new_prd[i] = new_prd[i];
        }
        R.step = ctl->timestep + 1;
        R.pad = 0;
        ctl->remap = R;
    }
    for (int i = 0; i < 3; ++i) {
        B.h_rate[i] = (new_prd[i] - old_prd[i]) / dt;
        B.lo[i] = lo_t[i];
        B.hi[i] = hi_t[i];
        B.prd[i] = B.hi[i] - B.lo[i];
    }
    B.h_rate[3] = B.h_rate[4] = B.h_rate[5] = (0.0 - 0.0) / dt;
    B.vremap = S.remap == 1 ? 1 : 0;
    ctl->box = B;
}

}  // namespace tmps
