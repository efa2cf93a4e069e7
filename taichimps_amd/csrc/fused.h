// Pieces shared by the fused step passes of both precisions: the XCD-aware
// block order, the publication of each block's tree partial, and the serial
// part of the one-block close (k_step_close, k_batch_close) that reduces
// them and runs the servo.
#pragma once

#include "contact.h"
#include "servo.h"
#include "tmps_internal.h"

// Waves per SIMD the step kernels are compiled for. They are bound by the
// latency of neighbour gathers, so occupancy counts; the FP64 pass needs twice
// the registers per value and is held where its gather loop does not spill.
#ifndef TMPS_STEP_WAVES_F64
#define TMPS_STEP_WAVES_F64 3
#endif
#ifndef TMPS_STEP_WAVES_SPLIT
#define TMPS_STEP_WAVES_SPLIT 5
#endif
// The FP32-split pass with walls holds the FP64 wall contact beside its own
// state: at 5 waves it spills (scratch stores on every atom's path), at 4 it
// does not.
#ifndef TMPS_STEP_WAVES_SPLIT_WALL
#define TMPS_STEP_WAVES_SPLIT_WALL 4
#endif

namespace tmps {


// A block's share of a fused step pass: the pairwise-tree partial of its 256
// atoms for the 7 tree leaves (6 virial+kinetic components, touching-slot
// count) and, as an 8th component, how many of its atoms asked for a neighbor
// rebuild; thread k < 8 receives component k. Blocks are 256-atom aligned, so
// every partial is an aligned subtree of the pairwise tree over all atoms
// zero-padded to a power of two.
__device__ __forceinline__ double block_partial(const double leaf[7], int trig) {
    double l8[8];
#pragma unroll
    for (int k = 0; k < 7; ++k) l8[k] = leaf[k];
    l8[7] = trig ? 1.0 : 0.0;
    return block_sum8(l8);
}

// The block's partial as entry `lb` of `partial` (8 doubles), for a separate
// close kernel (k_step_close, k_batch_close) whose tree over the entries is
// bitwise the sum the CPU oracle forms. Plain stores: the kernel boundary
// publishes them.
__device__ __forceinline__ void publish_block(const double leaf[7], int trig, int lb,
                                              double* partial) {
    const double v = block_partial(leaf, trig);
    if (threadIdx.x < 8) partial[(size_t)lb * 8 + threadIdx.x] = v;
}

// The serial part of a step's close on `cs`, a copy of the control block,
// given the step's tree sums `out` (6 virial+kinetic components, touching
// slots, rebuild requests). A step whose atoms asked for a rebuild records
// the trigger and the abort that skips the rest of its segment; otherwise the
// sums are recorded, the servo runs and the step advances (an ensemble
// `member` also flips its buffer set and notes the end of its run). Returns
// whether `cs` is to be written back (not for a segment already aborted).
__device__ __forceinline__ bool close_update(Ctl& cs, const double out[8], double dt,
                                             int servo_on, int natoms, bool member) {
    if (cs.abort) return false;
    const int t = cs.timestep + 1;
    if (out[7] != 0.0) {
        // the skin check asked for a rebuild: the step's outputs are
        // discarded and the host redoes it after the build
        cs.trigger = t;
        cs.abort = 1;
        return true;
    }
    for (int k = 0; k < 6; ++k) cs.vsum[k] = out[k];
    cs.ntouch = (int)out[6];
    if (servo_on) {
        servo_update(&cs, out, dt, natoms);
        if (!cs.abort) update_split_consts(&cs, 0);
    }
    if (!cs.abort) {
        cs.timestep = t;
        if (member) {
            // an ensemble member: the next step reads the other set; its run
            // ends with step bend
            cs.bpar ^= 1;
            cs.bkick = 0;
            if (t == cs.bend) cs.bdone = 1;
        }
    }
    return true;
}

// neigh_modify delay / every for step t (neighbor.py:371-403)
__device__ __forceinline__ bool check_due(const Ctl* ctl, int t) {
    const int since = t - ctl->last_build_step;
    return since >= ctl->delay && since % ctl->every == 0;
}

}  // namespace tmps
