// Pieces shared by the one-kernel-per-step passes of both precisions: the
// XCD-aware block order and the end of the step, where the last block to
// finish reduces the virial(+kinetic) tree and runs the servo.
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

namespace tmps {

// MI355X hands workgroup w to XCD (w mod 8). Blocks of consecutive atoms share
// neighbours (and so L2 lines), so each XCD is given one contiguous range of
// blocks: a bijection of [0, nb) onto itself.
__device__ __forceinline__ int xcd_block(int w, int nb) {
    const int q = nb >> 3, r = nb & 7, x = w & 7, k = w >> 3;
    return x * q + (x < r ? x : r) + k;
}

// Device-coherent (agent scope) accesses: write-through stores and loads that
// see them, so one block can read what others published without flushing or
// invalidating the caches.
__device__ __forceinline__ void st_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Ends step t = ctl->timestep + 1 of a fused pass. Every thread of every
// block calls it once, after its own atom's work. `leaf` holds the atom's
// 7 tree leaves (6 virial+kinetic components, touching-slot count), `trig`
// whether the atom asked for a neighbor rebuild.
//
// Tree: the block's 256 leaves, then groups of 64 entries level by level
// (RedP levels), each as a pairwise tree in lane order. Blocks are 256-atom
// aligned and groups 64-entry aligned, so every partial is an aligned subtree
// of the pairwise tree over all atoms zero-padded to a power of two: the sum
// is bitwise the one the CPU oracle forms.
//
// Publication: an entry is written with agent-scope stores; its writer waits
// for them to complete before counting itself into its group, and the group's
// last arrival (the only reader) loads with agent-scope loads.
__device__ inline void finish_step(Ctl* ctl, const RedP& R, const double leaf[7], int trig, int lb,
                                   double dt, int servo_on, int natoms) {
    __shared__ double wsum[TMPS_BLOCK / 64][8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const double t = wave_tree_sum(leaf[k]);
        if (lane == 0) wsum[wave][k] = t;
    }
    const int any = __any(trig);
    if (lane == 0) wsum[wave][7] = any ? 1.0 : 0.0;
    __syncthreads();
    if (wave != 0) return;
    double e[8];
#pragma unroll
    for (int k = 0; k < 7; ++k) e[k] = (wsum[0][k] + wsum[1][k]) + (wsum[2][k] + wsum[3][k]);
    e[7] = fmax(fmax(wsum[0][7], wsum[1][7]), fmax(wsum[2][7], wsum[3][7]));

    int idx = lb, count = R.count0;
    for (int level = 0; count > 1; ++level) {
        const int g = idx >> 6;
        const int gsize = min(64, count - (g << 6));
        double* arr = R.lvl[level];
        unsigned prev = 0;
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) st_agent(arr + (size_t)idx * 8 + k, e[k]);
            __builtin_amdgcn_s_waitcnt(0);
            prev = __hip_atomic_fetch_add(R.cnt[level] + g, 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        }
        prev = (unsigned)__builtin_amdgcn_readlane((int)prev, 0);
        if (prev != (unsigned)(gsize - 1)) return;
        if (lane == 0)
            __hip_atomic_store(R.cnt[level] + g, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const size_t m = (size_t)(g << 6) + lane;
#pragma unroll
        for (int k = 0; k < 7; ++k) e[k] = wave_tree_sum(lane < gsize ? ld_agent(arr + m * 8 + k) : 0.0);
        e[7] = __any(lane < gsize && ld_agent(arr + m * 8 + 7) != 0.0) ? 1.0 : 0.0;
        idx = g;
        count = (count + 63) >> 6;
    }
    if (lane != 0) return;
    // the last block of the step
    const int t = ctl->timestep + 1;
    if (e[7] != 0.0) {
        // the skin check asked for a rebuild: the step's outputs are
        // discarded and the host redoes it after the build
        ctl->trigger = t;
        ctl->abort = 1;
        return;
    }
    for (int k = 0; k < 6; ++k) ctl->vsum[k] = e[k];
    ctl->ntouch = (int)e[6];
    if (servo_on) {
        servo_update(ctl, e, dt, natoms);
        if (!ctl->abort) update_split_consts(ctl, 0);
    }
    if (!ctl->abort) ctl->timestep = t;
}

// neigh_modify delay / every for step t (neighbor.py:371-403)
__device__ __forceinline__ bool check_due(const Ctl* ctl, int t) {
    const int since = t - ctl->last_build_step;
    return since >= ctl->delay && since % ctl->every == 0;
}

}  // namespace tmps
