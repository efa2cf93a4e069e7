// Pieces shared by the fused step passes of both precisions: the XCD-aware
// block order and the publication of each block's tree partial, which the
// one-block k_step_close reduces before it runs the servo.
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

// Ends a block's share of a fused step pass: the pairwise-tree partial of the
// block's 256 atoms for the 7 tree leaves (6 virial+kinetic components,
// touching-slot count) and, as an 8th component, how many of its atoms asked
// for a neighbor rebuild, written to entry `lb` of `partial` (8 doubles).
// Blocks are 256-atom aligned, so every entry is an aligned subtree of the
// pairwise tree over all atoms zero-padded to a power of two, and
// k_step_close's tree over the entries is bitwise the sum the CPU oracle
// forms. Plain stores: the kernel boundary publishes them.
__device__ __forceinline__ void publish_block(const double leaf[7], int trig, int lb,
                                              double* partial) {
    double l8[8];
#pragma unroll
    for (int k = 0; k < 7; ++k) l8[k] = leaf[k];
    l8[7] = trig ? 1.0 : 0.0;
    block_tree_sum_lds<8>(l8, partial + (size_t)lb * 8);
}

// neigh_modify delay / every for step t (neighbor.py:371-403)
__device__ __forceinline__ bool check_due(const Ctl* ctl, int t) {
    const int since = t - ctl->last_build_step;
    return since >= ctl->delay && since % ctl->every == 0;
}

}  // namespace tmps
