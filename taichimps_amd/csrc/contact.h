// Per-contact laws, evaluated for the pair (a, b) with a the lower atom index.
//
// Each function restates one reference kernel body for a single contact; the
// caller gathers the two atoms, reduces the separation to its minimum image
// and tests for contact. Evaluating every pair in this one orientation makes
// the result for atom b the exact negation of the result for atom a, so a full
// neighbor list with a per-atom gather needs no atomics and stays bitwise
// equal to a half-list pass that adds F to a and -F to b.
//
// Written once over the scalar type T: T = double is the FP64 path (compiled
// with -ffp-contract=off, every operation rounds as written), T = float the
// FP32-split path.
#pragma once

#include <hip/hip_runtime.h>

#include "tmps_internal.h"

namespace tmps {

template <typename T>
__device__ __forceinline__ T dot3(const T a[3], const T b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <typename T>
__device__ __forceinline__ void cross3(const T a[3], const T b[3], T o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__device__ __forceinline__ T tsqrt(T v) { return sqrt(v); }
template <>
__device__ __forceinline__ float tsqrt<float>(float v) { return sqrtf(v); }

// a / b: correctly rounded in FP64 (the parity path); in FP32 the hardware
// reciprocal (about 1 ulp), which the FP32-split error budget absorbs.
template <typename T>
__device__ __forceinline__ T tdiv(T a, T b) { return a / b; }
template <>
__device__ __forceinline__ float tdiv<float>(float a, float b) { return __fdividef(a, b); }

template <typename T>
struct ContactOut {
    T F[3];    // force on a (b receives -F)
    T ta[3];   // torque on a
    T tb[3];   // torque on b
    T vir[6];  // pairwise virial, added to both atoms
};

// gran/hooke/history (GRANULAR/hooke_history.py:107-191) and
// gran/hertz/history (GRANULAR/hertz_history.py:94-168). `hist` holds the
// slot's shear when `touching_before`, and receives the updated shear.
template <typename T, bool HERTZ>
__device__ __forceinline__ void legacy_history_contact(
    const T dpos[3], const T dvj[3], T rsq, T radsum, T ri, T rj, T mi, T mj, const T vi[3],
    const T vj[3], const T oi[3], const T oj[3], T kn, T kt, T gamman, T gammat, T xmu,
    int limit_damping, T dt, int shearupdate, bool touching_before, T hist[3],
    ContactOut<T>& o) {
    const T r = tsqrt(rsq);
    const T delta = radsum - r;
    T polyhertz = T(1);
    if (HERTZ) {
        const T reff = tdiv(ri * rj, ri + rj);
        polyhertz = tsqrt(reff * delta);
    }
    const T meff = tdiv(mi * mj, mi + mj);
    const T rsqinv = tdiv(T(1), rsq);
    const T rinv = tdiv(T(1), r);
    T vrel[3];
    for (int k = 0; k < 3; ++k) vrel[k] = vi[k] - (vj[k] + dvj[k]);
    const T vnnr = dot3(vrel, dpos);
    T vt[3], wr[3], dxw[3], vtr[3];
    for (int k = 0; k < 3; ++k) {
        const T vn = dpos[k] * (vnnr * rsqinv);
        vt[k] = vrel[k] - vn;
        wr[k] = (ri * oi[k] + rj * oj[k]) * rinv;
    }
    cross3(dpos, wr, dxw);
    for (int k = 0; k < 3; ++k) vtr[k] = vt[k] + dxw[k];
    const T damp = meff * gamman * vnnr * rsqinv;
    T ccel;
    if (HERTZ) ccel = (kn * delta * rinv - damp) * polyhertz;
    else ccel = kn * delta * rinv - damp;
    if (limit_damping == 1 && ccel < T(0)) ccel = T(0);

    T shear[3];
    if (touching_before) {
        shear[0] = hist[0]; shear[1] = hist[1]; shear[2] = hist[2];
    } else {
        shear[0] = shear[1] = shear[2] = T(0);
    }
    const T gammat_eff = meff * gammat;
    if (shearupdate != 0)
        for (int k = 0; k < 3; ++k) shear[k] = shear[k] + vtr[k] * dt;
    const T shrmag = tsqrt(dot3(shear, shear));
    if (shearupdate != 0) {
        const T rsht = dot3(shear, dpos) * rsqinv;
        for (int k = 0; k < 3; ++k) shear[k] = shear[k] - rsht * dpos[k];
    }
    T fs[3];
    if (HERTZ) {
        for (int k = 0; k < 3; ++k) fs[k] = -polyhertz * (kt * shear[k] + gammat_eff * vtr[k]);
    } else {
        for (int k = 0; k < 3; ++k) fs[k] = -(kt * shear[k] + gammat_eff * vtr[k]);
    }
    const T fs_mag = tsqrt(dot3(fs, fs));
    const T fn_coulomb = xmu * fabs(ccel * r);
    if (fs_mag > fn_coulomb) {
        if (shrmag != T(0)) {
            const T ratio = tdiv(fn_coulomb, fs_mag);
            for (int k = 0; k < 3; ++k) {
                const T damp_corr = tdiv(gammat_eff * vtr[k], kt);
                shear[k] = ratio * (shear[k] + damp_corr) - damp_corr;
                fs[k] = ratio * fs[k];
            }
        } else {
            fs[0] = fs[1] = fs[2] = T(0);
        }
    }
    hist[0] = shear[0]; hist[1] = shear[1]; hist[2] = shear[2];
    for (int k = 0; k < 3; ++k) o.F[k] = dpos[k] * ccel + fs[k];
    T c[3];
    cross3(dpos, fs, c);
    for (int k = 0; k < 3; ++k) {
        const T tor = rinv * c[k];
        o.ta[k] = -ri * tor;
        o.tb[k] = -rj * tor;
    }
    o.vir[0] = T(0.5) * (dpos[0] * o.F[0]);
    o.vir[1] = T(0.5) * (dpos[1] * o.F[1]);
    o.vir[2] = T(0.5) * (dpos[2] * o.F[2]);
    o.vir[3] = T(0.5) * (dpos[0] * o.F[1]);
    o.vir[4] = T(0.5) * (dpos[0] * o.F[2]);
    o.vir[5] = T(0.5) * (dpos[1] * o.F[2]);
}

// PairGranular._rotate_rescale, GRANULAR/granular.py:271-280
template <typename T>
__device__ __forceinline__ void rotate_rescale(T h[3], const T n[3]) {
    const T shrmag = tsqrt(dot3(h, h));
    const T hn = dot3(h, n);
    T out[3];
    for (int k = 0; k < 3; ++k) out[k] = h[k] - hn * n[k];
    const T prjmag = tsqrt(dot3(out, out));
    T scale = T(0);
    if (prjmag > T(0)) scale = tdiv(shrmag, prjmag);
    for (int k = 0; k < 3; ++k) h[k] = out[k] * scale;
}

// PairGranular.compute_kernel contact body, GRANULAR/granular.py:316-426.
template <typename T>
__device__ __forceinline__ void granular_contact(
    const T dpos[3], const T dvj[3], T rsq, T radsum, T ri, T rj, T mi, T mj, const T vi[3],
    const T vj[3], const T oi[3], const T oj[3], const PairP& P, T dt, int history_update,
    bool touching_before, T hist[3], ContactOut<T>& o) {
    const T r = tsqrt(rsq);
    const T rinv = tdiv(T(1), r);
    T nrm[3] = {dpos[0] * rinv, dpos[1] * rinv, dpos[2] * rinv};
    const T delta = radsum - r;
    const T reff = tdiv(ri * rj, radsum);
    const T meff = tdiv(mi * mj, mi + mj);
    const T a = tsqrt(delta * reff);
    T vr[3];
    for (int k = 0; k < 3; ++k) vr[k] = vi[k] - (vj[k] + dvj[k]);
    const T vnnr = dot3(vr, nrm);
    T vt[3], wr[3], wxn[3], vtr[3];
    for (int k = 0; k < 3; ++k) {
        vt[k] = vr[k] - vnnr * nrm[k];
        wr[k] = ri * oi[k] + rj * oj[k];
    }
    cross3(wr, nrm, wxn);
    for (int k = 0; k < 3; ++k) vtr[k] = vt[k] - wxn[k];
    const T vrel = tsqrt(dot3(vtr, vtr));

    const T gkn = T(P.g_kn), gdamp = T(P.g_damp);
    T fne;
    if (P.normal_model != 0) fne = gkn * a * delta;
    else fne = gkn * delta;

    T damp_prefactor = T(0);
    switch (P.damping_model) {
        case 1: damp_prefactor = gdamp; break;
        case 2: damp_prefactor = gdamp * meff; break;
        case 3: damp_prefactor = gdamp * meff * a; break;
        case 4:
        case 5: {
            T arg = T(0);
            if (delta > T(0)) arg = fmax(T(0), tdiv(meff * fne, delta));
            damp_prefactor = gdamp * tsqrt(arg);
        } break;
        default: break;
    }
    T fntot = fne - damp_prefactor * vnnr;
    if (P.limit_damping == 1) fntot = fmax(fntot, T(0));
    const T fscrit = T(P.g_mu) * fabs(fntot);

    T fs[3] = {T(0), T(0), T(0)};
    if (P.tangential_model == 1) {
        const T damp_t = T(P.g_xt) * damp_prefactor;
        T ft = T(0);
        if (vrel != T(0)) ft = tdiv(fmin(fscrit, damp_t * vrel), vrel);
        for (int k = 0; k < 3; ++k) fs[k] = -ft * vtr[k];
    } else if (P.uses_history) {
        T h[3];
        if (touching_before) {
            h[0] = hist[0]; h[1] = hist[1]; h[2] = hist[2];
        } else {
            h[0] = h[1] = h[2] = T(0);
        }
        T k_eff = T(P.g_kt);
        if (P.tangential_model == 3) k_eff = T(P.g_kt) * a;
        const T damp_t = T(P.g_xt) * damp_prefactor;
        if (history_update != 0) {
            const T rsht = dot3(h, nrm);
            if (fabs(rsht) * k_eff > T(1e-10) * fscrit) rotate_rescale(h, nrm);
            for (int k = 0; k < 3; ++k) h[k] = h[k] + vtr[k] * dt;
        }
        T fdamp[3];
        for (int k = 0; k < 3; ++k) {
            fdamp[k] = -damp_t * vtr[k];
            fs[k] = -k_eff * h[k] + fdamp[k];
        }
        const T magfs = tsqrt(dot3(fs, fs));
        if (magfs > fscrit) {
            if (tsqrt(dot3(h, h)) != T(0)) {
                for (int k = 0; k < 3; ++k) fs[k] = fs[k] * tdiv(fscrit, magfs);
                for (int k = 0; k < 3; ++k) h[k] = tdiv(-(fs[k] - fdamp[k]), k_eff);
            } else {
                fs[0] = fs[1] = fs[2] = T(0);
            }
        }
        hist[0] = h[0]; hist[1] = h[1]; hist[2] = h[2];
    }
    for (int k = 0; k < 3; ++k) o.F[k] = fntot * nrm[k] + fs[k];
    T c[3];
    cross3(nrm, fs, c);
    for (int k = 0; k < 3; ++k) {
        const T tor = -c[k];
        o.ta[k] = (ri - T(0.5) * delta) * tor;
        o.tb[k] = (rj - T(0.5) * delta) * tor;
    }
    o.vir[0] = T(0.5) * (dpos[0] * o.F[0]);
    o.vir[1] = T(0.5) * (dpos[1] * o.F[1]);
    o.vir[2] = T(0.5) * (dpos[2] * o.F[2]);
    o.vir[3] = T(0.5) * (dpos[0] * o.F[1]);
    o.vir[4] = T(0.5) * (dpos[0] * o.F[2]);
    o.vir[5] = T(0.5) * (dpos[1] * o.F[2]);
}

// ------------------------------------------------------ reductions (FP64)

// Pairwise sum of one leaf per lane over a 64-lane wavefront: lanes combine in
// adjacent pairs, then pairs of pairs, ... (xor offsets 1, 2, 4, ..., 32), the
// same binary tree a serial pairwise sum over the 64 leaves forms.
// One DPP move of a double across lanes (CTRL: a gfx9 dpp_ctrl code).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Pairwise tree over the 64 lanes of a wavefront in lane order (all lanes
// active): adjacent pairs, then pairs of pairs, ... -- the same additions as a
// pairwise tree over those 64 leaves. Lane 8k+4's value reaches lane 8k by
// row_ror:12, lane 16k+8's lane 16k by row_ror:8; the four 16-lane sums are
// combined as uniform values. Every lane receives the sum.
__device__ __forceinline__ double wave_tree_sum(double v) {
    v = v + dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]: lane i^1
    v = v + dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]: lane i^2
    v = v + dpp_f64<0x12C>(v);  // row_ror:12: lane i+4 within the row
    v = v + dpp_f64<0x128>(v);  // row_ror:8: lane i+8 within the row
    const double a = readlane_f64(v, 0), b = readlane_f64(v, 16);
    const double c = readlane_f64(v, 32), d = readlane_f64(v, 48);
    return (a + b) + (c + d);
}

// Block of TMPS_BLOCK (= 4 wavefronts) leaves -> one pairwise-tree sum per
// component, written to out[0..NC) by thread 0.
template <int NC>
__device__ __forceinline__ void block_tree_sum(double leaf[NC], double* out) {
    __shared__ double wsum[TMPS_BLOCK / 64][NC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        double t = wave_tree_sum(leaf[k]);
        if (lane == 0) wsum[wave][k] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NC; ++k)
            out[k] = (wsum[0][k] + wsum[1][k]) + (wsum[2][k] + wsum[3][k]);
    }
    __syncthreads();
}

// Lane k < 8 of the calling wavefront receives component k of the pairwise
// tree over its 64 lanes of x (8 components a lane), formed in registers:
// three reduce-scatter levels (lanes at distance 1, 2, 4 each add the half
// of their components they keep, so every lane ends with one component
// summed over its aligned group of 8 lanes), then butterflies at distances
// 8, 16, 32. Every addition joins two aligned sibling subtrees and
// floating-point addition is commutative, so each node of the pairwise tree
// gets the same bits as in a tree formed level by level. All 64 lanes must
// call it.
__device__ __forceinline__ double wave_sum8(const double x[8]) {
    const int lane = threadIdx.x & 63;
    const bool u1 = (lane & 1) != 0, u2 = (lane & 2) != 0, u4 = (lane & 4) != 0;
    double a4[4], a2[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double send = u1 ? x[j] : x[4 + j];
        const double keep = u1 ? x[4 + j] : x[j];
        a4[j] = keep + __shfl_xor(send, 1);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double send = u2 ? a4[j] : a4[2 + j];
        const double keep = u2 ? a4[2 + j] : a4[j];
        a2[j] = keep + __shfl_xor(send, 2);
    }
    double a1 = (u4 ? a2[1] : a2[0]) + __shfl_xor(u4 ? a2[0] : a2[1], 4);
    a1 = a1 + __shfl_xor(a1, 8);
    a1 = a1 + __shfl_xor(a1, 16);
    a1 = a1 + __shfl_xor(a1, 32);
    // lane l holds component 4(l & 1) + (l & 2) + ((l >> 2) & 1) of its group
    const int k = lane & 7;
    return __shfl(a1, ((k >> 2) & 1) | (k & 2) | ((k & 1) << 2));
}

// The same tree over a block's 256 threads: thread k < 8 receives component
// k, the wavefronts' subtrees joined as (w0 + w1) + (w2 + w3).
__device__ __forceinline__ double block_sum8(const double leaf[8]) {
    __shared__ double ws[TMPS_BLOCK / 64][8];
    const double a = wave_sum8(leaf);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < 8) ws[wave][lane] = a;
    __syncthreads();
    const int k = threadIdx.x & 7;
    return (ws[0][k] + ws[1][k]) + (ws[2][k] + ws[3][k]);
}

}  // namespace tmps
