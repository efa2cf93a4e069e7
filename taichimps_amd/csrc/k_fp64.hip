// FP64 kernels of the DEM step: contact gather, integrator, damping,
// reductions, the deform/pressure servo and the fused per-step kernels.
//
// Built with -ffp-contract=off: every expression rounds as written, so each
// result is bitwise equal to the CPU oracle's for the same inputs.
#include <algorithm>

#include <hip/hip_ext.h>

#include "contact.h"
#include "fused.h"
#include "pbc.h"
#include "servo.h"
#include "tmps_internal.h"
#include "wall.h"

using namespace tmps;

namespace {

struct F64Ptrs {
    double4* xr;  // x, y, z, radius
    double4* vm;  // vx, vy, vz, rmass (read by every contact of the step)
    double4* om;  // omega, 0
    double4* vm_out;  // where a fused force pass writes v / omega
    double4* om_out;
    double4* f;
    double4* tq;
    double4* x0;
    double* vir;  // [n][6]
    const int* nb_start;
    const int* nb_j;
    uint8_t* nb_flag;
    double4* nb_hist;
    int n;
    WallRefs W;  // the fused run's walls (nwall 0 elsewhere)
};

F64Ptrs ptrs(System& s, int in_buf = 0, int out_buf = 0) {
    F64Ptrs p;
    p.xr = s.xr; p.f = s.f; p.tq = s.tq; p.x0 = s.x0; p.vir = s.vir;
    p.vm = in_buf ? s.vm2 : s.vm;
    p.om = in_buf ? s.om2 : s.om;
    p.vm_out = out_buf ? s.vm2 : s.vm;
    p.om_out = out_buf ? s.om2 : s.om;
    p.nb_start = s.nb_start; p.nb_j = s.nb_j; p.nb_flag = s.nb_flag;
    p.nb_hist = static_cast<double4*>(s.nb_hist);
    p.n = s.n;
    p.W = WallRefs{};
    return p;
}

// A fused step's walls before (pass 0) or after (pass 1) damping/cundall on
// atom i at its drifted position, with its half-step v and omega
// (FixWallGran.post_force between the pair force and final_integrate,
// simulation.py:236-247).
__device__ __forceinline__ void f64_walls(const WallRefs& W, int pass, const double4& x,
                                          const double4& v, const double4& w, int i,
                                          double F[3], double T[3]) {
    const double xs[3] = {x.x, x.y, x.z};
    const double vd[3] = {v.x, v.y, v.z}, wd[3] = {w.x, w.y, w.z};
    walls_apply<double>(W, pass, xs, x.w, v.w, vd, wd, W.spec[0].dt, i, F, T);
}

// What the contact gather needs of the box.
struct ImgP {
    double prd[3], inv[3];
    int per[3], fast[3];
    int vremap;
    const double* h_rate;
};

__device__ __forceinline__ ImgP img_params(const Ctl* ctl) {
    ImgP I;
    for (int k = 0; k < 3; ++k) {
        I.prd[k] = ctl->box.prd[k];
        I.inv[k] = 1.0 / I.prd[k];
        I.per[k] = ctl->box.per[k];
        I.fast[k] = ctl->fast_img[k];
    }
    I.vremap = ctl->box.vremap;
    I.h_rate = ctl->box.h_rate;
    return I;
}

// Domain.minimum_image_and_vshift, orthogonal branch (domain.py:388-462).
// Where the dimension holds at least 3 list bins the image count is taken by
// reciprocal multiply: it can differ from round(d / prd) only for |d| within
// rounding of half a box, which is farther than any contact reaches, so every
// contact sees the reference's separation bit for bit.
__device__ __forceinline__ void min_image_vshift(const ImgP& B, double d[3], double dv[3]) {
    double nimg[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (B.per[k] == 1) nimg[k] = round(B.fast[k] ? d[k] * B.inv[k] : d[k] / B.prd[k]);
    dv[0] = dv[1] = dv[2] = 0.0;
    if (B.vremap != 0) {
        const double* h = B.h_rate;
        dv[0] = nimg[0] * h[0] + nimg[1] * h[5] + nimg[2] * h[4];
        dv[1] = nimg[1] * h[1] + nimg[2] * h[3];
        dv[2] = nimg[2] * h[2];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] -= nimg[k] * B.prd[k];
}

// Sum of every contact of atom i over its full-list row, partners in
// ascending index order; each pair is evaluated with the lower index first.
template <int STYLE>
__device__ __forceinline__ void gather_f64(int i, const double4& xi, const double4& vi,
                                           const double4& oi, const F64Ptrs& P, const ImgP& B,
                                           const PairP& pp, double dt, int shearupdate,
                                           double accF[3], double accT[3], double accV[6],
                                           int& ntouch) {
    const int s0 = P.nb_start[i], s1 = P.nb_start[i + 1];
    const bool uses_history = (STYLE != TMPS_PAIR_GRANULAR) || pp.uses_history;
    for (int s = s0; s < s1; ++s) {
        const int j = P.nb_j[s];
        const double4 xj = P.xr[j];
        const bool first = i < j;
        const double4 xa = first ? xi : xj;
        const double4 xb = first ? xj : xi;
        double d[3] = {xa.x - xb.x, xa.y - xb.y, xa.z - xb.z};
        double dvj[3];
        min_image_vshift(B, d, dvj);
        const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double ra = xa.w, rb = xb.w;
        const double radsum = ra + rb;
        if (rsq < radsum * radsum && rsq > 1e-28) {
            const double4 vj = P.vm[j];
            const double4 oj = P.om[j];
            const double4 va4 = first ? vi : vj;
            const double4 vb4 = first ? vj : vi;
            const double4 oa4 = first ? oi : oj;
            const double4 ob4 = first ? oj : oi;
            const double va[3] = {va4.x, va4.y, va4.z}, vb[3] = {vb4.x, vb4.y, vb4.z};
            const double oa[3] = {oa4.x, oa4.y, oa4.z}, ob[3] = {ob4.x, ob4.y, ob4.z};
            bool before = false;
            double h[3] = {0.0, 0.0, 0.0};
            if (uses_history) {
                before = P.nb_flag[s] != 0;
                if (before) {
                    const double4 hh = P.nb_hist[s];
                    h[0] = hh.x; h[1] = hh.y; h[2] = hh.z;
                }
            }
            ContactOut<double> o;
            if (STYLE == TMPS_PAIR_HOOKE_HISTORY)
                legacy_history_contact<double, false>(d, dvj, rsq, radsum, ra, rb, va4.w, vb4.w, va,
                                                      vb, oa, ob, pp.kn, pp.kt, pp.gamman,
                                                      pp.gammat, pp.xmu, pp.limit_damping, dt,
                                                      shearupdate, before, h, o);
            else if (STYLE == TMPS_PAIR_HERTZ_HISTORY)
                legacy_history_contact<double, true>(d, dvj, rsq, radsum, ra, rb, va4.w, vb4.w, va,
                                                     vb, oa, ob, pp.kn, pp.kt, pp.gamman, pp.gammat,
                                                     pp.xmu, pp.limit_damping, dt, shearupdate,
                                                     before, h, o);
            else
                granular_contact<double>(d, dvj, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa, ob,
                                         pp, dt, shearupdate, before, h, o);
            if (uses_history) {
                if (!before) P.nb_flag[s] = 1;
                P.nb_hist[s] = make_double4(h[0], h[1], h[2], 0.0);
            }
            if (first) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    accF[k] = accF[k] + o.F[k];
                    accT[k] = accT[k] + o.ta[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    accF[k] = accF[k] + (-o.F[k]);
                    accT[k] = accT[k] + o.tb[k];
                }
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) accV[k] = accV[k] + o.vir[k];
            ++ntouch;
        } else if (uses_history) {
            if (P.nb_flag[s] != 0) P.nb_flag[s] = 0;
        }
    }
}

// ----------------------------------------------------------- stage kernels

__global__ void k_clear_f64(F64Ptrs P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    P.f[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    P.tq[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = 0.0;
}

// FixNVESphere.initial_integrate_kernel, nve_sphere.py:43-76
__global__ void k_nve_initial_f64(F64Ptrs P, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double dtv = dt, dtf = 0.5 * dt;
    double4 x = P.xr[i], v = P.vm[i], w = P.om[i];
    const double4 f = P.f[i], t = P.tq[i];
    const double m = v.w, r = x.w;
    const double inertia = 0.4 * m * r * r;
    v.x += dtf * (f.x / m);
    v.y += dtf * (f.y / m);
    v.z += dtf * (f.z / m);
    x.x = x.x + dtv * v.x;
    x.y = x.y + dtv * v.y;
    x.z = x.z + dtv * v.z;
    w.x += dtf * (t.x / inertia);
    w.y += dtf * (t.y / inertia);
    w.z += dtf * (t.z / inertia);
    P.xr[i] = x;
    P.vm[i] = v;
    P.om[i] = w;
}

// FixNVESphere.final_integrate_kernel, nve_sphere.py:78-96
__global__ void k_nve_final_f64(F64Ptrs P, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double dtf = 0.5 * dt;
    double4 v = P.vm[i], w = P.om[i];
    const double4 f = P.f[i], t = P.tq[i];
    const double m = v.w, r = P.xr[i].w;
    const double inertia = 0.4 * m * r * r;
    v.x += dtf * (f.x / m);
    v.y += dtf * (f.y / m);
    v.z += dtf * (f.z / m);
    w.x += dtf * (t.x / inertia);
    w.y += dtf * (t.y / inertia);
    w.z += dtf * (t.z / inertia);
    P.vm[i] = v;
    P.om[i] = w;
}

// FixDampingCundall.post_force_kernel, damping_cundall.py:35-53
__device__ __forceinline__ double cundall_scale(double fv, double g) {
    const double sgn = fv >= 0.0 ? 1.0 : -1.0;
    return 1.0 - g * sgn;
}

__global__ void k_cundall_f64(F64Ptrs P, double gl, double ga) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    double4 f = P.f[i], t = P.tq[i];
    const double4 v = P.vm[i], w = P.om[i];
    f.x *= cundall_scale(f.x * v.x, gl);
    t.x *= cundall_scale(t.x * w.x, ga);
    f.y *= cundall_scale(f.y * v.y, gl);
    t.y *= cundall_scale(t.y * w.y, ga);
    f.z *= cundall_scale(f.z * v.z, gl);
    t.z *= cundall_scale(t.z * w.z, ga);
    P.f[i] = f;
    P.tq[i] = t;
}

// GranularPair.compute: every contact of atom i, written (or added) to its
// force, torque and per-atom virial.
template <int STYLE>
__global__ void k_pair_f64(F64Ptrs P, const Ctl* ctl, PairP pp, double dt, int shearupdate,
                           int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const ImgP B = img_params(ctl);
    double accF[3] = {0.0, 0.0, 0.0}, accT[3] = {0.0, 0.0, 0.0};
    double accV[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (accumulate) {
        const double4 f = P.f[i], t = P.tq[i];
        accF[0] = f.x; accF[1] = f.y; accF[2] = f.z;
        accT[0] = t.x; accT[1] = t.y; accT[2] = t.z;
        for (int k = 0; k < 6; ++k) accV[k] = P.vir[6 * i + k];
    }
    int nt = 0;
    gather_f64<STYLE>(i, P.xr[i], P.vm[i], P.om[i], P, B, pp, dt, shearupdate, accF, accT, accV,
                      nt);
    P.f[i] = make_double4(accF[0], accF[1], accF[2], 0.0);
    P.tq[i] = make_double4(accT[0], accT[1], accT[2], 0.0);
    for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = accV[k];
}

// NeighborList.check_displacement, neighbor.py:214-245 (a flag, not a stamp:
// the host clears it before the launch).
__global__ void k_check_f64(F64Ptrs P, Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double4 x = P.xr[i], x0 = P.x0[i];
    const double d0 = x.x - x0.x, d1 = x.y - x0.y, d2 = x.z - x0.z;
    if (d0 * d0 + d1 * d1 + d2 * d2 > ctl->trigger_sq) ctl->moved = 1;
}

// Per-atom leaves of compute_virial_kernel (thermo.py:62-109), one pairwise
// tree per block of atoms.
__global__ void k_leaves_f64(F64Ptrs P, int keflag, double* partial) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < P.n) {
        for (int k = 0; k < 6; ++k) leaf[k] = P.vir[6 * i + k];
        if (keflag) {
            const double4 v = P.vm[i];
            const double m = v.w;
            leaf[0] = leaf[0] + m * v.x * v.x;
            leaf[1] = leaf[1] + m * v.y * v.y;
            leaf[2] = leaf[2] + m * v.z * v.z;
            leaf[3] = leaf[3] + m * v.x * v.y;
            leaf[4] = leaf[4] + m * v.x * v.z;
            leaf[5] = leaf[5] + m * v.y * v.z;
        }
    }
    block_tree_sum<7>(leaf, partial + (size_t)blockIdx.x * 8);
}

// compute_ke_kernel leaves, thermo.py:46-60
__global__ void k_ke_leaves_f64(F64Ptrs P, double* partial) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[2] = {0.0, 0.0};
    if (i < P.n) {
        const double4 v = P.vm[i], w = P.om[i];
        const double r = P.xr[i].w, m = v.w;
        const double vsq = v.x * v.x + v.y * v.y + v.z * v.z;
        leaf[0] = 0.5 * m * vsq;
        const double im = 0.4 * m * r * r;
        const double wsq = w.x * w.x + w.y * w.y + w.z * w.z;
        leaf[1] = 0.5 * im * wsq;
    }
    block_tree_sum<2>(leaf, partial + (size_t)blockIdx.x * 8);
}

// ---------------------------------------------------- final tree + servo

#define FINAL_THREADS 1024

// Pairwise tree over nparts block partials (zero-padded to a power of two),
// by one block of FINAL_THREADS threads: thread t sums leaves [t*L, t*L+L) as
// an adjacent-pair tree in registers, then the threads' sums are combined the
// same way (wavefront butterfly, then the 16 wavefront sums). Thread 0
// receives the sums.
template <int L, int NC, int THREADS = FINAL_THREADS>
__device__ void final_tree_L(const double* partial, int nparts, double out[NC]) {
    __shared__ double ws[THREADS / 64][NC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = threadIdx.x * L;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        double t[L];
#pragma unroll
        for (int q = 0; q < L; ++q)
            t[q] = (base + q < nparts) ? partial[(size_t)(base + q) * 8 + k] : 0.0;
#pragma unroll
        for (int w = 1; w < L; w <<= 1)
#pragma unroll
            for (int q = 0; q < L; q += 2 * w) t[q] = t[q] + t[q + w];
        const double v = wave_tree_sum(t[0]);
        if (lane == 0) ws[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            double u[THREADS / 64];
#pragma unroll
            for (int q = 0; q < THREADS / 64; ++q) u[q] = ws[q][k];
#pragma unroll
            for (int w = 1; w < THREADS / 64; w <<= 1)
#pragma unroll
                for (int q = 0; q < THREADS / 64; q += 2 * w) u[q] = u[q] + u[q + w];
            out[k] = u[0];
        }
    }
    __syncthreads();
}

// Leaves per thread for nparts partials over `threads` threads (a power of
// two, at least 1). Leaf chunks, wavefronts and the block are all aligned
// power-of-two subtrees, so the thread count does not change the tree.
inline int final_leaves(int nparts, int threads = FINAL_THREADS) {
    int p2 = 1;
    while (p2 < nparts) p2 <<= 1;
    return p2 <= threads ? 1 : p2 / threads;
}

// The step close kernel's block (two wavefronts per SIMD) and how many rounds
// of one block partial per thread it loads at once.
#define CLOSE_THREADS 512
#define CLOSE_BATCH 8

// Partials one close block reduces: CLOSE_BATCH rounds of CLOSE_THREADS.
#define CLOSE_SPAN (CLOSE_THREADS * CLOSE_BATCH)


// The tree of final_tree_L over 8 components for up to CLOSE_SPAN partials.
// Thread t owns component t & 7 of the chunk t >> 3 of 8L consecutive
// partials, which it loads (eight threads read one partial's 64-byte row) and
// sums pairwise in registers; chunks are then joined across lanes 8, 16 and
// 32 apart and the wavefronts' sums pairwise through LDS. Chunks, wavefront
// ranges and the block's range are aligned power-of-two subtrees of the
// 512L-leaf range (zero leaves past nparts), so this is final_tree_L's tree.
// Every thread receives the sums.
template <int L>
__device__ __forceinline__ void close_tree_load(const double* partial, int nparts,
                                                double (&v)[8 * L]) {
    static_assert(L <= CLOSE_BATCH, "a close block reduces at most CLOSE_SPAN partials");
    constexpr int S = 8 * L;
    const int t = threadIdx.x, k = t & 7, c = t >> 3;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int p = c * S + q;
        v[q] = p < nparts ? partial[(size_t)p * 8 + k] : 0.0;
    }
}

template <int L>
__device__ __forceinline__ void close_tree_sum(double (&v)[8 * L], double out[8]) {
    constexpr int S = 8 * L;
    constexpr int NWAVE = CLOSE_THREADS / 64;
    __shared__ double ws[NWAVE][8];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int w = 1; w < S; w <<= 1)
#pragma unroll
        for (int q = 0; q < S; q += 2 * w) v[q] = v[q] + v[q + w];
    double a = v[0];
    a = a + __shfl_xor(a, 8);
    a = a + __shfl_xor(a, 16);
    a = a + __shfl_xor(a, 32);
    if (lane < 8) ws[wave][lane] = a;
    __syncthreads();
    if (t < 8) {
        double u[NWAVE];
#pragma unroll
        for (int q = 0; q < NWAVE; ++q) u[q] = ws[q][t];
#pragma unroll
        for (int w = 1; w < NWAVE; w <<= 1)
#pragma unroll
            for (int q = 0; q < NWAVE; q += 2 * w) u[q] = u[q] + u[q + w];
        ws[0][t] = u[0];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) out[q] = ws[0][q];
    __syncthreads();
}

template <int L>
__device__ void close_tree(const double* partial, int nparts, double out[8]) {
    double v[8 * L];
    close_tree_load<L>(partial, nparts, v);
    close_tree_sum<L>(v, out);
}

// Above CLOSE_SPAN block partials, one block per aligned group of CLOSE_SPAN
// reduces the group to one 8-double entry of `level` first; groups are
// aligned subtrees, so the close kernel's tree over the entries is the same
// tree.
__global__ void __launch_bounds__(CLOSE_THREADS) k_tree_level(const double* partial, int nparts,
                                                               double* level) {
    const int g = blockIdx.x;
    const int count = min(CLOSE_SPAN, nparts - g * CLOSE_SPAN);
    double out[8];
    close_tree<CLOSE_BATCH>(partial + (size_t)g * CLOSE_SPAN * 8, count, out);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) level[(size_t)g * 8 + k] = out[k];
}

template <int L, int NC>
__global__ void __launch_bounds__(FINAL_THREADS) k_final_tree(Ctl* ctl, const double* partial,
                                                               int nparts) {
    double out[NC];
    final_tree_L<L, NC>(partial, nparts, out);
    if (threadIdx.x == 0)
        for (int k = 0; k < NC; ++k) ctl->vsum[k] = out[k];
}

// Tree-sums the force pass's block partials and, when enabled, runs the
// servo. `advance` marks a fused step: it honours the segment's abort and
// rebuild-trigger flags and completes the step's bookkeeping.
template <int L>
__device__ __forceinline__ void step_end_block(Ctl* ctl, const double* partial, int nparts,
                                               double dt, int advance, int apply_servo,
                                               int natoms) {
    if (advance) {
        if (ctl->abort) return;
        const int t = ctl->timestep + 1;
        if (ctl->trigger == t) {
            if (threadIdx.x == 0) ctl->abort = 1;
            return;
        }
    }
    double out[7];
    final_tree_L<L, 7>(partial, nparts, out);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; ++k) ctl->vsum[k] = out[k];
        ctl->ntouch = (int)out[6];
        if (apply_servo) {
            servo_update(ctl, out, dt, natoms);
            if (!ctl->abort) update_split_consts(ctl, 0);
        }
        if (advance && !ctl->abort) ctl->timestep = ctl->timestep + 1;
    }
}

// step_end_block on the close's tree (the same tree): one block of
// CLOSE_THREADS over at most CLOSE_SPAN partials (the launcher reduces
// aligned groups first above that).
template <int L>
__global__ void __launch_bounds__(CLOSE_THREADS) k_step_end(Ctl* ctl, const double* partial,
                                                             int nparts, double dt, int advance,
                                                             int apply_servo, int natoms) {
    if (advance) {
        if (ctl->abort) return;
        const int t = ctl->timestep + 1;
        if (ctl->trigger == t) {
            if (threadIdx.x == 0) ctl->abort = 1;
            return;
        }
    }
    double out[8];
    close_tree<L>(partial, nparts, out);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; ++k) ctl->vsum[k] = out[k];
        ctl->ntouch = (int)out[6];
        if (apply_servo) {
            servo_update(ctl, out, dt, natoms);
            if (!ctl->abort) update_split_consts(ctl, 0);
        }
        if (advance && !ctl->abort) ctl->timestep = ctl->timestep + 1;
    }
}

// End of a fused step, in one block, for passes that publish their block
// partials (publish_block): the pairwise tree over them, then close_update.
// The control block is copied into LDS and written back whole, both as one
// parallel round trip, so the serial servo arithmetic in between waits on no
// global memory. An aborted segment is left untouched.
template <int L>
__device__ __forceinline__ void step_close_block(Ctl* ctl, const double* partial, int nparts,
                                                 double dt, int servo_on, int natoms,
                                                 bool member) {
    static_assert(sizeof(Ctl) % 4 == 0, "the control block is copied in 4-byte words");
    constexpr int NW = (int)(sizeof(Ctl) / 4);
    __shared__ Ctl cs;
    __shared__ int writeback;
    // a segment cut short by a rebuild request skips its remaining closes
    static_assert(NW <= CLOSE_THREADS, "one control-block word per thread");
    if (ctl->abort) return;
    const unsigned* src = reinterpret_cast<const unsigned*>(ctl);
    unsigned* lds = reinterpret_cast<unsigned*>(&cs);
    // the control block's words and the partials are requested together
    const unsigned word = threadIdx.x < NW ? src[threadIdx.x] : 0u;
    double v[8 * L];
    close_tree_load<L>(partial, nparts, v);
    if (threadIdx.x < NW) lds[threadIdx.x] = word;
    double out[8];
    close_tree_sum<L>(v, out);  // its barriers complete cs
    if (threadIdx.x == 0) writeback = close_update(cs, out, dt, servo_on, natoms, member) ? 1 : 0;
    __syncthreads();
    if (writeback) {
        unsigned* dst = reinterpret_cast<unsigned*>(ctl);
        for (int k = threadIdx.x; k < NW; k += CLOSE_THREADS) dst[k] = lds[k];
    }
}

template <int L>
__global__ void __launch_bounds__(CLOSE_THREADS) k_step_close(Ctl* ctl, const double* partial,
                                                               int nparts, double dt,
                                                               int servo_on, int natoms) {
    step_close_block<L>(ctl, partial, nparts, dt, servo_on, natoms, false);
}

// The close of every running ensemble member, one block each.
template <int L>
__global__ void __launch_bounds__(CLOSE_THREADS) k_batch_close(const CloseEntry* table, double dt) {
    const CloseEntry& e = table[blockIdx.x];
    if (e.ctl->bdone) return;
    step_close_block<L>(e.ctl, e.partial, e.nparts, dt, e.servo_on, e.natoms, true);
}

// remap_positions, deform_pressure.py:493-530 (zero tilt)
__device__ __forceinline__ void remap_point(const RemapRec& R, double4& x) {
    const double d0 = x.x - R.old_lo[0], d1 = x.y - R.old_lo[1], d2 = x.z - R.old_lo[2];
    const double l0 = d0 / R.old_prd[0] + R.oh[0] * d1 + R.oh[1] * d2;
    const double l1 = d1 / R.old_prd[1] + R.oh[2] * d2;
    const double l2 = d2 / R.old_prd[2];
    x.x = R.new_prd[0] * l0 + 0.0 * l1 + 0.0 * l2 + R.new_lo[0];
    x.y = R.new_prd[1] * l1 + 0.0 * l2 + R.new_lo[1];
    x.z = R.new_prd[2] * l2 + R.new_lo[2];
}

// Applies a pending remap: in stage mode whatever is pending, in fused mode
// only the one recorded by the step just completed.
__global__ void k_remap_apply(F64Ptrs P, const Ctl* ctl, int fused) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const RemapRec& R = ctl->remap;
    if (R.step == -1) return;
    if (fused && (ctl->abort || R.step != ctl->timestep)) return;
    double4 x = P.xr[i];
    remap_point(R, x);
    P.xr[i] = x;
}

__global__ void k_remap_clear(Ctl* ctl, int honor_abort) {
    if (honor_abort && ctl->abort) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->remap.step = -1;
}

__device__ __forceinline__ void ctl_rebuilt(Ctl* ctl, int step, int split) {
    ctl->trigger = -1;
    ctl->abort = 0;
    ctl->last_build_step = step;
    if (split) update_split_consts(ctl, 0);
}

// (a build that recorded an error keeps its abort: the rest of the redo
// then leaves the step's state for the host). A negative step is the step
// being redone, read on the device: the aborted step left the timestep one
// short of it, and nothing in the redo advances it before the step's end.
__global__ void k_ctl_rebuilt(Ctl* ctl, int step, int split) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !ctl->error)
        ctl_rebuilt(ctl, step < 0 ? ctl->timestep + 1 : step, split);
}

// The host's edited mirror of the control block, stored into the device's
// copy in stream order. The value travels as the kernel's argument, so the
// host may edit its mirror again as soon as the launch returns.
static_assert(sizeof(Ctl) % 4 == 0 && sizeof(Ctl) <= 3072,
              "the control block travels as one kernel argument");
__global__ void k_ctl_store(Ctl* ctl, Ctl value) {
    constexpr int NW = (int)(sizeof(Ctl) / 4);
    const unsigned* src = reinterpret_cast<const unsigned*>(&value);
    unsigned* dst = reinterpret_cast<unsigned*>(ctl);
    for (int k = threadIdx.x; k < NW; k += blockDim.x) dst[k] = src[k];
}

// A step cut short by a rebuild request is about to be redone.
__global__ void k_ctl_resume(Ctl* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->trigger = -1;
        ctl->abort = 0;
    }
}

// What a list build records in the control block: its step, per axis whether
// it has at least 3 bins, and (FP32-split) the bins of the encoding.
// (FP32-split: with the encoding's bins, and the split constants of the
// build box, which becomes the skin check's reference)
__device__ __forceinline__ void ctl_built(Ctl* ctl, int step, int3 fast, int3 gi, int split) {
    ctl->last_build_step = step;
    ctl->fast_img[0] = fast.x;
    ctl->fast_img[1] = fast.y;
    ctl->fast_img[2] = fast.z;
    if (split) {
        ctl->sc.gi[0] = gi.x;
        ctl->sc.gi[1] = gi.y;
        ctl->sc.gi[2] = gi.z;
        update_split_consts(ctl, 1);
    }
}

__global__ void k_ctl_built(Ctl* ctl, int step, int3 fast, int3 gi, int split) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl_built(ctl, step, fast, gi, split);
}

// ctl_built for a list build decided on the device, from its grid; the
// build box is recorded in both precisions (the host derives the reported
// grid from it after the run).
// (a negative step: the step being redone, as for k_ctl_rebuilt)
__global__ void k_ctl_built_dev(Ctl* ctl, int step, int split) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || ctl->error) return;
    const GridP& G = ctl->lgrid;
    ctl_built(ctl, step < 0 ? ctl->timestep + 1 : step, make_int3(G.g[0] >= 3, G.g[1] >= 3, G.g[2] >= 3),
              make_int3(G.g[0], G.g[1], G.g[2]), split);
    if (!split)
        for (int k = 0; k < 3; ++k) {
            ctl->lo_build[k] = ctl->box.lo[k];
            ctl->prd_build[k] = ctl->box.prd[k];
        }
}

// ---------------------------------------------------- fused step kernels

// Start of step t: the pending end-of-step remap of step t-1, the
// initial_integrate half-kick (only for a segment's first step; later steps
// receive it from the previous force pass), the drift, and the skin check.
__device__ __forceinline__ void step_begin_f64_atom(int i, const F64Ptrs& P, Ctl* ctl, double dt,
                                                    int kick) {
    const int t = ctl->timestep + 1;
    double4 x = P.xr[i];
    const RemapRec& R = ctl->remap;
    if (R.step != -1 && R.step == ctl->timestep) remap_point(R, x);
    double4 v = P.vm[i];
    const double dtf = 0.5 * dt;
    if (kick) {
        const double4 f = P.f[i];
        const double m = v.w;
        v.x += dtf * (f.x / m);
        v.y += dtf * (f.y / m);
        v.z += dtf * (f.z / m);
    }
    x.x = x.x + dt * v.x;
    x.y = x.y + dt * v.y;
    x.z = x.z + dt * v.z;
    if (kick) {
        double4 w = P.om[i];
        const double4 tq = P.tq[i];
        const double m = v.w, r = x.w;
        const double inertia = 0.4 * m * r * r;
        w.x += dtf * (tq.x / inertia);
        w.y += dtf * (tq.y / inertia);
        w.z += dtf * (tq.z / inertia);
        P.om[i] = w;
        P.vm[i] = v;
    }
    P.xr[i] = x;
    const int since = t - ctl->last_build_step;
    if (since >= ctl->delay && since % ctl->every == 0) {
        if (!ctl->check) {
            if (i == 0) ctl->trigger = t;
        } else {
            const double4 x0 = P.x0[i];
            const double d0 = x.x - x0.x, d1 = x.y - x0.y, d2 = x.z - x0.z;
            if (d0 * d0 + d1 * d1 + d2 * d2 > ctl->trigger_sq) ctl->trigger = t;
        }
    }
}

__global__ void k_step_begin_f64(F64Ptrs P, Ctl* ctl, double dt, int kick) {
    if (ctl->abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.n) step_begin_f64_atom(i, P, ctl, dt, kick);
}

// The start of a redone step in one pass over the atoms: k_step_begin_f64
// (no kick) and the rebuild's wrap (k_pbc), each as its own kernel computes
// it.
__global__ void k_redo_begin_f64(F64Ptrs P, Ctl* ctl, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    if (!ctl->abort) step_begin_f64_atom(i, P, ctl, dt, 0);
    pbc_atom(i, P.xr, P.vm, ctl);
}

// Step t after the drift: contact forces, damping/cundall, final_integrate,
// the block's share of the virial(+kinetic) tree, and either the next step's
// half-kick (mid-segment) or the stored f/torque/virial (segment end).
template <int STYLE, int CUNDALL, int FINAL, int WALL>
__global__ void k_step_force_f64(F64Ptrs P, Ctl* ctl, PairP pp, double dt, double gl, double ga,
                                 double* partial) {
    if (ctl->abort) return;
    const int t = ctl->timestep + 1;
    if (ctl->trigger == t) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < P.n) {
        const ImgP B = img_params(ctl);
        const double4 xi = P.xr[i];
        double4 v = P.vm[i];
        double4 w = P.om[i];
        double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
        double V[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int nt = 0;
        gather_f64<STYLE>(i, xi, v, w, P, B, pp, dt, 1, F, T, V, nt);
        if (WALL) f64_walls(P.W, 0, xi, v, w, i, F, T);
        if (CUNDALL) {
            F[0] *= cundall_scale(F[0] * v.x, gl);
            T[0] *= cundall_scale(T[0] * w.x, ga);
            F[1] *= cundall_scale(F[1] * v.y, gl);
            T[1] *= cundall_scale(T[1] * w.y, ga);
            F[2] *= cundall_scale(F[2] * v.z, gl);
            T[2] *= cundall_scale(T[2] * w.z, ga);
        }
        if (WALL) f64_walls(P.W, 1, xi, v, w, i, F, T);
        const double dtf = 0.5 * dt;
        const double m = v.w, r = xi.w;
        const double inertia = 0.4 * m * r * r;
        const double am[3] = {F[0] / m, F[1] / m, F[2] / m};
        const double ai[3] = {T[0] / inertia, T[1] / inertia, T[2] / inertia};
        v.x += dtf * am[0];
        v.y += dtf * am[1];
        v.z += dtf * am[2];
        w.x += dtf * ai[0];
        w.y += dtf * ai[1];
        w.z += dtf * ai[2];
        leaf[0] = V[0] + m * v.x * v.x;
        leaf[1] = V[1] + m * v.y * v.y;
        leaf[2] = V[2] + m * v.z * v.z;
        leaf[3] = V[3] + m * v.x * v.y;
        leaf[4] = V[4] + m * v.x * v.z;
        leaf[5] = V[5] + m * v.y * v.z;
        leaf[6] = (double)nt;
        if (FINAL) {
            P.f[i] = make_double4(F[0], F[1], F[2], 0.0);
            P.tq[i] = make_double4(T[0], T[1], T[2], 0.0);
            if (FINAL == 1) {
                for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = V[k];
            } else {
                // FINAL 2, a rebuild step redone inside a run: the next
                // step's half-kick, k_kick_f64's expressions on the stored
                // force and torque; no virial (only a run's last step
                // leaves it for the getters)
                v.x += dtf * (F[0] / m);
                v.y += dtf * (F[1] / m);
                v.z += dtf * (F[2] / m);
                w.x += dtf * (T[0] / inertia);
                w.y += dtf * (T[1] / inertia);
                w.z += dtf * (T[2] / inertia);
            }
        } else {
            v.x += dtf * am[0];
            v.y += dtf * am[1];
            v.z += dtf * am[2];
            w.x += dtf * ai[0];
            w.y += dtf * ai[1];
            w.z += dtf * ai[2];
        }
        P.vm_out[i] = v;
        P.om_out[i] = w;
    }
    block_tree_sum<7>(leaf, partial + (size_t)blockIdx.x * 8);
}

__global__ void k_copy_vw(double4* vdst, double4* wdst, const double4* vsrc, const double4* wsrc,
                          int n, const Ctl* ctl, int honor_abort) {
    if (honor_abort && ctl->abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vdst[i] = vsrc[i];
    wdst[i] = wsrc[i];
}

// ------------------------------------------------ one kernel per step

#ifndef F64_BATCH
#define F64_BATCH 4
#endif

// Buffers of one fused step: it reads set `in` (step t-1's results) and
// writes set `out`.
struct Fused64 {
    const double4 *xr_in, *vm_in, *om_in;
    double4 *xr_out, *vm_out, *om_out;
    const uint8_t* fl_in;
    uint8_t* fl_out;
    const double4* h_in;
    double4* h_out;
    double4 *f, *tq;
    const double4* x0;
    double* vir;
    const int *nb_start, *nb_j;
    int n;
    WallRefs W;  // the run's walls (nwall 0 without)
};

// Where atom j is at step t: step t-1's pending remap, then the drift with
// its half-step velocity -- the arithmetic of k_step_begin_f64, so every
// thread that needs atom j computes the same bits.
__device__ __forceinline__ double4 step_position(const RemapRec& R, bool remap, double dt,
                                                 double4 x, const double4& v) {
    if (remap) remap_point(R, x);
    x.x = x.x + dt * v.x;
    x.y = x.y + dt * v.y;
    x.z = x.z + dt * v.z;
    return x;
}

// Step t in one pass: drift (own atom and, on the fly, each neighbour), skin
// check, contact gather, damping/cundall, final_integrate, the next step's
// half-kick (or, at a segment's end, the stored f/torque/virial), and the
// block's tree partial (publish_block); k_step_close ends the step. Nothing of
// set `in` is written.
template <int STYLE, int CUNDALL, int FINAL, int WALL>
__global__ void __launch_bounds__(TMPS_BLOCK) __attribute__((amdgpu_waves_per_eu(TMPS_STEP_WAVES_F64))) k_step_f64(Fused64 P, Ctl* ctl, PairP pp, double* partial,
                                                          double dt, double gl, double ga) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int i = lb * TMPS_BLOCK + threadIdx.x;
    // the atom's own streams and row bounds are requested before the control
    // block is read, so the two memory round trips overlap
    const int ia = i < P.n ? i : P.n - 1;
    double4 v = P.vm_in[ia];
    double4 w = P.om_in[ia];
    const double4 xr_i = P.xr_in[ia];
    const double4 x0 = P.x0[ia];
    const int s0 = P.nb_start[ia], s1 = P.nb_start[ia + 1];
    if (ctl->abort) return;
    const int t = ctl->timestep + 1;
    double leaf[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int trig = 0;
    if (i < P.n) {
        const RemapRec RM = ctl->remap;
        const bool remap = RM.step != -1 && RM.step == ctl->timestep;
        const ImgP B = img_params(ctl);
        const double4 xi = step_position(RM, remap, dt, xr_i, v);
        if (check_due(ctl, t)) {
            if (!ctl->check) {
                trig = 1;
            } else {
                const double d0 = xi.x - x0.x, d1 = xi.y - x0.y, d2 = xi.z - x0.z;
                if (d0 * d0 + d1 * d1 + d2 * d2 > ctl->trigger_sq) trig = 1;
            }
        }
        double F[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
        double V[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int nt = 0;
        const bool uses_history = (STYLE != TMPS_PAIR_GRANULAR) || pp.uses_history;
        // Pass 1, every slot: separation and contact test, neighbour states
        // requested a batch at a time (see k_step_split). Touching slots go
        // into `tmask` (first 64 of the row; the rest stay in the flags).
        uint64_t tmask = 0;
        for (int sb = s0; sb < s1; sb += F64_BATCH) {
            int jb[F64_BATCH];
#pragma unroll
            for (int q = 0; q < F64_BATCH; ++q) jb[q] = sb + q < s1 ? P.nb_j[sb + q] : i;
            double4 vq[F64_BATCH], xq[F64_BATCH];
#pragma unroll
            for (int q = 0; q < F64_BATCH; ++q) {
                vq[q] = P.vm_in[jb[q]];
                xq[q] = P.xr_in[jb[q]];
            }
#pragma unroll
            for (int q = 0; q < F64_BATCH; ++q) {
                if (sb + q < s1) {
                    const double4 xj = step_position(RM, remap, dt, xq[q], vq[q]);
                    // i - j: the exact negation of the canonical (lower index
                    // first) separation, so rsq and the test are the same bits
                    double d[3] = {xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
                    double dvj[3];
                    min_image_vshift(B, d, dvj);
                    const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    const double radsum = xi.w + xj.w;
                    const bool touch = rsq < radsum * radsum && rsq > 1e-28;
                    const int k = sb + q - s0;
                    if (touch && k < 64) tmask |= 1ull << k;
                    P.fl_out[sb + q] = touch ? 1 : 0;
                }
            }
        }
        // Pass 2, touching slots in ascending order: the contact law with the
        // neighbour fetched again (the same bits as in pass 1).
        const int nlong = s1 - s0 > 64 ? s1 - s0 - 64 : 0;
        for (int k = 0; tmask != 0 || k < nlong; ) {
            int s;
            if (tmask != 0) {
                s = s0 + __builtin_ctzll(tmask);
                tmask &= tmask - 1;
            } else {
                s = s0 + 64 + k++;
                if (P.fl_out[s] == 0) continue;
            }
            const int j = P.nb_j[s];
            const double4 vj = P.vm_in[j];
            const double4 xj = step_position(RM, remap, dt, P.xr_in[j], vj);
            const bool first = i < j;
            const double4 xa = first ? xi : xj;
            const double4 xb = first ? xj : xi;
            double d[3] = {xa.x - xb.x, xa.y - xb.y, xa.z - xb.z};
            double dvj[3];
            min_image_vshift(B, d, dvj);
            const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double ra = xa.w, rb = xb.w;
            const double radsum = ra + rb;
            const double4 oj = P.om_in[j];
            const double4 va4 = first ? v : vj;
            const double4 vb4 = first ? vj : v;
            const double4 oa4 = first ? w : oj;
            const double4 ob4 = first ? oj : w;
            const double va[3] = {va4.x, va4.y, va4.z}, vb[3] = {vb4.x, vb4.y, vb4.z};
            const double oa[3] = {oa4.x, oa4.y, oa4.z}, ob[3] = {ob4.x, ob4.y, ob4.z};
            bool before = false;
            double h[3] = {0.0, 0.0, 0.0};
            if (uses_history) {
                before = P.fl_in[s] != 0;
                if (before) {
                    const double4 hh = P.h_in[s];
                    h[0] = hh.x; h[1] = hh.y; h[2] = hh.z;
                }
            }
            ContactOut<double> o;
            if (STYLE == TMPS_PAIR_HOOKE_HISTORY)
                legacy_history_contact<double, false>(d, dvj, rsq, radsum, ra, rb, va4.w, vb4.w, va,
                                                      vb, oa, ob, pp.kn, pp.kt, pp.gamman,
                                                      pp.gammat, pp.xmu, pp.limit_damping, dt, 1,
                                                      before, h, o);
            else if (STYLE == TMPS_PAIR_HERTZ_HISTORY)
                legacy_history_contact<double, true>(d, dvj, rsq, radsum, ra, rb, va4.w, vb4.w, va,
                                                     vb, oa, ob, pp.kn, pp.kt, pp.gamman, pp.gammat,
                                                     pp.xmu, pp.limit_damping, dt, 1, before, h, o);
            else
                granular_contact<double>(d, dvj, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa, ob,
                                         pp, dt, 1, before, h, o);
            if (uses_history) P.h_out[s] = make_double4(h[0], h[1], h[2], 0.0);
            if (first) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    F[c] = F[c] + o.F[c];
                    T[c] = T[c] + o.ta[c];
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    F[c] = F[c] + (-o.F[c]);
                    T[c] = T[c] + o.tb[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) V[c] = V[c] + o.vir[c];
            ++nt;
        }
        if (WALL) f64_walls(P.W, 0, xi, v, w, i, F, T);
        if (CUNDALL) {
            F[0] *= cundall_scale(F[0] * v.x, gl);
            T[0] *= cundall_scale(T[0] * w.x, ga);
            F[1] *= cundall_scale(F[1] * v.y, gl);
            T[1] *= cundall_scale(T[1] * w.y, ga);
            F[2] *= cundall_scale(F[2] * v.z, gl);
            T[2] *= cundall_scale(T[2] * w.z, ga);
        }
        if (WALL) f64_walls(P.W, 1, xi, v, w, i, F, T);
        const double dtf = 0.5 * dt;
        const double m = v.w, r = xi.w;
        const double inertia = 0.4 * m * r * r;
        const double am[3] = {F[0] / m, F[1] / m, F[2] / m};
        const double ai[3] = {T[0] / inertia, T[1] / inertia, T[2] / inertia};
        v.x += dtf * am[0];
        v.y += dtf * am[1];
        v.z += dtf * am[2];
        w.x += dtf * ai[0];
        w.y += dtf * ai[1];
        w.z += dtf * ai[2];
        leaf[0] = V[0] + m * v.x * v.x;
        leaf[1] = V[1] + m * v.y * v.y;
        leaf[2] = V[2] + m * v.z * v.z;
        leaf[3] = V[3] + m * v.x * v.y;
        leaf[4] = V[4] + m * v.x * v.z;
        leaf[5] = V[5] + m * v.y * v.z;
        leaf[6] = (double)nt;
        if (FINAL) {
            P.f[i] = make_double4(F[0], F[1], F[2], 0.0);
            P.tq[i] = make_double4(T[0], T[1], T[2], 0.0);
            for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = V[k];
        } else {
            v.x += dtf * am[0];
            v.y += dtf * am[1];
            v.z += dtf * am[2];
            w.x += dtf * ai[0];
            w.y += dtf * ai[1];
            w.z += dtf * ai[2];
        }
        P.xr_out[i] = xi;
        P.vm_out[i] = v;
        P.om_out[i] = w;
    }
    publish_block(leaf, trig, lb, partial);
}

// initial_integrate's half-kick of a segment's first step (the later steps
// receive theirs from the previous step's pass), in place.
__global__ void k_kick_f64(double4* vm, double4* om, const double4* xr, const double4* f,
                           const double4* tq, int n, const Ctl* ctl, double dt) {
    if (ctl->abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dtf = 0.5 * dt;
    double4 v = vm[i];
    const double4 ff = f[i];
    const double m = v.w;
    v.x += dtf * (ff.x / m);
    v.y += dtf * (ff.y / m);
    v.z += dtf * (ff.z / m);
    double4 w = om[i];
    const double4 t = tq[i];
    const double r = xr[i].w;
    const double inertia = 0.4 * m * r * r;
    w.x += dtf * (t.x / inertia);
    w.y += dtf * (t.y / inertia);
    w.z += dtf * (t.z / inertia);
    vm[i] = v;
    om[i] = w;
}

// Buffer set 1 -> set 0: per-atom 16/32-byte records and per-slot flags and
// history, one grid-stride pass.
template <class T4>
__device__ __forceinline__ void copy_state_items(int first, int stride, T4* x0, T4* v0, T4* w0,
                                                 const T4* x1, const T4* v1, const T4* w1, int n,
                                                 uint8_t* fl0, const uint8_t* fl1, T4* h0,
                                                 const T4* h1, const int* nb_start) {
    // the slot count of the current list, read on the device: a captured
    // segment outlives the list it was captured with
    const int nslots = nb_start[n];
    for (int i = first; i < n; i += stride) {
        x0[i] = x1[i];
        v0[i] = v1[i];
        w0[i] = w1[i];
    }
    // a slot's history is read only where its touch flag is set (a step
    // writes it only there), so only those entries are copied
    for (int k = first; k < nslots; k += stride) {
        const uint8_t f = fl1[k];
        fl0[k] = f;
        if (f) h0[k] = h1[k];
    }
}

template <class T4>
__global__ void k_copy_state(T4* x0, T4* v0, T4* w0, const T4* x1, const T4* v1, const T4* w1,
                             int n, uint8_t* fl0, const uint8_t* fl1, T4* h0, const T4* h1,
                             const int* nb_start, const Ctl* ctl, int honor_abort) {
    if (honor_abort && ctl->abort) return;
    copy_state_items<T4>(blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, x0, v0, w0,
                         x1, v1, w1, n, fl0, fl1, h0, h1, nb_start);
}

// The fused run's wall histories, set W.in -> set W.out (with k_copy_state):
// entries flagged possibly nonzero are copied, an out entry whose in entry
// is zero is cleared where it may be nonzero, the rest are left as they are
// (zero on both sides).
__global__ void k_copy_walls(WallRefs W, int n, const Ctl* ctl, int honor_abort) {
    if (honor_abort && ctl->abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned nz0 = W.nz[i];
    unsigned nz = nz0;
    for (int q = 0; q < W.nwall; ++q) {
        const unsigned bin = wall_nz_bit(q, W.in), bout = wall_nz_bit(q, W.out);
        if (nz & bin) {
            W.h_out[q][i] = W.h_in[q][i];
            nz |= bout;
        } else if (nz & bout) {
            W.h_out[q][i] = make_double4(0.0, 0.0, 0.0, 0.0);
            nz &= ~bout;
        }
    }
    if (nz != nz0) W.nz[i] = (uint8_t)nz;
}

// WallRefs::nz from the walls' two history sets, at the start of a fused run
// (other code writes the histories densely and keeps no flags).
__global__ void k_walls_nz(WallRefs W, const double4* h0_0, const double4* h0_1,
                           const double4* h1_0, const double4* h1_1, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4* h[TMPS_RUN_WALLS][2] = {{h0_0, h0_1}, {h1_0, h1_1}};
    unsigned nz = 0;
    for (int q = 0; q < W.nwall; ++q)
        for (int b = 0; b < 2; ++b) {
            const double4 e = h[q][b][i];
            // bitwise: a negative zero counts as nonzero (a set bit may cover a zero)
            const unsigned long long a = __double_as_longlong(e.x) | __double_as_longlong(e.y) |
                                         __double_as_longlong(e.z) | __double_as_longlong(e.w);
            if (a) nz |= wall_nz_bit(q, b);
        }
    W.nz[i] = (uint8_t)nz;
}


// ------------------------------------------------ batched redo (batch.hip)
//
// The control-block steps, buffer-set copy and step end of ensemble members'
// redone rebuild steps, many members per launch (the bodies of the
// single-system kernels above).

// entries that resume from set 1: set 1 -> set 0 (honor_abort 0)
__global__ void kb_redo_copy_state(const RedoEntry* T, const int2* map) {
    const int2 mb = map[blockIdx.x];
    const RedoEntry& e = T[mb.x];
    copy_state_items<float4>(mb.y * TMPS_BLOCK + threadIdx.x, e.copy_blocks * TMPS_BLOCK, e.up,
                             e.vmf, e.omf, e.up2, e.vmf2, e.omf2, e.n, e.nb_flag, e.nb_flag2,
                             e.nb_hist, e.nb_hist2, e.nb_start);
}

// one block per entry (k_ctl_resume)
__global__ void kb_redo_resume(const RedoEntry* T) {
    if (threadIdx.x != 0) return;
    Ctl* ctl = T[blockIdx.x].ctl;
    ctl->trigger = -1;
    ctl->abort = 0;
}

// The build record (k_ctl_built), then the resume after the rebuild
// (k_ctl_rebuilt). Between the two, the single-system redo runs the split
// encoding, the slot codes and the history restore, which read neither
// what k_ctl_rebuilt writes nor anything it derives from.
__global__ void kb_redo_built(const RedoEntry* T) {
    if (threadIdx.x != 0) return;
    const RedoEntry& e = T[blockIdx.x];
    ctl_built(e.ctl, e.step, make_int3(e.fast[0], e.fast[1], e.fast[2]),
              make_int3(e.gi[0], e.gi[1], e.gi[2]), 1);
    ctl_rebuilt(e.ctl, e.step, 1);
}

// one block per entry (k_step_end with advance)
template <int L>
__global__ void __launch_bounds__(FINAL_THREADS) kb_redo_end(const RedoEntry* T, double dt) {
    const RedoEntry& e = T[blockIdx.x];
    const int nparts = ((e.n > 0 ? e.n : 1) + TMPS_BLOCK - 1) / TMPS_BLOCK;
    step_end_block<L>(e.ctl, e.partial, nparts, dt, 1, e.servo_on, e.n);
}

// The pending remap cleared unless the step aborted (k_remap_clear), and the
// member rejoins the segments: its next step reads set 0 after a half-kick.
__global__ void kb_redo_rejoin(const RedoEntry* T) {
    if (threadIdx.x != 0) return;
    const RedoEntry& e = T[blockIdx.x];
    Ctl* ctl = e.ctl;
    if (!ctl->abort) ctl->remap.step = -1;
    ctl->bpar = 0;
    ctl->bkick = 1;
    ctl->bdone = e.done;
}

// ---------------------------------------------------- per-contact computes

// Domain.minimum_image, orthogonal branch (domain.py:481-486), with the image
// count taken as min_image_vshift takes it.
__device__ __forceinline__ void min_image(const ImgP& B, double d[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (B.per[k] == 1) d[k] -= round(B.fast[k] ? d[k] * B.inv[k] : d[k] / B.prd[k]) * B.prd[k];
}

// Separation of the pair (a, b) = (min(i, j), max(i, j)) at the positions in
// xr, as the half-list computes take it: d = x_a - x_b, minimum image.
__device__ __forceinline__ double pair_sep(const ImgP& B, const double4& xa, const double4& xb,
                                           double d[3]) {
    d[0] = xa.x - xb.x;
    d[1] = xa.y - xb.y;
    d[2] = xa.z - xb.z;
    min_image(B, d);
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// ComputeContactAtom.compute_kernel, contact_atom.py:19-52: the partners of
// atom i that overlap it at the current positions (rsq < (ri + rj)^2 and
// rsq > 1e-18). The full-list row holds each pair from both sides, so the
// count is the reference's two-sided tally of its half list, without atomics.
__global__ void k_contact_atom(const double4* xr, const int* nb_start, const int* nb_j,
                               const Ctl* ctl, int n, int* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ImgP B = img_params(ctl);
    const double4 xi = xr[i];
    int c = 0;
    for (int s = nb_start[i], e = nb_start[i + 1]; s < e; ++s) {
        const int j = nb_j[s];
        const double4 xj = xr[j];
        const bool first = i < j;
        double d[3];
        const double rsq = pair_sep(B, first ? xi : xj, first ? xj : xi, d);
        const double radsum = first ? xi.w + xj.w : xj.w + xi.w;
        if (rsq < radsum * radsum && rsq > 1e-18) ++c;
    }
    count[i] = c;
}

// ComputeFabric.compute_kernel, fabric.py:26-78: per atom i, the sums of
// n_a n_b over its touching partners j > i (ascending j; n = d / |d|) and
// their number, as leaves of one pairwise tree per block.
__global__ void k_fabric_leaves(const double4* xr, const int* nb_start, const int* nb_j,
                                const Ctl* ctl, int n, double* partial) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const ImgP B = img_params(ctl);
        const double4 xi = xr[i];
        for (int s = nb_start[i], e = nb_start[i + 1]; s < e; ++s) {
            const int j = nb_j[s];
            if (j <= i) continue;
            const double4 xj = xr[j];
            double d[3];
            const double rsq = pair_sep(B, xi, xj, d);
            const double radsum = xi.w + xj.w;
            if (rsq < radsum * radsum && rsq > 1e-18) {
                const double r = sqrt(rsq);
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
    block_tree_sum<7>(leaf, partial + (size_t)blockIdx.x * 8);
}

// ComputeStressAtom.compute_stress_kernel, stress_atom.py:30-52: s = -virial
// (- m v_a v_b with keflag) per atom, and the pairwise-tree sum of s over the
// atoms (the macro stress before the division by the volume). Velocities come
// from vm64 (FP64 systems) or vm32 (FP32-split systems); .w is the mass.
__global__ void k_stress_atom(const double* vir, const double4* vm64, const float4* vm32,
                              int keflag, int n, double* stress, double* partial) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] = -vir[6 * (size_t)i + k];
        if (keflag) {
            double m, vx, vy, vz;
            if (vm64) {
                const double4 v = vm64[i];
                m = v.w; vx = v.x; vy = v.y; vz = v.z;
            } else {
                const float4 v = vm32[i];
                m = v.w; vx = v.x; vy = v.y; vz = v.z;
            }
            s[0] = s[0] - m * vx * vx;
            s[1] = s[1] - m * vy * vy;
            s[2] = s[2] - m * vz * vz;
            s[3] = s[3] - m * vx * vy;
            s[4] = s[4] - m * vx * vz;
            s[5] = s[5] - m * vy * vz;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) stress[6 * (size_t)i + k] = s[k];
    }
    block_tree_sum<6>(s, partial + (size_t)blockIdx.x * 8);
}

// ------------------------------------------------------------------ walls

// FixWallGran.post_force_kernel, wall_gran.py:120-211 (wall_plane_contact),
// and FixWallGranRegion.post_force_kernel, wall_gran_region.py:114-271
// (wall_cyl_contact), one thread per atom: the wall's force and torque added
// to the atom's, the shear history (xyz) and touch flag (w) in wh[i].
// V: double4 (FP64 system) or float4 (FP32-split: v, omega, force and torque
// in FP32, positions decoded to FP64 in xr); the contact is FP64 either way.
template <typename V>
__device__ __forceinline__ void wall_stage_add(V* f, V* tq, int i, const double dF[3],
                                               const double dT[3]) {
    V fo = f[i], to = tq[i];
    fo.x = fo.x + dF[0];
    fo.y = fo.y + dF[1];
    fo.z = fo.z + dF[2];
    to.x = to.x + dT[0];
    to.y = to.y + dT[1];
    to.z = to.z + dT[2];
    f[i] = fo;
    tq[i] = to;
}

template <typename V>
__global__ void k_wall_gran(const double4* xr, const V* vm, const V* om, V* f, V* tq,
                            double4* wh, int n, tmps_wall_params P, double dt,
                            double coord, double vwall_n, double vwall_s, int history_update) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 x4 = xr[i];
    const double xs[3] = {x4.x, x4.y, x4.z};
    const V v4 = vm[i], o4 = om[i];
    const double v[3] = {v4.x, v4.y, v4.z}, w[3] = {o4.x, o4.y, o4.z};
    double dF[3], dT[3];
    double4 hn;
    const bool touched = wall_plane_contact(P, coord, vwall_n, vwall_s, xs, x4.w, v4.w, v, w, dt,
                                            history_update, wh + i, hn, dF, dT);
    if (P.use_history) wh[i] = hn;
    if (touched) wall_stage_add(f, tq, i, dF, dT);
}

template <typename V>
__global__ void k_wall_cyl(const double4* xr, const V* vm, const V* om, V* f, V* tq,
                           double4* wh, int n, tmps_wall_params P, double dt,
                           int history_update) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 x4 = xr[i];
    const double pos[3] = {x4.x, x4.y, x4.z};
    const V v4 = vm[i], o4 = om[i];
    const double vi[3] = {v4.x, v4.y, v4.z}, oi[3] = {o4.x, o4.y, o4.z};
    double dF[3], dT[3];
    double4 hn;
    const bool touched =
        wall_cyl_contact(P, pos, x4.w, v4.w, vi, oi, dt, history_update, wh + i, hn, dF, dT);
    if (P.use_history) wh[i] = hn;
    if (touched) wall_stage_add(f, tq, i, dF, dT);
}

// The final pairwise tree over nparts block partials, NC components, into out.
template <int L, int NC>
__global__ void __launch_bounds__(FINAL_THREADS) k_final_tree_out(const double* partial,
                                                                   int nparts, double* out) {
    double o[NC];
    final_tree_L<L, NC>(partial, nparts, o);
    if (threadIdx.x == 0)
        for (int k = 0; k < NC; ++k) out[k] = o[k];
}

}  // namespace

// ================================================================ launchers

#define LAUNCH_N(kernel, n, ...)                                                          \
    do {                                                                                  \
        if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(nblocks_for(n)), dim3(TMPS_BLOCK), 0, \
                                        s.stream, __VA_ARGS__);                           \
        HIP_TRY(hipGetLastError());                                                       \
    } while (0)

int launch_clear_f64(System& s) {
    LAUNCH_N(k_clear_f64, s.n, ptrs(s));
    return 0;
}

int launch_nve_initial_f64(System& s, double dt) {
    LAUNCH_N(k_nve_initial_f64, s.n, ptrs(s), dt);
    return 0;
}

int launch_nve_final_f64(System& s, double dt) {
    LAUNCH_N(k_nve_final_f64, s.n, ptrs(s), dt);
    return 0;
}

int launch_cundall_f64(System& s, double gl, double ga) {
    LAUNCH_N(k_cundall_f64, s.n, ptrs(s), gl, ga);
    return 0;
}

int launch_pair_f64(System& s, double dt, int shearupdate, int accumulate) {
    switch (s.pair.style) {
        case TMPS_PAIR_HOOKE_HISTORY:
            LAUNCH_N(k_pair_f64<TMPS_PAIR_HOOKE_HISTORY>, s.n, ptrs(s), s.ctl, s.pair, dt,
                     shearupdate, accumulate);
            break;
        case TMPS_PAIR_HERTZ_HISTORY:
            LAUNCH_N(k_pair_f64<TMPS_PAIR_HERTZ_HISTORY>, s.n, ptrs(s), s.ctl, s.pair, dt,
                     shearupdate, accumulate);
            break;
        default:
            LAUNCH_N(k_pair_f64<TMPS_PAIR_GRANULAR>, s.n, ptrs(s), s.ctl, s.pair, dt, shearupdate,
                     accumulate);
            break;
    }
    return 0;
}

int launch_check_f64(System& s) {
    LAUNCH_N(k_check_f64, s.n, ptrs(s), s.ctl);
    return 0;
}

#define TREE_CASE(Lv)                                                                          \
    case Lv:                                                                                   \
        if (nc == 2)                                                                           \
            hipLaunchKernelGGL((k_final_tree<Lv, 2>), dim3(1), dim3(FINAL_THREADS), 0, s.stream, \
                               s.ctl, s.partial, nparts);                                      \
        else                                                                                   \
            hipLaunchKernelGGL((k_final_tree<Lv, 6>), dim3(1), dim3(FINAL_THREADS), 0, s.stream, \
                               s.ctl, s.partial, nparts);                                      \
        break;

int launch_final_tree(System& s, int nc, int nparts) {
    switch (final_leaves(nparts)) {
        TREE_CASE(1) TREE_CASE(2) TREE_CASE(4) TREE_CASE(8) TREE_CASE(16) TREE_CASE(32)
        TREE_CASE(64)
        default: return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for the reduction tree");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Block partials of the per-atom virial (+ kinetic) leaves; the caller
// finishes the tree (launch_final_tree) or runs the servo (launch_step_end).
int launch_leaves_f64(System& s, int keflag) {
    const int nb = nblocks_for(s.n);
    hipLaunchKernelGGL(k_leaves_f64, dim3(nb > 0 ? nb : 1), dim3(TMPS_BLOCK), 0, s.stream, ptrs(s),
                       keflag, s.partial);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_ke_f64(System& s, double* out2) {
    (void)out2;
    const int nb = nblocks_for(s.n);
    hipLaunchKernelGGL(k_ke_leaves_f64, dim3(nb > 0 ? nb : 1), dim3(TMPS_BLOCK), 0, s.stream,
                       ptrs(s), s.partial);
    HIP_TRY(hipGetLastError());
    return launch_final_tree(s, 2, nb > 0 ? nb : 1);
}

int launch_step_begin_f64(System& s, double dt, int kick, int buf) {
    LAUNCH_N(k_step_begin_f64, s.n, ptrs(s, buf, buf), s.ctl, dt, kick);
    return 0;
}

// ---------------------------------------------- fields through device memory

// A field's values into a caller's device buffer as tmps_get_field lays them
// out: x, y, z of each record (count x 3 doubles) or its w (scalar), from
// FP64 or FP32 records.
__global__ void k_field_pack(const double4* d4, const float4* f4, int n, int scalar,
                             double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x, y, z, w;
    if (d4) {
        const double4 a = d4[i];
        x = a.x; y = a.y; z = a.z; w = a.w;
    } else {
        const float4 a = f4[i];
        x = a.x; y = a.y; z = a.z; w = a.w;
    }
    if (scalar) {
        out[i] = w;
    } else {
        out[3 * i] = x;
        out[3 * i + 1] = y;
        out[3 * i + 2] = z;
    }
}

// A caller's device buffer of count x 3 doubles into the x, y, z of FP64 or
// FP32 records (rounded to FP32 as the host setter does); w kept, or taken
// from wsrc.
__global__ void k_field_unpack(double4* d4, float4* f4, int n, const double* in,
                               const double* wsrc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (d4) {
        double4 a = d4[i];
        a.x = in[3 * i];
        a.y = in[3 * i + 1];
        a.z = in[3 * i + 2];
        if (wsrc) a.w = wsrc[i];
        d4[i] = a;
    } else {
        float4 a = f4[i];
        a.x = (float)in[3 * i];
        a.y = (float)in[3 * i + 1];
        a.z = (float)in[3 * i + 2];
        f4[i] = a;
    }
}

int launch_field_pack(System& s, const double4* d4, const float4* f4, int n, int scalar,
                      double* out) {
    LAUNCH_N(k_field_pack, n, d4, f4, n, scalar, out);
    return 0;
}

int launch_field_unpack(System& s, double4* d4, float4* f4, int n, const double* in,
                        const double* wsrc) {
    LAUNCH_N(k_field_unpack, n, d4, f4, n, in, wsrc);
    return 0;
}

int launch_redo_begin_f64(System& s, double dt) {
    LAUNCH_N(k_redo_begin_f64, s.n, ptrs(s, 0, 0), s.ctl, dt);
    return 0;
}

int launch_copy_vw(System& s, int honor_abort) {
    LAUNCH_N(k_copy_vw, s.n, s.vm, s.om, s.vm2, s.om2, s.n, s.ctl, honor_abort);
    return 0;
}

template <int STYLE, int WALL>
static int launch_force_walls(System& s, const F64Ptrs& P, double dt, int final_mode, int cundall,
                              double gl, double ga) {
    const dim3 g(nblocks_for(s.n)), b(TMPS_BLOCK);
    if (final_mode == 2) {
        if (cundall)
            hipLaunchKernelGGL((k_step_force_f64<STYLE, 1, 2, WALL>), g, b, 0, s.stream, P, s.ctl,
                               s.pair, dt, gl, ga, s.partial);
        else
            hipLaunchKernelGGL((k_step_force_f64<STYLE, 0, 2, WALL>), g, b, 0, s.stream, P, s.ctl,
                               s.pair, dt, gl, ga, s.partial);
    } else if (final_mode) {
        if (cundall)
            hipLaunchKernelGGL((k_step_force_f64<STYLE, 1, 1, WALL>), g, b, 0, s.stream, P, s.ctl,
                               s.pair, dt, gl, ga, s.partial);
        else
            hipLaunchKernelGGL((k_step_force_f64<STYLE, 0, 1, WALL>), g, b, 0, s.stream, P, s.ctl,
                               s.pair, dt, gl, ga, s.partial);
    } else {
        if (cundall)
            hipLaunchKernelGGL((k_step_force_f64<STYLE, 1, 0, WALL>), g, b, 0, s.stream, P, s.ctl,
                               s.pair, dt, gl, ga, s.partial);
        else
            hipLaunchKernelGGL((k_step_force_f64<STYLE, 0, 0, WALL>), g, b, 0, s.stream, P, s.ctl,
                               s.pair, dt, gl, ga, s.partial);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// (a fused run's walls, when it has any, in place on set 0: the redo of a
// step runs its force pass there)
template <int STYLE>
static int launch_force_style(System& s, double dt, int final_mode, int cundall, double gl,
                              double ga, int in_buf, int out_buf) {
    F64Ptrs P = ptrs(s, in_buf, out_buf);
    if (s.nrun_walls > 0) {
        P.W = run_walls(s, 0, 0);
        return launch_force_walls<STYLE, 1>(s, P, dt, final_mode, cundall, gl, ga);
    }
    return launch_force_walls<STYLE, 0>(s, P, dt, final_mode, cundall, gl, ga);
}

int launch_step_force_f64(System& s, double dt, int final_mode, int cundall, double gl,
                          double ga, int in_buf, int out_buf) {
    if (s.n <= 0) return 0;
    if (in_buf == out_buf) return tmps_fail(TMPS_ERR_ARG, "force pass cannot update v in place");
    switch (s.pair.style) {
        case TMPS_PAIR_HOOKE_HISTORY:
            return launch_force_style<TMPS_PAIR_HOOKE_HISTORY>(s, dt, final_mode, cundall, gl, ga,
                                                               in_buf, out_buf);
        case TMPS_PAIR_HERTZ_HISTORY:
            return launch_force_style<TMPS_PAIR_HERTZ_HISTORY>(s, dt, final_mode, cundall, gl, ga,
                                                               in_buf, out_buf);
        default:
            return launch_force_style<TMPS_PAIR_GRANULAR>(s, dt, final_mode, cundall, gl, ga,
                                                          in_buf, out_buf);
    }
}

#define END_CASE(Lv)                                                                          \
    case Lv:                                                                                  \
        hipLaunchKernelGGL((k_step_end<Lv>), dim3(1), dim3(CLOSE_THREADS), 0, s.stream, s.ctl, \
                           src, nparts, dt, advance, apply_servo, s.n);                       \
        break;

int launch_step_end(System& s, double dt, int nparts, int advance, int apply_servo) {
    if (nparts < 1) nparts = 1;
    const double* src = s.partial;
    if (nparts > CLOSE_SPAN) {
        const int groups = (nparts + CLOSE_SPAN - 1) / CLOSE_SPAN;
        if (groups > CLOSE_SPAN || groups > s.level_cap)
            return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for the reduction tree");
        hipLaunchKernelGGL(k_tree_level, dim3(groups), dim3(CLOSE_THREADS), 0, s.stream, s.partial,
                           nparts, s.level);
        HIP_TRY(hipGetLastError());
        src = s.level;
        nparts = groups;
    }
    switch (final_leaves(nparts, CLOSE_THREADS)) {
        END_CASE(1) END_CASE(2) END_CASE(4) END_CASE(8)
        default: return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for the reduction tree");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

#define CLOSE_CASE(Lv)                                                                          \
    case Lv:                                                                                    \
        hipLaunchKernelGGL((k_step_close<Lv>), dim3(1), dim3(CLOSE_THREADS), 0, s.stream, s.ctl, \
                           src, nparts, dt, servo_on, s.n);                                     \
        break;

int launch_step_close(System& s, double dt, int servo_on) {
    int nparts = nblocks_for(std::max(s.n, 1));
    const double* src = s.partial;
    if (nparts > CLOSE_SPAN) {
        const int groups = (nparts + CLOSE_SPAN - 1) / CLOSE_SPAN;
        if (groups > CLOSE_SPAN || groups > s.level_cap)
            return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for the reduction tree");
        hipLaunchKernelGGL(k_tree_level, dim3(groups), dim3(CLOSE_THREADS), 0, s.stream, s.partial,
                           nparts, s.level);
        HIP_TRY(hipGetLastError());
        src = s.level;
        nparts = groups;
    }
    switch (final_leaves(nparts, CLOSE_THREADS)) {
        CLOSE_CASE(1) CLOSE_CASE(2) CLOSE_CASE(4) CLOSE_CASE(8)
        default: return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for the reduction tree");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

#define BATCH_CLOSE_CASE(Lv)                                                                 \
    case Lv:                                                                                 \
        hipLaunchKernelGGL((k_batch_close<Lv>), dim3(ninst), dim3(CLOSE_THREADS), 0, st, table, \
                           dt);                                                              \
        break;

int launch_batch_close(const CloseEntry* table, int ninst, int max_parts, double dt,
                       hipStream_t st) {
    if (ninst <= 0) return 0;
    if (max_parts > CLOSE_SPAN)
        return tmps_fail(TMPS_ERR_CAPACITY, "an ensemble member has too many atoms");
    switch (final_leaves(max_parts, CLOSE_THREADS)) {
        BATCH_CLOSE_CASE(1) BATCH_CLOSE_CASE(2) BATCH_CLOSE_CASE(4) BATCH_CLOSE_CASE(8)
        default: return tmps_fail(TMPS_ERR_CAPACITY, "an ensemble member has too many atoms");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_remap_apply(System& s) {
    if (s.precision == TMPS_FP64) {
        LAUNCH_N(k_remap_apply, s.n, ptrs(s), s.ctl, 0);
    }
    return 0;
}

int launch_remap_apply_fused(System& s) {
    if (s.precision == TMPS_FP64) {
        LAUNCH_N(k_remap_apply, s.n, ptrs(s), s.ctl, 1);
    }
    return 0;
}

int launch_remap_clear(System& s, int honor_abort) {
    hipLaunchKernelGGL(k_remap_clear, dim3(1), dim3(64), 0, s.stream, s.ctl, honor_abort);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_ctl_rebuilt(System& s, int step) {
    hipLaunchKernelGGL(k_ctl_rebuilt, dim3(1), dim3(64), 0, s.stream, s.ctl, step,
                       s.precision == TMPS_FP32_SPLIT ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_ctl_store(System& s) {
    hipLaunchKernelGGL(k_ctl_store, dim3(1), dim3(256), 0, s.stream, s.ctl, *s.hctl);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_ctl_resume(System& s) {
    hipLaunchKernelGGL(k_ctl_resume, dim3(1), dim3(64), 0, s.stream, s.ctl);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_ctl_built_dev(System& s, int step) {
    hipLaunchKernelGGL(k_ctl_built_dev, dim3(1), dim3(64), 0, s.stream, s.ctl, step,
                       s.precision == TMPS_FP32_SPLIT ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_ctl_built(System& s, int step, const int fast[3], const int gi[3], int split) {
    hipLaunchKernelGGL(k_ctl_built, dim3(1), dim3(64), 0, s.stream, s.ctl, step,
                       make_int3(fast[0], fast[1], fast[2]), make_int3(gi[0], gi[1], gi[2]),
                       split);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ------------------------------------------------ one kernel per step

template <int STYLE, int CUNDALL, int FINAL, int WALL>
static int fused_f64_go(System& s, const Fused64& P, double dt, double gl, double ga) {
    const int nb = nblocks_for(s.n);
    if (s.kstart || s.kstop) {
        hipExtLaunchKernelGGL((k_step_f64<STYLE, CUNDALL, FINAL, WALL>), dim3(nb),
                              dim3(TMPS_BLOCK), 0, s.stream, s.kstart, s.kstop, 0, P, s.ctl,
                              s.pair, s.partial, dt, gl, ga);
        s.kstart = s.kstop = nullptr;
    } else {
        hipLaunchKernelGGL((k_step_f64<STYLE, CUNDALL, FINAL, WALL>), dim3(nb), dim3(TMPS_BLOCK),
                           0, s.stream, P, s.ctl, s.pair, s.partial, dt, gl, ga);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int STYLE, int WALL>
static int fused_f64_walls(System& s, const Fused64& P, double dt, int fin, int cundall, double gl,
                           double ga) {
    if (fin) {
        if (cundall) return fused_f64_go<STYLE, 1, 1, WALL>(s, P, dt, gl, ga);
        return fused_f64_go<STYLE, 0, 1, WALL>(s, P, dt, gl, ga);
    }
    if (cundall) return fused_f64_go<STYLE, 1, 0, WALL>(s, P, dt, gl, ga);
    return fused_f64_go<STYLE, 0, 0, WALL>(s, P, dt, gl, ga);
}

template <int STYLE>
static int fused_f64_style(System& s, const Fused64& P, double dt, int fin, int cundall, double gl,
                           double ga) {
    if (P.W.nwall > 0) return fused_f64_walls<STYLE, 1>(s, P, dt, fin, cundall, gl, ga);
    return fused_f64_walls<STYLE, 0>(s, P, dt, fin, cundall, gl, ga);
}

int launch_fused_f64(System& s, int in_buf, double dt, int fin, int cundall, double gl,
                     double ga) {
    if (s.n <= 0) return 0;
    if (nblocks_for(s.n) > (1 << 30) / TMPS_BLOCK)
        return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for one step launch");
    const int ob = 1 - in_buf;
    Fused64 P;
    P.xr_in = in_buf ? s.xr2 : s.xr;
    P.vm_in = in_buf ? s.vm2 : s.vm;
    P.om_in = in_buf ? s.om2 : s.om;
    P.xr_out = ob ? s.xr2 : s.xr;
    P.vm_out = ob ? s.vm2 : s.vm;
    P.om_out = ob ? s.om2 : s.om;
    P.fl_in = in_buf ? s.nb_flag2 : s.nb_flag;
    P.fl_out = ob ? s.nb_flag2 : s.nb_flag;
    P.h_in = static_cast<const double4*>(in_buf ? s.nb_hist2 : s.nb_hist);
    P.h_out = static_cast<double4*>(ob ? s.nb_hist2 : s.nb_hist);
    P.f = s.f;
    P.tq = s.tq;
    P.x0 = s.x0;
    P.vir = s.vir;
    P.nb_start = s.nb_start;
    P.nb_j = s.nb_j;
    P.n = s.n;
    P.W = run_walls(s, in_buf, ob);
    switch (s.pair.style) {
        case TMPS_PAIR_HOOKE_HISTORY:
            return fused_f64_style<TMPS_PAIR_HOOKE_HISTORY>(s, P, dt, fin, cundall, gl, ga);
        case TMPS_PAIR_HERTZ_HISTORY:
            return fused_f64_style<TMPS_PAIR_HERTZ_HISTORY>(s, P, dt, fin, cundall, gl, ga);
        default:
            return fused_f64_style<TMPS_PAIR_GRANULAR>(s, P, dt, fin, cundall, gl, ga);
    }
}

int launch_kick_f64(System& s, double dt) {
    LAUNCH_N(k_kick_f64, s.n, s.vm, s.om, s.xr, s.f, s.tq, s.n, s.ctl, dt);
    return 0;
}

int launch_copy_state(System& s, int honor_abort) {
    if (s.n <= 0) return 0;
    // grid-stride over atoms and slots; sized for the list's capacity, since
    // the slot count can grow while a captured segment is replayed
    const int work = (int)std::min<size_t>(std::max<size_t>((size_t)s.n, s.nb_cap), 1u << 30);
    const int nb = std::min(nblocks_for(work), 4096);
    if (s.precision == TMPS_FP64)
        hipLaunchKernelGGL(k_copy_state<double4>, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.xr,
                           s.vm, s.om, s.xr2, s.vm2, s.om2, s.n, s.nb_flag, s.nb_flag2,
                           static_cast<double4*>(s.nb_hist), static_cast<const double4*>(s.nb_hist2),
                           (const int*)s.nb_start, s.ctl, honor_abort);
    else
        hipLaunchKernelGGL(k_copy_state<float4>, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.up,
                           s.vmf, s.omf, s.up2, s.vmf2, s.omf2, s.n, s.nb_flag, s.nb_flag2,
                           static_cast<float4*>(s.nb_hist), static_cast<const float4*>(s.nb_hist2),
                           (const int*)s.nb_start, s.ctl, honor_abort);
    HIP_TRY(hipGetLastError());
    if (s.nrun_walls > 0)
        LAUNCH_N(k_copy_walls, s.n, run_walls(s, 1, 0), s.n, s.ctl, honor_abort);
    return 0;
}

int launch_walls_nz(System& s) {
    if (s.nrun_walls <= 0 || s.n <= 0) return 0;
    static_assert(TMPS_RUN_WALLS == 2, "k_walls_nz takes two walls' histories");
    const WallRefs W = run_walls(s, 0, 1);
    LAUNCH_N(k_walls_nz, s.n, W, s.run_wh[0][0], s.run_wh[0][1], s.run_wh[1][0], s.run_wh[1][1],
             s.n);
    return 0;
}

// ------------------------------------------------ per-contact computes

#define TREE_OUT_CASE(Lv)                                                                     \
    case Lv:                                                                                  \
        if (nc == 6)                                                                          \
            hipLaunchKernelGGL((k_final_tree_out<Lv, 6>), dim3(1), dim3(FINAL_THREADS), 0,     \
                               s.stream, s.partial, nparts, out);                             \
        else                                                                                  \
            hipLaunchKernelGGL((k_final_tree_out<Lv, 7>), dim3(1), dim3(FINAL_THREADS), 0,     \
                               s.stream, s.partial, nparts, out);                             \
        break;

// out: 8 doubles. Up to CLOSE_SPAN partials one k_tree_level block (coalesced
// loads, LDS transpose) reduces all 8 components; above, final_tree_L's
// kernel. Both are the same pairwise tree, so the bits agree.
static int launch_tree_out(System& s, int nc, int nparts, double* out) {
    if (nparts <= CLOSE_SPAN) {
        hipLaunchKernelGGL(k_tree_level, dim3(1), dim3(CLOSE_THREADS), 0, s.stream, s.partial,
                           nparts, out);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    switch (final_leaves(nparts)) {
        TREE_OUT_CASE(1) TREE_OUT_CASE(2) TREE_OUT_CASE(4) TREE_OUT_CASE(8) TREE_OUT_CASE(16)
        TREE_OUT_CASE(32) TREE_OUT_CASE(64)
        default: return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for the reduction tree");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_contact_atom(System& s, int* d_count) {
    LAUNCH_N(k_contact_atom, s.n, s.xr, s.nb_start, s.nb_j, s.ctl, s.n, d_count);
    return 0;
}

int launch_fabric(System& s, double* d_sum7) {
    const int nb = std::max(nblocks_for(s.n), 1);
    hipLaunchKernelGGL(k_fabric_leaves, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.xr, s.nb_start,
                       s.nb_j, s.ctl, s.n, s.partial);
    HIP_TRY(hipGetLastError());
    return launch_tree_out(s, 7, nb, d_sum7);
}

int launch_fabric_into(System& s, double* partial, double* d_sum7) {
    const int nb = std::max(nblocks_for(s.n), 1);
    if (nb > CLOSE_SPAN) return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for one tree block");
    hipLaunchKernelGGL(k_fabric_leaves, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.xr, s.nb_start,
                       s.nb_j, s.ctl, s.n, partial);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_tree_level, dim3(1), dim3(CLOSE_THREADS), 0, s.stream, partial, nb,
                       d_sum7);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_stress_atom(System& s, int keflag, double* d_stress, double* d_sum6) {
    const int nb = std::max(nblocks_for(s.n), 1);
    const bool f64 = s.precision == TMPS_FP64;
    hipLaunchKernelGGL(k_stress_atom, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.vir,
                       f64 ? s.vm : nullptr, f64 ? nullptr : s.vmf, keflag, s.n, d_stress,
                       s.partial);
    HIP_TRY(hipGetLastError());
    return launch_tree_out(s, 6, nb, d_sum6);
}

int launch_wall_gran(System& s, double4* wh, const tmps_wall_params& p, double dt, double coord,
                     double vwall_n, double vwall_s, int history_update) {
    const bool f64 = s.precision == TMPS_FP64;
    if (p.geometry == TMPS_WALL_CYLINDER) {
        if (f64)
            LAUNCH_N(k_wall_cyl<double4>, s.n, s.xr, s.vm, s.om, s.f, s.tq, wh, s.n, p, dt,
                     history_update);
        else
            LAUNCH_N(k_wall_cyl<float4>, s.n, s.xr, s.vmf, s.omf, s.ff, s.tqf, wh, s.n, p, dt,
                     history_update);
    } else {
        if (f64)
            LAUNCH_N(k_wall_gran<double4>, s.n, s.xr, s.vm, s.om, s.f, s.tq, wh, s.n, p, dt, coord,
                     vwall_n, vwall_s, history_update);
        else
            LAUNCH_N(k_wall_gran<float4>, s.n, s.xr, s.vmf, s.omf, s.ff, s.tqf, wh, s.n, p, dt,
                     coord, vwall_n, vwall_s, history_update);
    }
    return 0;
}

// ------------------------------------------------ batched redo (batch.hip)

int launch_redo_resume(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    if (M.n_copy > 0) {
        hipLaunchKernelGGL(kb_redo_copy_state, dim3(M.n_copy), dim3(TMPS_BLOCK), 0, st, T, M.copy);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(kb_redo_resume, dim3(M.entries), dim3(64), 0, st, T);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_redo_built(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    hipLaunchKernelGGL(kb_redo_built, dim3(M.entries), dim3(64), 0, st, T);
    HIP_TRY(hipGetLastError());
    return 0;
}

#define REDO_END_CASE(Lv)                                                                       \
    case Lv:                                                                                    \
        hipLaunchKernelGGL((kb_redo_end<Lv>), dim3(M.entries), dim3(FINAL_THREADS), 0, st, T, dt); \
        break;

// The tree's leaves per thread follow the largest member; zero padding and
// aligned subtrees leave every member's tree its own.
int launch_redo_end(const RedoEntry* T, const RedoMaps& M, double dt, hipStream_t st) {
    switch (final_leaves(std::max(M.max_parts, 1))) {
        REDO_END_CASE(1) REDO_END_CASE(2) REDO_END_CASE(4) REDO_END_CASE(8) REDO_END_CASE(16)
        REDO_END_CASE(32) REDO_END_CASE(64)
        default: return tmps_fail(TMPS_ERR_CAPACITY, "too many atoms for the reduction tree");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_redo_rejoin(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    hipLaunchKernelGGL(kb_redo_rejoin, dim3(M.entries), dim3(64), 0, st, T);
    HIP_TRY(hipGetLastError());
    return 0;
}
