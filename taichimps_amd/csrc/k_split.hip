// FP32-split kernels of the DEM step.
//
// Position encoding:  x = lo + (prd / g) * (c + 0.5 + u)
//   c  integer bin of the last neighbor build (packed 11/11/10 bits),
//   u  FP32 offset in bin widths (|u| <= 0.5 + drift since the build).
// The affine `remap x` of fix deform/pressure keeps every atom's fractional
// position, so under it the encoding needs no per-atom update at all: the new
// box alone moves the atoms. Contact and integration arithmetic is FP32; the
// per-atom virial, the kinetic tensor, their reduction and the servo stay FP64.
#include <hip/hip_ext.h>

#include <cstring>

#include "contact.h"
#include "fused.h"
#include "pbc.h"
#include "tmps_internal.h"
#include "wall.h"

using namespace tmps;

namespace {

struct SplitPtrs {
    float4* up;   // u.xyz, radius
    float4* vm;   // v.xyz, rmass (read by every contact of the step)
    float4* om;   // omega.xyz, 1 / (0.4 m r^2)
    float4* vm_out;  // where a fused force pass writes v / omega
    float4* om_out;
    float4* f;
    float4* tq;
    float4* u0;   // u at the last build
    const uint32_t* cell;
    double* vir;  // [n][6]
    const int* nb_start;
    const int* nb_j;
    uint8_t* nb_flag;
    float4* nb_hist;
    int n;
    WallRefs W;  // the fused run's walls (nwall 0 elsewhere)
};

SplitPtrs sptrs(System& s, int in_buf = 0, int out_buf = 0) {
    SplitPtrs p;
    p.up = s.up; p.f = s.ff; p.tq = s.tqf; p.u0 = s.u0;
    p.vm = in_buf ? s.vmf2 : s.vmf;
    p.om = in_buf ? s.omf2 : s.omf;
    p.vm_out = out_buf ? s.vmf2 : s.vmf;
    p.om_out = out_buf ? s.omf2 : s.omf;
    p.cell = s.cell; p.vir = s.vir;
    p.nb_start = s.nb_start; p.nb_j = s.nb_j; p.nb_flag = s.nb_flag;
    p.nb_hist = static_cast<float4*>(s.nb_hist);
    p.n = s.n;
    p.W = WallRefs{};
    return p;
}

__device__ __forceinline__ void unpack_cell(uint32_t c, int& cx, int& cy, int& cz) {
    cx = (int)(c & 0x7FFu);
    cy = (int)((c >> 11) & 0x7FFu);
    cz = (int)(c >> 22);
}

// A cell difference reduced to the periodic image nearest zero.
__device__ __forceinline__ int wrap_cells(int dc, int g) {
    if (2 * dc > g) dc -= g;
    else if (2 * dc < -g) dc += g;
    return dc;
}

// Separation of a minus b in length units, minimum image taken in bin units.
__device__ __forceinline__ void split_delta(const SplitConst& sc, const int per[3], uint32_t ca,
                                            const float4& ua, uint32_t cb, const float4& ub,
                                            float d[3]) {
    int ax, ay, az, bx, by, bz;
    unpack_cell(ca, ax, ay, az);
    unpack_cell(cb, bx, by, bz);
    // the cell difference is wrapped as an integer first, so the FP32 offset
    // difference is added to a small number and keeps its precision
    int dc[3] = {ax - bx, ay - by, az - bz};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (per[k] == 1) dc[k] = wrap_cells(dc[k], sc.gi[k]);
    float e[3] = {(float)dc[0] + (ua.x - ub.x), (float)dc[1] + (ua.y - ub.y),
                  (float)dc[2] + (ua.z - ub.z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (per[k] == 1) e[k] -= sc.g[k] * rintf(e[k] * (1.0f / sc.g[k]));
        d[k] = e[k] * sc.cw[k];
    }
}

// The touch test of gather_split for slot partner j (cell cj, offset uj):
// the contact condition the contact law runs under, by its expressions.
__device__ __forceinline__ bool split_touch(const SplitConst& sc, const int per[3], int i,
                                            uint32_t ci, const float4& ui, int j, uint32_t cj,
                                            const float4& uj) {
    const bool first = i < j;
    float d[3];
    if (first) split_delta(sc, per, ci, ui, cj, uj, d);
    else split_delta(sc, per, cj, uj, ci, ui, d);
    const float rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float ra = first ? ui.w : uj.w, rb = first ? uj.w : ui.w;
    const float radsum = ra + rb;
    return rsq < radsum * radsum && rsq > 1e-28f;
}

// slots whose partner loads one gather_split batch issues together
#define GATHER_BATCH 4

template <int STYLE>
__device__ __forceinline__ void gather_split(int i, uint32_t ci, const float4& ui, const float4& vi,
                                             const float4& oi, const SplitPtrs& P,
                                             const SplitConst& sc, const int per[3],
                                             const PairP& pp, float dt, int shearupdate,
                                             float accF[3], float accT[3], double accV[6],
                                             int& ntouch) {
    const int s0 = P.nb_start[i], s1 = P.nb_start[i + 1];
    const bool uses_history = (STYLE != TMPS_PAIR_GRANULAR) || pp.uses_history;
    const float zero3[3] = {0.f, 0.f, 0.f};
    for (int sb = s0; sb < s1; sb += GATHER_BATCH) {
        // a batch of slots: the partners' offsets, cells and touch flags are
        // requested together, then the touch test; a slot that does not
        // touch has its flag cleared here
        int jb[GATHER_BATCH];
        // (components in plain arrays: an array of float4 stays in scratch)
        float ux[GATHER_BATCH], uy[GATHER_BATCH], uz[GATHER_BATCH], uw[GATHER_BATCH];
        uint32_t cb[GATHER_BATCH];
        uint8_t fb[GATHER_BATCH];
#pragma unroll
        for (int b = 0; b < GATHER_BATCH; ++b) jb[b] = sb + b < s1 ? P.nb_j[sb + b] : i;
#pragma unroll
        for (int b = 0; b < GATHER_BATCH; ++b) {
            const float4 u = P.up[jb[b]];
            ux[b] = u.x;
            uy[b] = u.y;
            uz[b] = u.z;
            uw[b] = u.w;
            cb[b] = P.cell[jb[b]];
            fb[b] = uses_history && sb + b < s1 ? P.nb_flag[sb + b] : 0;
        }
        unsigned touch = 0;
#pragma unroll
        for (int b = 0; b < GATHER_BATCH; ++b) {
            const bool in_row = sb + b < s1;
            if (in_row && split_touch(sc, per, i, ci, ui, jb[b], cb[b],
                                      make_float4(ux[b], uy[b], uz[b], uw[b])))
                touch |= 1u << b;
            else if (fb[b] != 0) P.nb_flag[sb + b] = 0;
        }
        // the contact law on the batch's touching slots, in slot order (the
        // separation recomputed by the same expressions)
        while (touch) {
            const int s = sb + __builtin_ctz(touch);
            touch &= touch - 1;
            const int j = P.nb_j[s];
            const float4 uj = P.up[j];
            const uint32_t cj = P.cell[j];
            const bool first = i < j;
            float d[3];
            if (first) split_delta(sc, per, ci, ui, cj, uj, d);
            else split_delta(sc, per, cj, uj, ci, ui, d);
            const float rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const float ra = first ? ui.w : uj.w, rb = first ? uj.w : ui.w;
            const float radsum = ra + rb;
            {
            const float4 vj = P.vm[j];
            const float4 oj = P.om[j];
            const float4 va4 = first ? vi : vj;
            const float4 vb4 = first ? vj : vi;
            const float4 oa4 = first ? oi : oj;
            const float4 ob4 = first ? oj : oi;
            const float va[3] = {va4.x, va4.y, va4.z}, vb[3] = {vb4.x, vb4.y, vb4.z};
            const float oa[3] = {oa4.x, oa4.y, oa4.z}, ob[3] = {ob4.x, ob4.y, ob4.z};
            bool before = false;
            float h[3] = {0.f, 0.f, 0.f};
            if (uses_history) {
                before = P.nb_flag[s] != 0;
                if (before) {
                    const float4 hh = P.nb_hist[s];
                    h[0] = hh.x; h[1] = hh.y; h[2] = hh.z;
                }
            }
            ContactOut<float> o;
            if (STYLE == TMPS_PAIR_HOOKE_HISTORY)
                legacy_history_contact<float, false>(
                    d, zero3, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa, ob, (float)pp.kn,
                    (float)pp.kt, (float)pp.gamman, (float)pp.gammat, (float)pp.xmu,
                    pp.limit_damping, dt, shearupdate, before, h, o);
            else if (STYLE == TMPS_PAIR_HERTZ_HISTORY)
                legacy_history_contact<float, true>(
                    d, zero3, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa, ob, (float)pp.kn,
                    (float)pp.kt, (float)pp.gamman, (float)pp.gammat, (float)pp.xmu,
                    pp.limit_damping, dt, shearupdate, before, h, o);
            else
                granular_contact<float>(d, zero3, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa, ob,
                                        pp, dt, shearupdate, before, h, o);
            if (uses_history) {
                if (!before) P.nb_flag[s] = 1;
                P.nb_hist[s] = make_float4(h[0], h[1], h[2], 0.f);
            }
            if (first) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    accF[k] += o.F[k];
                    accT[k] += o.ta[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    accF[k] -= o.F[k];
                    accT[k] += o.tb[k];
                }
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) accV[k] += (double)o.vir[k];
            ++ntouch;
            }
        }
    }
}

__device__ __forceinline__ float cundall_scale_f(float fv, float g) {
    return 1.0f - g * (fv >= 0.0f ? 1.0f : -1.0f);
}

// ---------------------------------------------------------- conversions

// x (FP64) of every atom from its split encoding, into xr (x, y, z, r).
// The FP64 position of an encoded one (cell, offset u) in the current box.
__device__ __forceinline__ void split_position(const DevBox& B, const SplitConst& sc,
                                               uint32_t cell, const float4& u, double x[3]) {
    int c[3];
    unpack_cell(cell, c[0], c[1], c[2]);
    const float uu[3] = {u.x, u.y, u.z};
    for (int k = 0; k < 3; ++k) {
        const double cw = B.prd[k] / (double)sc.gi[k];
        x[k] = B.lo[k] + cw * ((double)c[k] + 0.5 + (double)uu[k]);
    }
}

__device__ __forceinline__ void split_to_x64_atom(int i, const SplitPtrs& P, double4* xr,
                                                  const Ctl* ctl) {
    const float4 u = P.up[i];
    double x[3];
    split_position(ctl->box, ctl->sc, P.cell[i], u, x);
    xr[i] = make_double4(x[0], x[1], x[2], (double)u.w);
}

// A fused step's walls before (pass 0) or after (pass 1) damping/cundall on
// atom i, at its drifted position (cell ci, offset u), with its half-step v
// and omega (FixWallGran.post_force between the pair force and
// final_integrate, simulation.py:236-247).
__device__ __forceinline__ void split_walls(const WallRefs& W, int pass, const Ctl* ctl,
                                            uint32_t ci, const float4& u, const float4& v,
                                            const float4& w, int i, float F[3], float T[3]) {
    double xs[3];
    split_position(ctl->box, ctl->sc, ci, u, xs);
    const double vd[3] = {v.x, v.y, v.z}, wd[3] = {w.x, w.y, w.z};
    walls_apply<float>(W, pass, xs, (double)u.w, (double)v.w, vd, wd, W.spec[0].dt, i, F, T);
}

__global__ void k_split_to_x64(SplitPtrs P, double4* xr, const Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.n) split_to_x64_atom(i, P, xr, ctl);
}

// Split encoding of every atom from xr on grid G; u0 encodes x0 (the
// positions of the last build) in the same bins, relative to the build box.
__device__ __forceinline__ void x64_to_split_atom(int i, const SplitPtrs& P, const double4* xr,
                                                  const double4* x0, const Ctl* ctl,
                                                  const GridP& G, uint32_t* cell) {
    const DevBox& B = ctl->box;
    const double4 p = xr[i], q = x0[i];
    const double xs[3] = {p.x, p.y, p.z}, x0s[3] = {q.x, q.y, q.z};
    int c[3];
    float u[3], u0[3];
    for (int k = 0; k < 3; ++k) {
        const double lam = (xs[k] - B.lo[k]) / B.prd[k];
        int ck = (int)floor(lam / G.csize[k]);
        ck = max(0, min(ck, G.g[k] - 1));
        c[k] = ck;
        const double g = (double)G.g[k];
        u[k] = (float)(lam * g - ((double)ck + 0.5));
        const double lam0 = (x0s[k] - ctl->lo_build[k]) / ctl->prd_build[k];
        u0[k] = (float)(lam0 * g - ((double)ck + 0.5));
    }
    cell[i] = (uint32_t)c[0] | ((uint32_t)c[1] << 11) | ((uint32_t)c[2] << 22);
    float4 up = P.up[i];
    up.x = u[0]; up.y = u[1]; up.z = u[2];
    P.up[i] = up;
    P.u0[i] = make_float4(u0[0], u0[1], u0[2], 0.f);
}

// Gd: the grid of a build decided on the device (ctl->lgrid), else G
// x0_store: a build's x0, stored here from xr (the encoding of u0 then reads
// what it stores) instead of by a copy kernel of its own
__global__ void k_x64_to_split(SplitPtrs P, const double4* xr, const double4* x0, const Ctl* ctl,
                               GridP G, const GridP* Gd, uint32_t* cell, double4* x0_store) {
    if (Gd && ctl->error) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    if (x0_store) x0_store[i] = xr[i];
    x64_to_split_atom(i, P, xr, x0, ctl, Gd ? *Gd : G, cell);
}

// Per slot of the neighbor list: the cell of i minus the cell of j, wrapped
// to the nearest periodic image, 3 bits per axis (offset by 3). Between two
// builds the split cells do not change, so the one-kernel step reads this
// instead of gathering and unpacking the neighbour's cell. List partners sit
// within one cell of each other; a re-encode without a rebuild moves an atom
// by at most one more, so |difference| <= 3 always fits.
__device__ __forceinline__ void slot_codes_atom(int i, const int* nb_start, const int* nb_j,
                                                const uint32_t* cell, const Ctl* ctl,
                                                uint16_t* code) {
    int ci[3];
    unpack_cell(cell[i], ci[0], ci[1], ci[2]);
    const int per[3] = {ctl->box.per[0], ctl->box.per[1], ctl->box.per[2]};
    const int g[3] = {ctl->sc.gi[0], ctl->sc.gi[1], ctl->sc.gi[2]};
    for (int s = nb_start[i]; s < nb_start[i + 1]; ++s) {
        int cj[3];
        unpack_cell(cell[nb_j[s]], cj[0], cj[1], cj[2]);
        unsigned c = 0;
        for (int k = 0; k < 3; ++k) {
            int dc = ci[k] - cj[k];
            if (per[k] == 1) dc = wrap_cells(dc, g[k]);
            dc = max(-3, min(3, dc));
            c |= (unsigned)(dc + 3) << (3 * k);
        }
        code[s] = (uint16_t)c;
    }
}

__global__ void k_slot_codes(int n, const int* nb_start, const int* nb_j, const uint32_t* cell,
                             const Ctl* ctl, uint16_t* code) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) slot_codes_atom(i, nb_start, nb_j, cell, ctl, code);
}

// ------------------------------------------------------------ stage kernels

__global__ void k_clear_split(SplitPtrs P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    P.f[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.tq[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = 0.0;
}

__global__ void k_nve_initial_split(SplitPtrs P, const Ctl* ctl, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const SplitConst& sc = ctl->sc;
    const float dtf = 0.5f * dt;
    float4 u = P.up[i], v = P.vm[i], w = P.om[i];
    const float4 f = P.f[i], t = P.tq[i];
    const float im = 1.0f / v.w;
    v.x += dtf * f.x * im; v.y += dtf * f.y * im; v.z += dtf * f.z * im;
    u.x += dt * sc.inv_cw[0] * v.x; u.y += dt * sc.inv_cw[1] * v.y; u.z += dt * sc.inv_cw[2] * v.z;
    w.x += dtf * t.x * w.w; w.y += dtf * t.y * w.w; w.z += dtf * t.z * w.w;
    P.up[i] = u; P.vm[i] = v; P.om[i] = w;
}

__global__ void k_nve_final_split(SplitPtrs P, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const float dtf = 0.5f * dt;
    float4 v = P.vm[i], w = P.om[i];
    const float4 f = P.f[i], t = P.tq[i];
    const float im = 1.0f / v.w;
    v.x += dtf * f.x * im; v.y += dtf * f.y * im; v.z += dtf * f.z * im;
    w.x += dtf * t.x * w.w; w.y += dtf * t.y * w.w; w.z += dtf * t.z * w.w;
    P.vm[i] = v; P.om[i] = w;
}

__global__ void k_cundall_split(SplitPtrs P, float gl, float ga) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    float4 f = P.f[i], t = P.tq[i];
    const float4 v = P.vm[i], w = P.om[i];
    f.x *= cundall_scale_f(f.x * v.x, gl); t.x *= cundall_scale_f(t.x * w.x, ga);
    f.y *= cundall_scale_f(f.y * v.y, gl); t.y *= cundall_scale_f(t.y * w.y, ga);
    f.z *= cundall_scale_f(f.z * v.z, gl); t.z *= cundall_scale_f(t.z * w.z, ga);
    P.f[i] = f; P.tq[i] = t;
}

template <int STYLE>
__global__ void k_pair_split(SplitPtrs P, const Ctl* ctl, PairP pp, float dt, int shearupdate,
                             int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const SplitConst sc = ctl->sc;
    const int per[3] = {ctl->box.per[0], ctl->box.per[1], ctl->box.per[2]};
    float F[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f};
    double V[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (accumulate) {
        const float4 f = P.f[i], t = P.tq[i];
        F[0] = f.x; F[1] = f.y; F[2] = f.z;
        T[0] = t.x; T[1] = t.y; T[2] = t.z;
        for (int k = 0; k < 6; ++k) V[k] = P.vir[6 * i + k];
    }
    int nt = 0;
    const float4 ui = P.up[i], vi = P.vm[i], oi = P.om[i];
    gather_split<STYLE>(i, P.cell[i], ui, vi, oi, P, sc, per, pp, dt, shearupdate, F, T, V, nt);
    P.f[i] = make_float4(F[0], F[1], F[2], 0.f);
    P.tq[i] = make_float4(T[0], T[1], T[2], 0.f);
    for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = V[k];
}

// |x - x0|^2 in the split encoding: the box change since the build plus the
// atom's own drift.
__device__ __forceinline__ float split_disp_sq(const SplitConst& sc, uint32_t c, const float4& u,
                                               const float4& u0) {
    int cc[3];
    unpack_cell(c, cc[0], cc[1], cc[2]);
    const float uu[3] = {u.x, u.y, u.z}, u00[3] = {u0.x, u0.y, u0.z};
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float d = sc.dlo[k] + sc.dcw[k] * ((float)cc[k] + 0.5f + uu[k]) +
                        sc.cw_build[k] * (uu[k] - u00[k]);
        s += d * d;
    }
    return s;
}

__global__ void k_check_split(SplitPtrs P, Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    if (split_disp_sq(ctl->sc, P.cell[i], P.up[i], P.u0[i]) > (float)ctl->trigger_sq) ctl->moved = 1;
}

__global__ void k_leaves_split(SplitPtrs P, int keflag, double* partial) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < P.n) {
        for (int k = 0; k < 6; ++k) leaf[k] = P.vir[6 * i + k];
        if (keflag) {
            const float4 v4 = P.vm[i];
            const double m = v4.w, vx = v4.x, vy = v4.y, vz = v4.z;
            leaf[0] += m * vx * vx; leaf[1] += m * vy * vy; leaf[2] += m * vz * vz;
            leaf[3] += m * vx * vy; leaf[4] += m * vx * vz; leaf[5] += m * vy * vz;
        }
    }
    block_tree_sum<7>(leaf, partial + (size_t)blockIdx.x * 8);
}

__global__ void k_ke_leaves_split(SplitPtrs P, double* partial) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[2] = {0.0, 0.0};
    if (i < P.n) {
        const float4 v4 = P.vm[i], w4 = P.om[i];
        const double m = v4.w, r = P.up[i].w;
        const double vsq = (double)v4.x * v4.x + (double)v4.y * v4.y + (double)v4.z * v4.z;
        leaf[0] = 0.5 * m * vsq;
        const double wsq = (double)w4.x * w4.x + (double)w4.y * w4.y + (double)w4.z * w4.z;
        leaf[1] = 0.5 * (0.4 * m * r * r) * wsq;
    }
    block_tree_sum<2>(leaf, partial + (size_t)blockIdx.x * 8);
}

// ------------------------------------------------------------ fused steps

__device__ __forceinline__ void step_begin_split_atom(int i, const SplitPtrs& P, Ctl* ctl,
                                                      float dt, int kick) {
    const int t = ctl->timestep + 1;
    const SplitConst& sc = ctl->sc;
    float4 u = P.up[i];
    float4 v = P.vm[i];
    const float dtf = 0.5f * dt;
    if (kick) {
        const float4 f = P.f[i], tq = P.tq[i];
        const float im = 1.0f / v.w;
        v.x += dtf * f.x * im; v.y += dtf * f.y * im; v.z += dtf * f.z * im;
        float4 w = P.om[i];
        w.x += dtf * tq.x * w.w; w.y += dtf * tq.y * w.w; w.z += dtf * tq.z * w.w;
        P.om[i] = w;
        P.vm[i] = v;
    }
    u.x += dt * sc.inv_cw[0] * v.x;
    u.y += dt * sc.inv_cw[1] * v.y;
    u.z += dt * sc.inv_cw[2] * v.z;
    P.up[i] = u;
    const int since = t - ctl->last_build_step;
    if (since >= ctl->delay && since % ctl->every == 0) {
        if (!ctl->check) {
            if (i == 0) ctl->trigger = t;
        } else if (split_disp_sq(sc, P.cell[i], u, P.u0[i]) > (float)ctl->trigger_sq) {
            ctl->trigger = t;
        }
    }
}

__global__ void k_step_begin_split(SplitPtrs P, Ctl* ctl, float dt, int kick) {
    if (ctl->abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.n) step_begin_split_atom(i, P, ctl, dt, kick);
}

// The start of a redone step in one pass over the atoms: the drift and skin
// check of k_step_begin_split (no kick), the FP64 position decoded from the
// drifted encoding (k_split_to_x64) and wrapped into the box (k_pbc), each
// as its own kernel computes it.
__global__ void k_redo_begin_split(SplitPtrs P, Ctl* ctl, float dt, double4* xr, double4* vm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    if (!ctl->abort) step_begin_split_atom(i, P, ctl, dt, 0);
    split_to_x64_atom(i, P, xr, ctl);
    pbc_atom(i, xr, vm, ctl);
}

// Block `blk` of the force pass (its partial: partial[blk]).
template <int STYLE, int CUNDALL, int FINAL, int WALL = 0>
__device__ __forceinline__ void step_force_split_block(const SplitPtrs& P, Ctl* ctl,
                                                       const PairP& pp, float dt, float gl,
                                                       float ga, double* partial, int blk) {
    if (ctl->abort) return;
    const int t = ctl->timestep + 1;
    if (ctl->trigger == t) return;
    const int i = blk * TMPS_BLOCK + threadIdx.x;
    double leaf[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < P.n) {
        const SplitConst sc = ctl->sc;
        const int per[3] = {ctl->box.per[0], ctl->box.per[1], ctl->box.per[2]};
        const float4 ui = P.up[i];
        float4 v = P.vm[i];
        float4 w = P.om[i];
        float F[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f};
        double V[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int nt = 0;
        gather_split<STYLE>(i, P.cell[i], ui, v, w, P, sc, per, pp, dt, 1, F, T, V, nt);
        if (WALL) split_walls(P.W, 0, ctl, P.cell[i], ui, v, w, i, F, T);
        if (CUNDALL) {
            F[0] *= cundall_scale_f(F[0] * v.x, gl); T[0] *= cundall_scale_f(T[0] * w.x, ga);
            F[1] *= cundall_scale_f(F[1] * v.y, gl); T[1] *= cundall_scale_f(T[1] * w.y, ga);
            F[2] *= cundall_scale_f(F[2] * v.z, gl); T[2] *= cundall_scale_f(T[2] * w.z, ga);
        }
        if (WALL) split_walls(P.W, 1, ctl, P.cell[i], ui, v, w, i, F, T);
        const float dtf = 0.5f * dt;
        const float im = 1.0f / v.w;
        const float am[3] = {F[0] * im, F[1] * im, F[2] * im};
        const float ai[3] = {T[0] * w.w, T[1] * w.w, T[2] * w.w};
        v.x += dtf * am[0]; v.y += dtf * am[1]; v.z += dtf * am[2];
        w.x += dtf * ai[0]; w.y += dtf * ai[1]; w.z += dtf * ai[2];
        const double m = v.w, vx = v.x, vy = v.y, vz = v.z;
        leaf[0] = V[0] + m * vx * vx;
        leaf[1] = V[1] + m * vy * vy;
        leaf[2] = V[2] + m * vz * vz;
        leaf[3] = V[3] + m * vx * vy;
        leaf[4] = V[4] + m * vx * vz;
        leaf[5] = V[5] + m * vy * vz;
        leaf[6] = (double)nt;
        if (FINAL) {
            P.f[i] = make_float4(F[0], F[1], F[2], 0.f);
            P.tq[i] = make_float4(T[0], T[1], T[2], 0.f);
            if (FINAL == 1) {
                for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = V[k];
            } else {
                // FINAL 2, a rebuild step redone inside a run: the next
                // step's half-kick as k_kick_split computes it from the
                // stored force and torque (its expressions, so the bits
                // are the kick kernel's); no virial, which only a run's
                // last step leaves for the getters
                v.x += dtf * F[0] * im; v.y += dtf * F[1] * im; v.z += dtf * F[2] * im;
                w.x += dtf * T[0] * w.w; w.y += dtf * T[1] * w.w; w.z += dtf * T[2] * w.w;
            }
        } else {
            v.x += dtf * am[0]; v.y += dtf * am[1]; v.z += dtf * am[2];
            w.x += dtf * ai[0]; w.y += dtf * ai[1]; w.z += dtf * ai[2];
        }
        P.vm_out[i] = v;
        P.om_out[i] = w;
    }
    block_tree_sum<7>(leaf, partial + (size_t)blk * 8);
}

template <int STYLE, int CUNDALL, int FINAL, int WALL>
__global__ void k_step_force_split(SplitPtrs P, Ctl* ctl, PairP pp, float dt, float gl, float ga,
                                   double* partial) {
    step_force_split_block<STYLE, CUNDALL, FINAL, WALL>(P, ctl, pp, dt, gl, ga, partial,
                                                        blockIdx.x);
}

__global__ void k_copy_vw_split(float4* vdst, float4* wdst, const float4* vsrc, const float4* wsrc,
                                int n, const Ctl* ctl, int honor_abort) {
    if (honor_abort && ctl->abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vdst[i] = vsrc[i];
    wdst[i] = wsrc[i];
}

// ------------------------------------------------ one kernel per step

#ifndef SPLIT_BATCH
#define SPLIT_BATCH 4
#endif

// Separation i - j in length units from the slot's cell offset (i's cell
// minus j's, nearest image) and the two drifted offsets.
__device__ __forceinline__ void slot_separation(const SplitConst& sc, const bool fix[3],
                                                const float ginv[3], unsigned code,
                                                const float4& ui, const float4& uj, float d[3]) {
    float e[3] = {(float)((int)(code & 7u) - 3) + (ui.x - uj.x),
                  (float)((int)((code >> 3) & 7u) - 3) + (ui.y - uj.y),
                  (float)((int)((code >> 6) & 7u) - 3) + (ui.z - uj.z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (fix[k]) e[k] -= sc.g[k] * rintf(e[k] * ginv[k]);
        d[k] = e[k] * sc.cw[k];
    }
}

struct FusedS {
    const float4 *up_in, *vm_in, *om_in;
    float4 *up_out, *vm_out, *om_out;
    const uint8_t* fl_in;
    uint8_t* fl_out;
    const float4* h_in;
    float4* h_out;
    float4 *f, *tq;
    const float4* u0;
    const uint32_t* cell;
    double* vir;
    const int *nb_start, *nb_j;
    const uint16_t* nb_code;
    int n;
    WallRefs W;  // the run's walls (nwall 0 without)
};

// Where atom j is at step t in bin units: under remap x the offsets need no
// update, so the drift is all (dtc = dt / bin width of the current box).
__device__ __forceinline__ float4 step_offset(const float dtc[3], float4 u, const float4& v) {
    u.x = fmaf(dtc[0], v.x, u.x);
    u.y = fmaf(dtc[1], v.y, u.y);
    u.z = fmaf(dtc[2], v.z, u.z);
    return u;
}

// Step t in one pass (see k_step_f64) for the atoms of block `lb`: FP32
// contacts and integration, FP64 leaves; k_step_close finishes the tree and
// runs the servo. `fin`: the run's last step, which stores f / torque /
// virial instead of applying the next step's half-kick. Shared by the
// single-system kernel and the batched one (k_batch.hip).
template <int STYLE, int CUNDALL, int WALL = 0>
__device__ __forceinline__ void split_step_block(const FusedS& P, Ctl* ctl, const PairP& pp,
                                                 double* partial, float dt, float gl, float ga,
                                                 int lb, bool fin) {
    const int i = lb * TMPS_BLOCK + threadIdx.x;
    // The atom's own streams and row bounds are requested before the control
    // block is read, so the two memory round trips overlap.
    const int ia = i < P.n ? i : P.n - 1;
    float4 v = P.vm_in[ia];
    float4 w = P.om_in[ia];
    const uint32_t ci = P.cell[ia];
    const float4 up_i = P.up_in[ia];
    const float4 u0_i = P.u0[ia];
    const int s0 = P.nb_start[ia], s1 = P.nb_start[ia + 1];
    if (ctl->abort) return;
    const int t = ctl->timestep + 1;
    double leaf[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int trig = 0;
    if (i < P.n) {
        const SplitConst sc = ctl->sc;
        const int per[3] = {ctl->box.per[0], ctl->box.per[1], ctl->box.per[2]};
        const float dtc[3] = {dt * sc.inv_cw[0], dt * sc.inv_cw[1], dt * sc.inv_cw[2]};
        const float4 ui = step_offset(dtc, up_i, v);
        if (check_due(ctl, t)) {
            if (!ctl->check) trig = 1;
            else if (split_disp_sq(sc, ci, ui, u0_i) > (float)ctl->trigger_sq) trig = 1;
        }
        float F[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f};
        double V[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int nt = 0;
        const bool uses_history = (STYLE != TMPS_PAIR_GRANULAR) || pp.uses_history;
        const float zero3[3] = {0.f, 0.f, 0.f};
        // Per axis: offsets of list partners can be wider than half the box
        // only where the box holds fewer than 9 bins.
        bool fix[3];
        float ginv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            fix[k] = per[k] == 1 && sc.gi[k] < 9;
            ginv[k] = fix[k] ? 1.0f / sc.g[k] : 0.0f;
        }
        // Pass 1, every slot: separation (i minus j) and the contact test.
        // Indices and cell offsets of a batch, then all the partners' states,
        // are requested before any is used, so a row costs two memory round
        // trips, not two per slot. Touching slots are marked in `tmask` (the
        // row's first 64; a longer row keeps the rest in the touch flags).
        uint64_t tmask = 0;
        for (int sb = s0; sb < s1; sb += SPLIT_BATCH) {
            int jb[SPLIT_BATCH];
            uint16_t cb[SPLIT_BATCH];
#pragma unroll
            for (int q = 0; q < SPLIT_BATCH; ++q) {
                const bool in = sb + q < s1;
                jb[q] = in ? P.nb_j[sb + q] : i;
                cb[q] = in ? P.nb_code[sb + q] : (uint16_t)0xDB;
            }
            float4 vq[SPLIT_BATCH], uq[SPLIT_BATCH];
#pragma unroll
            for (int q = 0; q < SPLIT_BATCH; ++q) {
                vq[q] = P.vm_in[jb[q]];
                uq[q] = P.up_in[jb[q]];
            }
#pragma unroll
            for (int q = 0; q < SPLIT_BATCH; ++q) {
                if (sb + q < s1) {
                    const float4 uj = step_offset(dtc, uq[q], vq[q]);
                    float d[3];
                    slot_separation(sc, fix, ginv, cb[q], ui, uj, d);
                    const float rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    const float radsum = ui.w + uj.w;
                    const bool touch = rsq < radsum * radsum && rsq > 1e-28f;
                    const int k = sb + q - s0;
                    if (touch && k < 64) tmask |= 1ull << k;
                    P.fl_out[sb + q] = touch ? 1 : 0;
                }
            }
        }
        // Pass 2, touching slots in ascending order: the contact law for the
        // pair in canonical orientation (lower index first; the separation of
        // the other orientation is the exact negation), the partner's state
        // fetched again.
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
            const float4 vj = P.vm_in[j];
            const float4 uj = step_offset(dtc, P.up_in[j], vj);
            const bool first = i < j;
            float d[3];
            slot_separation(sc, fix, ginv, P.nb_code[s], ui, uj, d);
            if (!first) {
                d[0] = -d[0];
                d[1] = -d[1];
                d[2] = -d[2];
            }
            const float rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const float ra = first ? ui.w : uj.w, rb = first ? uj.w : ui.w;
            const float radsum = ra + rb;
            const float4 oj = P.om_in[j];
            const float4 va4 = first ? v : vj;
            const float4 vb4 = first ? vj : v;
            const float4 oa4 = first ? w : oj;
            const float4 ob4 = first ? oj : w;
            const float va[3] = {va4.x, va4.y, va4.z}, vb[3] = {vb4.x, vb4.y, vb4.z};
            const float oa[3] = {oa4.x, oa4.y, oa4.z}, ob[3] = {ob4.x, ob4.y, ob4.z};
            bool before = false;
            float h[3] = {0.f, 0.f, 0.f};
            if (uses_history) {
                before = P.fl_in[s] != 0;
                if (before) {
                    const float4 hh = P.h_in[s];
                    h[0] = hh.x; h[1] = hh.y; h[2] = hh.z;
                }
            }
            ContactOut<float> o;
            if (STYLE == TMPS_PAIR_HOOKE_HISTORY)
                legacy_history_contact<float, false>(
                    d, zero3, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa, ob, (float)pp.kn,
                    (float)pp.kt, (float)pp.gamman, (float)pp.gammat, (float)pp.xmu,
                    pp.limit_damping, dt, 1, before, h, o);
            else if (STYLE == TMPS_PAIR_HERTZ_HISTORY)
                legacy_history_contact<float, true>(
                    d, zero3, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa, ob, (float)pp.kn,
                    (float)pp.kt, (float)pp.gamman, (float)pp.gammat, (float)pp.xmu,
                    pp.limit_damping, dt, 1, before, h, o);
            else
                granular_contact<float>(d, zero3, rsq, radsum, ra, rb, va4.w, vb4.w, va, vb, oa,
                                        ob, pp, dt, 1, before, h, o);
            if (uses_history) P.h_out[s] = make_float4(h[0], h[1], h[2], 0.f);
            if (first) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    F[c] += o.F[c];
                    T[c] += o.ta[c];
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    F[c] -= o.F[c];
                    T[c] += o.tb[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) V[c] += (double)o.vir[c];
            ++nt;
        }
        if (WALL) split_walls(P.W, 0, ctl, ci, ui, v, w, i, F, T);
        if (CUNDALL) {
            F[0] *= cundall_scale_f(F[0] * v.x, gl); T[0] *= cundall_scale_f(T[0] * w.x, ga);
            F[1] *= cundall_scale_f(F[1] * v.y, gl); T[1] *= cundall_scale_f(T[1] * w.y, ga);
            F[2] *= cundall_scale_f(F[2] * v.z, gl); T[2] *= cundall_scale_f(T[2] * w.z, ga);
        }
        if (WALL) split_walls(P.W, 1, ctl, ci, ui, v, w, i, F, T);
        const float dtf = 0.5f * dt;
        const float im = 1.0f / v.w;
        const float am[3] = {F[0] * im, F[1] * im, F[2] * im};
        const float ai[3] = {T[0] * w.w, T[1] * w.w, T[2] * w.w};
        v.x += dtf * am[0]; v.y += dtf * am[1]; v.z += dtf * am[2];
        w.x += dtf * ai[0]; w.y += dtf * ai[1]; w.z += dtf * ai[2];
        const double m = v.w, vx = v.x, vy = v.y, vz = v.z;
        leaf[0] = V[0] + m * vx * vx;
        leaf[1] = V[1] + m * vy * vy;
        leaf[2] = V[2] + m * vz * vz;
        leaf[3] = V[3] + m * vx * vy;
        leaf[4] = V[4] + m * vx * vz;
        leaf[5] = V[5] + m * vy * vz;
        leaf[6] = (double)nt;
        if (fin) {
            P.f[i] = make_float4(F[0], F[1], F[2], 0.f);
            P.tq[i] = make_float4(T[0], T[1], T[2], 0.f);
            for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = V[k];
        } else {
            v.x += dtf * am[0]; v.y += dtf * am[1]; v.z += dtf * am[2];
            w.x += dtf * ai[0]; w.y += dtf * ai[1]; w.z += dtf * ai[2];
        }
        P.up_out[i] = ui;
        P.vm_out[i] = v;
        P.om_out[i] = w;
    }
    publish_block(leaf, trig, lb, partial);
}

template <int STYLE, int CUNDALL, int FINAL, int WALL>
__global__ void __launch_bounds__(TMPS_BLOCK) __attribute__((amdgpu_waves_per_eu(WALL ? TMPS_STEP_WAVES_SPLIT_WALL : TMPS_STEP_WAVES_SPLIT))) k_step_split(FusedS P, Ctl* ctl, PairP pp, double* partial,
                                                            float dt, float gl, float ga) {
    split_step_block<STYLE, CUNDALL, WALL>(P, ctl, pp, partial, dt, gl, ga,
                                           xcd_block(blockIdx.x, gridDim.x), FINAL != 0);
}

// ------------------------------------------------ ensemble (batch.hip)

// One ensemble member: its step's buffers for either input set, its control
// block, pair coefficients and tree partials.
struct BatchSplit {
    FusedS P[2];  // reading set 0 / set 1
    Ctl* ctl;
    PairP pp;
    double* partial;
    float4 *vm0, *om0, *f, *tq;  // set 0 and the stored forces (the half-kick)
    int n;
};

// Member b.x, its local block b.y: the step of a member that is running and
// not cut short by a rebuild request, on the set its previous step left. The
// per-atom work is split_step_block's, so a member's results are bitwise
// those of its single-system run.
template <int STYLE>
__global__ void __launch_bounds__(TMPS_BLOCK) __attribute__((amdgpu_waves_per_eu(TMPS_STEP_WAVES_SPLIT))) k_batch_step_split(const BatchSplit* table, const int2* blocks, float dt) {
    const int2 b = blocks[blockIdx.x];
    const BatchSplit& E = table[b.x];
    Ctl* ctl = E.ctl;
    if (ctl->bdone) return;
    const int par = ctl->bpar;
    const bool fin = ctl->timestep + 1 == ctl->bend;
    split_step_block<STYLE, 0>(E.P[par], ctl, E.pp, E.partial, dt, 0.f, 0.f, b.y, fin);
}

// The segment-start half-kick of members that need one (their run's start
// or a redone step), in set 0.
__global__ void k_batch_kick_split(const BatchSplit* table, const int2* blocks, float dt) {
    const int2 b = blocks[blockIdx.x];
    const BatchSplit& E = table[b.x];
    const Ctl* ctl = E.ctl;
    if (ctl->bdone || !ctl->bkick || ctl->abort) return;
    const int i = b.y * TMPS_BLOCK + threadIdx.x;
    if (i >= E.n) return;
    const float dtf = 0.5f * dt;
    float4 v = E.vm0[i], w = E.om0[i];
    const float4 ff = E.f[i], tq4 = E.tq[i];
    const float im = 1.0f / v.w;
    v.x += dtf * ff.x * im; v.y += dtf * ff.y * im; v.z += dtf * ff.z * im;
    w.x += dtf * tq4.x * w.w; w.y += dtf * tq4.y * w.w; w.z += dtf * tq4.z * w.w;
    E.vm0[i] = v;
    E.om0[i] = w;
}

// A finished member whose last step wrote set 1: its state back into set 0
// (as the single-system segment tail does).
__global__ void k_batch_tail_split(const BatchSplit* table, const int2* blocks) {
    const int2 b = blocks[blockIdx.x];
    const BatchSplit& E = table[b.x];
    const Ctl* ctl = E.ctl;
    if (!ctl->bdone || ctl->bpar != 1) return;
    const FusedS& Q = E.P[0];  // reads set 0, writes set 1
    const int i = b.y * TMPS_BLOCK + threadIdx.x;
    if (i < E.n) {
        const_cast<float4*>(Q.up_in)[i] = Q.up_out[i];
        const_cast<float4*>(Q.vm_in)[i] = Q.vm_out[i];
        const_cast<float4*>(Q.om_in)[i] = Q.om_out[i];
    }
    // the member's slots, strided over its blocks
    const int nb = (E.n + TMPS_BLOCK - 1) / TMPS_BLOCK;
    const int ns = Q.nb_start[E.n];
    for (int k = b.y * TMPS_BLOCK + threadIdx.x; k < ns; k += nb * TMPS_BLOCK) {
        const_cast<uint8_t*>(Q.fl_in)[k] = Q.fl_out[k];
        const_cast<float4*>(Q.h_in)[k] = Q.h_out[k];
    }
}

__global__ void k_kick_split(float4* vm, float4* om, const float4* f, const float4* tq, int n,
                             const Ctl* ctl, float dt) {
    if (ctl->abort) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dtf = 0.5f * dt;
    float4 v = vm[i], w = om[i];
    const float4 ff = f[i], tq4 = tq[i];
    const float im = 1.0f / v.w;
    v.x += dtf * ff.x * im; v.y += dtf * ff.y * im; v.z += dtf * ff.z * im;
    w.x += dtf * tq4.x * w.w; w.y += dtf * tq4.y * w.w; w.z += dtf * tq4.z * w.w;
    vm[i] = v;
    om[i] = w;
}

// ------------------------------------------------ batched redo (batch.hip)
//
// An ensemble member's redone rebuild step, many members per launch: each
// kernel runs the per-atom body of the stage kernel the single-system redo
// launches, over a block map of (entry, local block) pairs.

// The member's buffers with set in_buf read and set out_buf written.
__device__ __forceinline__ SplitPtrs redo_ptrs(const RedoEntry& e, int in_buf, int out_buf) {
    SplitPtrs p;
    p.up = e.up; p.f = e.ff; p.tq = e.tqf; p.u0 = e.u0;
    p.vm = in_buf ? e.vmf2 : e.vmf;
    p.om = in_buf ? e.omf2 : e.omf;
    p.vm_out = out_buf ? e.vmf2 : e.vmf;
    p.om_out = out_buf ? e.omf2 : e.omf;
    p.cell = e.cell; p.vir = e.vir;
    p.nb_start = e.nb_start; p.nb_j = e.nb_j; p.nb_flag = e.nb_flag;
    p.nb_hist = e.nb_hist;
    p.n = e.n;
    return p;
}

// the step's start in set 0 (no half-kick), then xr from the encoding
__global__ void kb_redo_begin_split(const RedoEntry* T, const int2* map, float dt) {
    const int2 mb = map[blockIdx.x];
    const RedoEntry& e = T[mb.x];
    const int i = mb.y * TMPS_BLOCK + threadIdx.x;
    if (i >= e.n) return;
    const SplitPtrs P = redo_ptrs(e, 0, 0);
    if (!e.ctl->abort) step_begin_split_atom(i, P, e.ctl, dt, 0);
    if (e.to_x64) split_to_x64_atom(i, P, e.xr, e.ctl);
}

__global__ void kb_redo_encode_split(const RedoEntry* T, const int2* map) {
    const int2 mb = map[blockIdx.x];
    const RedoEntry& e = T[mb.x];
    const int i = mb.y * TMPS_BLOCK + threadIdx.x;
    if (i < e.n) x64_to_split_atom(i, redo_ptrs(e, 0, 0), e.xr, e.x0, e.ctl, e.grid, e.cell);
}

__global__ void kb_redo_codes_split(const RedoEntry* T, const int2* map) {
    const int2 mb = map[blockIdx.x];
    const RedoEntry& e = T[mb.x];
    const int i = mb.y * TMPS_BLOCK + threadIdx.x;
    if (e.codes && i < e.n) slot_codes_atom(i, e.nb_start, e.nb_j, e.cell, e.ctl, e.nb_code);
}

// the force pass of the redone step: reads set 0, writes v / omega to set 1
template <int STYLE>
__global__ void kb_redo_force_split(const RedoEntry* T, const int2* map, float dt) {
    const int2 mb = map[blockIdx.x];
    const RedoEntry& e = T[mb.x];
    step_force_split_block<STYLE, 0, 1>(redo_ptrs(e, 0, 1), e.ctl, e.pair, dt, 0.f, 0.f,
                                        e.partial, mb.y);
}

__global__ void kb_redo_copy_vw_split(const RedoEntry* T, const int2* map) {
    const int2 mb = map[blockIdx.x];
    const RedoEntry& e = T[mb.x];
    if (e.ctl->abort) return;
    const int i = mb.y * TMPS_BLOCK + threadIdx.x;
    if (i >= e.n) return;
    e.vmf[i] = e.vmf2[i];
    e.omf[i] = e.omf2[i];
}

}  // namespace

// ================================================================ launchers

#define LAUNCH_S(kernel, n, ...)                                                           \
    do {                                                                                   \
        if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(nblocks_for(n)), dim3(TMPS_BLOCK), 0, \
                                        s.stream, __VA_ARGS__);                            \
        HIP_TRY(hipGetLastError());                                                        \
    } while (0)

int launch_split_to_x64(System& s) {
    LAUNCH_S(k_split_to_x64, s.n, sptrs(s), s.xr, s.ctl);
    return 0;
}

int launch_x64_to_split(System& s) {
    LAUNCH_S(k_x64_to_split, s.n, sptrs(s), s.xr, s.x0, s.ctl, s.grid, nullptr, s.cell,
             (double4*)nullptr);
    // the cells may have changed: refresh the list's per-slot cell offsets
    if (s.nslots > 0 && s.nb_code)
        LAUNCH_S(k_slot_codes, s.n, s.n, s.nb_start, s.nb_j, s.cell, s.ctl, s.nb_code);
    return 0;
}

// The split encoding after a list build decided on the device: its grid is
// the control block's, the build's x0 is stored on the way (launch_build_dev
// leaves it to this pass), and the slot offsets follow the rows the build
// wrote (an emptied list has none).
int launch_x64_to_split_dev(System& s) {
    LAUNCH_S(k_x64_to_split, s.n, sptrs(s), s.xr, s.x0, s.ctl, s.grid, &s.ctl->lgrid, s.cell,
             s.x0);
    if (s.nb_code) LAUNCH_S(k_slot_codes, s.n, s.n, s.nb_start, s.nb_j, s.cell, s.ctl, s.nb_code);
    return 0;
}

int launch_clear_split(System& s) {
    LAUNCH_S(k_clear_split, s.n, sptrs(s));
    return 0;
}

int launch_nve_initial_split(System& s, double dt) {
    LAUNCH_S(k_nve_initial_split, s.n, sptrs(s), s.ctl, (float)dt);
    return 0;
}

int launch_nve_final_split(System& s, double dt) {
    LAUNCH_S(k_nve_final_split, s.n, sptrs(s), (float)dt);
    return 0;
}

int launch_cundall_split(System& s, double gl, double ga) {
    LAUNCH_S(k_cundall_split, s.n, sptrs(s), (float)gl, (float)ga);
    return 0;
}

int launch_pair_split(System& s, double dt, int shearupdate, int accumulate) {
    switch (s.pair.style) {
        case TMPS_PAIR_HOOKE_HISTORY:
            LAUNCH_S(k_pair_split<TMPS_PAIR_HOOKE_HISTORY>, s.n, sptrs(s), s.ctl, s.pair, (float)dt,
                     shearupdate, accumulate);
            break;
        case TMPS_PAIR_HERTZ_HISTORY:
            LAUNCH_S(k_pair_split<TMPS_PAIR_HERTZ_HISTORY>, s.n, sptrs(s), s.ctl, s.pair, (float)dt,
                     shearupdate, accumulate);
            break;
        default:
            LAUNCH_S(k_pair_split<TMPS_PAIR_GRANULAR>, s.n, sptrs(s), s.ctl, s.pair, (float)dt,
                     shearupdate, accumulate);
            break;
    }
    return 0;
}

int launch_check_split(System& s) {
    LAUNCH_S(k_check_split, s.n, sptrs(s), s.ctl);
    return 0;
}

int launch_leaves_split(System& s, int keflag) {
    const int nb = nblocks_for(s.n);
    hipLaunchKernelGGL(k_leaves_split, dim3(nb > 0 ? nb : 1), dim3(TMPS_BLOCK), 0, s.stream,
                       sptrs(s), keflag, s.partial);
    HIP_TRY(hipGetLastError());
    return launch_final_tree(s, 6, nb > 0 ? nb : 1);
}

int launch_ke_split(System& s, double* out2) {
    (void)out2;
    const int nb = nblocks_for(s.n);
    hipLaunchKernelGGL(k_ke_leaves_split, dim3(nb > 0 ? nb : 1), dim3(TMPS_BLOCK), 0, s.stream,
                       sptrs(s), s.partial);
    HIP_TRY(hipGetLastError());
    return launch_final_tree(s, 2, nb > 0 ? nb : 1);
}

__global__ void k_split_consts(Ctl* ctl, int set_build) {
    if (threadIdx.x == 0 && blockIdx.x == 0) update_split_consts(ctl, set_build);
}

int launch_split_consts(System& s, int set_build) {
    hipLaunchKernelGGL(k_split_consts, dim3(1), dim3(64), 0, s.stream, s.ctl, set_build);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_step_begin_split(System& s, double dt, int kick, int buf) {
    LAUNCH_S(k_step_begin_split, s.n, sptrs(s, buf, buf), s.ctl, (float)dt, kick);
    return 0;
}

int launch_redo_begin_split(System& s, double dt) {
    LAUNCH_S(k_redo_begin_split, s.n, sptrs(s, 0, 0), s.ctl, (float)dt, s.xr, s.vm);
    return 0;
}

int launch_copy_vw_split(System& s, int honor_abort) {
    LAUNCH_S(k_copy_vw_split, s.n, s.vmf, s.omf, s.vmf2, s.omf2, s.n, s.ctl, honor_abort);
    return 0;
}

template <int STYLE, int WALL>
static int launch_force_split_walls(System& s, const SplitPtrs& P, double dt, int final_mode,
                                    int cundall, double gl, double ga) {
    const dim3 g(nblocks_for(s.n)), b(TMPS_BLOCK);
    const float fdt = (float)dt, fgl = (float)gl, fga = (float)ga;
    if (final_mode == 2) {
        if (cundall)
            hipLaunchKernelGGL((k_step_force_split<STYLE, 1, 2, WALL>), g, b, 0, s.stream, P,
                               s.ctl, s.pair, fdt, fgl, fga, s.partial);
        else
            hipLaunchKernelGGL((k_step_force_split<STYLE, 0, 2, WALL>), g, b, 0, s.stream, P,
                               s.ctl, s.pair, fdt, fgl, fga, s.partial);
    } else if (final_mode) {
        if (cundall)
            hipLaunchKernelGGL((k_step_force_split<STYLE, 1, 1, WALL>), g, b, 0, s.stream, P,
                               s.ctl, s.pair, fdt, fgl, fga, s.partial);
        else
            hipLaunchKernelGGL((k_step_force_split<STYLE, 0, 1, WALL>), g, b, 0, s.stream, P,
                               s.ctl, s.pair, fdt, fgl, fga, s.partial);
    } else {
        if (cundall)
            hipLaunchKernelGGL((k_step_force_split<STYLE, 1, 0, WALL>), g, b, 0, s.stream, P,
                               s.ctl, s.pair, fdt, fgl, fga, s.partial);
        else
            hipLaunchKernelGGL((k_step_force_split<STYLE, 0, 0, WALL>), g, b, 0, s.stream, P,
                               s.ctl, s.pair, fdt, fgl, fga, s.partial);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// (a fused run's walls, when it has any, in place on set 0: the redo of a
// step runs its force pass there)
template <int STYLE>
static int launch_force_split_style(System& s, double dt, int final_mode, int cundall, double gl,
                                    double ga, int in_buf, int out_buf) {
    SplitPtrs P = sptrs(s, in_buf, out_buf);
    if (s.nrun_walls > 0) {
        P.W = run_walls(s, 0, 0);
        return launch_force_split_walls<STYLE, 1>(s, P, dt, final_mode, cundall, gl, ga);
    }
    return launch_force_split_walls<STYLE, 0>(s, P, dt, final_mode, cundall, gl, ga);
}

int launch_step_force_split(System& s, double dt, int final_mode, int cundall, double gl,
                            double ga, int in_buf, int out_buf) {
    if (s.n <= 0) return 0;
    if (in_buf == out_buf) return tmps_fail(TMPS_ERR_ARG, "force pass cannot update v in place");
    switch (s.pair.style) {
        case TMPS_PAIR_HOOKE_HISTORY:
            return launch_force_split_style<TMPS_PAIR_HOOKE_HISTORY>(s, dt, final_mode, cundall, gl,
                                                                     ga, in_buf, out_buf);
        case TMPS_PAIR_HERTZ_HISTORY:
            return launch_force_split_style<TMPS_PAIR_HERTZ_HISTORY>(s, dt, final_mode, cundall, gl,
                                                                     ga, in_buf, out_buf);
        default:
            return launch_force_split_style<TMPS_PAIR_GRANULAR>(s, dt, final_mode, cundall, gl, ga,
                                                                in_buf, out_buf);
    }
}

// ------------------------------------------------ one kernel per step

template <int STYLE, int CUNDALL, int FINAL, int WALL>
static int fused_split_go(System& s, const FusedS& P, double dt, double gl, double ga) {
    const int nb = nblocks_for(s.n);
    if (s.kstart || s.kstop) {
        hipExtLaunchKernelGGL((k_step_split<STYLE, CUNDALL, FINAL, WALL>), dim3(nb),
                              dim3(TMPS_BLOCK), 0, s.stream, s.kstart, s.kstop, 0, P, s.ctl,
                              s.pair, s.partial, (float)dt, (float)gl, (float)ga);
        s.kstart = s.kstop = nullptr;
    } else {
        hipLaunchKernelGGL((k_step_split<STYLE, CUNDALL, FINAL, WALL>), dim3(nb),
                           dim3(TMPS_BLOCK), 0, s.stream, P, s.ctl, s.pair, s.partial, (float)dt,
                           (float)gl, (float)ga);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int STYLE, int WALL>
static int fused_split_walls(System& s, const FusedS& P, double dt, int fin, int cundall,
                             double gl, double ga) {
    if (fin) {
        if (cundall) return fused_split_go<STYLE, 1, 1, WALL>(s, P, dt, gl, ga);
        return fused_split_go<STYLE, 0, 1, WALL>(s, P, dt, gl, ga);
    }
    if (cundall) return fused_split_go<STYLE, 1, 0, WALL>(s, P, dt, gl, ga);
    return fused_split_go<STYLE, 0, 0, WALL>(s, P, dt, gl, ga);
}

template <int STYLE>
static int fused_split_style(System& s, const FusedS& P, double dt, int fin, int cundall,
                             double gl, double ga) {
    if (P.W.nwall > 0) return fused_split_walls<STYLE, 1>(s, P, dt, fin, cundall, gl, ga);
    return fused_split_walls<STYLE, 0>(s, P, dt, fin, cundall, gl, ga);
}

int launch_fused_split(System& s, int in_buf, double dt, int fin, int cundall, double gl,
                       double ga) {
    if (s.n <= 0) return 0;
    const int ob = 1 - in_buf;
    FusedS P;
    P.up_in = in_buf ? s.up2 : s.up;
    P.vm_in = in_buf ? s.vmf2 : s.vmf;
    P.om_in = in_buf ? s.omf2 : s.omf;
    P.up_out = ob ? s.up2 : s.up;
    P.vm_out = ob ? s.vmf2 : s.vmf;
    P.om_out = ob ? s.omf2 : s.omf;
    P.fl_in = in_buf ? s.nb_flag2 : s.nb_flag;
    P.fl_out = ob ? s.nb_flag2 : s.nb_flag;
    P.h_in = static_cast<const float4*>(in_buf ? s.nb_hist2 : s.nb_hist);
    P.h_out = static_cast<float4*>(ob ? s.nb_hist2 : s.nb_hist);
    P.f = s.ff;
    P.tq = s.tqf;
    P.u0 = s.u0;
    P.cell = s.cell;
    P.vir = s.vir;
    P.nb_start = s.nb_start;
    P.nb_j = s.nb_j;
    P.nb_code = s.nb_code;
    P.n = s.n;
    P.W = run_walls(s, in_buf, ob);
    switch (s.pair.style) {
        case TMPS_PAIR_HOOKE_HISTORY:
            return fused_split_style<TMPS_PAIR_HOOKE_HISTORY>(s, P, dt, fin, cundall, gl, ga);
        case TMPS_PAIR_HERTZ_HISTORY:
            return fused_split_style<TMPS_PAIR_HERTZ_HISTORY>(s, P, dt, fin, cundall, gl, ga);
        default:
            return fused_split_style<TMPS_PAIR_GRANULAR>(s, P, dt, fin, cundall, gl, ga);
    }
}

int launch_kick_split(System& s, double dt) {
    LAUNCH_S(k_kick_split, s.n, s.vmf, s.omf, s.ff, s.tqf, s.n, s.ctl, (float)dt);
    return 0;
}

// ------------------------------------------------ ensemble (batch.hip)

size_t batch_split_entry_size() { return sizeof(BatchSplit); }

static FusedS fused_split_ptrs(System& s, int in_buf) {
    const int ob = 1 - in_buf;
    FusedS P;
    P.up_in = in_buf ? s.up2 : s.up;
    P.vm_in = in_buf ? s.vmf2 : s.vmf;
    P.om_in = in_buf ? s.omf2 : s.omf;
    P.up_out = ob ? s.up2 : s.up;
    P.vm_out = ob ? s.vmf2 : s.vmf;
    P.om_out = ob ? s.omf2 : s.omf;
    P.fl_in = in_buf ? s.nb_flag2 : s.nb_flag;
    P.fl_out = ob ? s.nb_flag2 : s.nb_flag;
    P.h_in = static_cast<const float4*>(in_buf ? s.nb_hist2 : s.nb_hist);
    P.h_out = static_cast<float4*>(ob ? s.nb_hist2 : s.nb_hist);
    P.f = s.ff;
    P.tq = s.tqf;
    P.u0 = s.u0;
    P.cell = s.cell;
    P.vir = s.vir;
    P.nb_start = s.nb_start;
    P.nb_j = s.nb_j;
    P.nb_code = s.nb_code;
    P.n = s.n;
    P.W = WallRefs{};  // ensemble members have no walls
    return P;
}

int batch_split_entry(System& s, void* host_entry) {
    if (s.precision != TMPS_FP32_SPLIT)
        return tmps_fail(TMPS_ERR_UNSUPPORTED, "an ensemble runs FP32-split systems");
    BatchSplit E;
    E.P[0] = fused_split_ptrs(s, 0);
    E.P[1] = fused_split_ptrs(s, 1);
    E.ctl = s.ctl;
    E.pp = s.pair;
    E.partial = s.partial;
    E.vm0 = s.vmf;
    E.om0 = s.omf;
    E.f = s.ff;
    E.tq = s.tqf;
    E.n = s.n;
    std::memcpy(host_entry, &E, sizeof E);
    return 0;
}

int launch_batch_split_kick(const void* table, const int2* blocks, int nblocks, float dt,
                            hipStream_t st) {
    if (nblocks <= 0) return 0;
    hipLaunchKernelGGL(k_batch_kick_split, dim3(nblocks), dim3(TMPS_BLOCK), 0, st,
                       static_cast<const BatchSplit*>(table), blocks, dt);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_batch_split_step(int style, const void* table, const int2* blocks, int nblocks,
                            float dt, hipStream_t st) {
    if (nblocks <= 0) return 0;
    const BatchSplit* T = static_cast<const BatchSplit*>(table);
    switch (style) {
        case TMPS_PAIR_HOOKE_HISTORY:
            hipLaunchKernelGGL(k_batch_step_split<TMPS_PAIR_HOOKE_HISTORY>, dim3(nblocks),
                               dim3(TMPS_BLOCK), 0, st, T, blocks, dt);
            break;
        case TMPS_PAIR_HERTZ_HISTORY:
            hipLaunchKernelGGL(k_batch_step_split<TMPS_PAIR_HERTZ_HISTORY>, dim3(nblocks),
                               dim3(TMPS_BLOCK), 0, st, T, blocks, dt);
            break;
        default:
            hipLaunchKernelGGL(k_batch_step_split<TMPS_PAIR_GRANULAR>, dim3(nblocks),
                               dim3(TMPS_BLOCK), 0, st, T, blocks, dt);
            break;
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_batch_split_tail(const void* table, const int2* blocks, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return 0;
    hipLaunchKernelGGL(k_batch_tail_split, dim3(nblocks), dim3(TMPS_BLOCK), 0, st,
                       static_cast<const BatchSplit*>(table), blocks);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ------------------------------------------------ batched redo (batch.hip)

int launch_redo_begin_split(const RedoEntry* T, const RedoMaps& M, float dt, hipStream_t st) {
    if (M.n_atoms <= 0) return 0;
    hipLaunchKernelGGL(kb_redo_begin_split, dim3(M.n_atoms), dim3(TMPS_BLOCK), 0, st, T, M.atoms,
                       dt);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_redo_encode_split(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    if (M.n_atoms <= 0) return 0;
    hipLaunchKernelGGL(kb_redo_encode_split, dim3(M.n_atoms), dim3(TMPS_BLOCK), 0, st, T, M.atoms);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(kb_redo_codes_split, dim3(M.n_atoms), dim3(TMPS_BLOCK), 0, st, T, M.atoms);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_redo_force_split(const RedoEntry* T, const RedoMaps& M, float dt, hipStream_t st) {
    if (M.n_style[TMPS_PAIR_HOOKE_HISTORY] > 0)
        hipLaunchKernelGGL(kb_redo_force_split<TMPS_PAIR_HOOKE_HISTORY>,
                           dim3(M.n_style[TMPS_PAIR_HOOKE_HISTORY]), dim3(TMPS_BLOCK), 0, st, T,
                           M.style[TMPS_PAIR_HOOKE_HISTORY], dt);
    HIP_TRY(hipGetLastError());
    if (M.n_style[TMPS_PAIR_HERTZ_HISTORY] > 0)
        hipLaunchKernelGGL(kb_redo_force_split<TMPS_PAIR_HERTZ_HISTORY>,
                           dim3(M.n_style[TMPS_PAIR_HERTZ_HISTORY]), dim3(TMPS_BLOCK), 0, st, T,
                           M.style[TMPS_PAIR_HERTZ_HISTORY], dt);
    HIP_TRY(hipGetLastError());
    if (M.n_style[TMPS_PAIR_GRANULAR] > 0)
        hipLaunchKernelGGL(kb_redo_force_split<TMPS_PAIR_GRANULAR>,
                           dim3(M.n_style[TMPS_PAIR_GRANULAR]), dim3(TMPS_BLOCK), 0, st, T,
                           M.style[TMPS_PAIR_GRANULAR], dt);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_redo_copy_vw_split(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    if (M.n_atoms <= 0) return 0;
    hipLaunchKernelGGL(kb_redo_copy_vw_split, dim3(M.n_atoms), dim3(TMPS_BLOCK), 0, st, T,
                       M.atoms);
    HIP_TRY(hipGetLastError());
    return 0;
}
