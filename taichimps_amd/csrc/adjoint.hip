// Hand-written adjoint of the FP64 DEM step (contact + velocity Verlet +
// deform/pressure servo with remap x), with segment checkpointing.
//
// Forward (tmps_adjoint_forward): Verlet::setup (the setup force), then the
// device-resident runner in segments of `every` steps; the full state is
// checkpointed before each segment. Backward (tmps_adjoint_backward): for each
// segment, last first, the checkpoint is restored and the segment replayed
// through the stage path (bitwise the same trajectory), recording every
// step's inputs; then the steps are reversed:
//
//   end of step t : X_t = lo' + (xh - lo) prd'/prd       (remap x)
//                   cum' = cum + dt rate(P_t) prd          (servo, pressure styles)
//                   P_t = (sum vir + sum m V' V'^T) / vol
//                   V' = vh + dt/2 F_t/m, W' = wh + dt/2 T_t/I
//   force         : F_t, T_t, S_t = contact(xh, vh, wh, S_{t-1}, box)
//   start of step : vh = V + dt/2 F_{t-1}/m, xh = X + dt vh, wh = W + dt/2 T_{t-1}/I
//                   (+ at a rebuild: PBC wrap of xh, history carried to the new list)
//
// Adjoint state between steps: per atom (X, V, W, F, T), per slot S, and the
// servo's cumulative strain cum (the box is lo_start - cum/2, hi_start + cum/2
// on pressure axes). Per-atom mass and radius adjoints accumulate over all
// steps; the radius gradient includes the mass at fixed density
// (m = 4/3 pi rho r^3) and the moment of inertia 0.4 m r^2.
//
// History: a pair's shear is stored in both of its slots (the full list) and
// evaluated once in canonical orientation; its adjoint is carried by the
// lower-index atom's slot only, so the two copies never double count.
//
// The contact set, neighbor lists, rebuild steps and Coulomb-slip branch
// choices are those of the forward pass (piecewise constant in the inputs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "adjoint.h"
#include "tmps_internal.h"

using namespace tmps;

namespace {

#define ADJ_THREADS 256

// The adjoint's scalar state on the device.
struct AdjCtl {
    double cumb[3];   // adjoint of the servo's cumulative strain
    double lob[3];    // pending adjoint of lo of the current step's box
    double prdb[3];   // pending adjoint of prd of the current step's box
    double vsb[6];    // adjoint of the reduced virial(+kinetic) sum of the step
    double grad[5];   // kn, kt, gamman, gammat, xmu
};

// What a step saw of the box and the servo.
struct BoxRec {
    double lo[3], prd[3];      // box of the step (forces, remap source)
    double lo_n[3], prd_n[3];  // box after the step's servo update (remap target)
    double tensor[6];          // pressure tensor of the servo update
    int per[3], fast[3];
};

struct ListVer {
    const int* start = nullptr;
    const int* j = nullptr;
    int nslots = 0;
    const int* src = nullptr;   // per new slot: the old slot its history came from, or -1
    const int* wrap = nullptr;  // per atom, 4 ints: PBC image shift of the rebuild
};

struct StepRec {
    int list = 0;     // version of the list the step's forces used
    int rebuilt = 0;  // the step rebuilt its list (list > 0 then)
    size_t slot_off = 0;
    BoxRec box;
};

// A bump allocator over device chunks, reset per segment.
struct Arena {
    struct Chunk {
        char* p = nullptr;
        size_t cap = 0, used = 0;
    };
    std::vector<Chunk> chunks;
    int cur = 0;
    size_t min_chunk = (size_t)256 << 20;
    void reset() {
        for (auto& c : chunks) c.used = 0;
        cur = 0;
    }
    void* get(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        while (cur < (int)chunks.size()) {
            Chunk& c = chunks[cur];
            if (c.used + bytes <= c.cap) {
                void* r = c.p + c.used;
                c.used += bytes;
                return r;
            }
            ++cur;
        }
        Chunk c;
        c.cap = std::max(bytes, min_chunk);
        if (hipMalloc(&c.p, c.cap) != hipSuccess) return nullptr;
        c.used = bytes;
        chunks.push_back(c);
        cur = (int)chunks.size() - 1;
        return c.p;
    }
    void release() {
        for (auto& c : chunks)
            if (c.p) (void)hipFree(c.p);
        chunks.clear();
        cur = 0;
    }
};

struct Checkpoint {
    Arena mem{{}, 0, (size_t)16 << 20};
    double4 *xr = nullptr, *vm = nullptr, *om = nullptr, *f = nullptr, *tq = nullptr,
            *x0 = nullptr;
    int *nb_start = nullptr, *nb_j = nullptr;
    uint8_t* nb_flag = nullptr;
    double4* nb_hist = nullptr;
    int nslots = 0;
    Ctl ctl{};
    int timestep = 0, current_step = 0, last_build_step = 0, has_built = 0, build_count = 0;
    int npairs = 0, setup_done = 0, servo_step_count = 0, servo_nsteps = 0;
    GridP grid{}, ref_grid{};
};

}  // namespace

struct AdjContext {
    int nsteps = 0, every = 0, done_forward = 0;
    double dt = 0.0;
    std::vector<Checkpoint*> cps;  // before each segment, then the end state
    // the setup force's inputs: list of the setup, touching flags and stored
    // shear before it
    uint8_t* setup_flb = nullptr;
    double4* setup_sold = nullptr;
    Arena setup_mem{{}, 0, (size_t)16 << 20};
    // segment recording
    Arena rec;
    std::vector<StepRec> steps;
    std::vector<ListVer> lists;
    double4 *XH = nullptr, *VH = nullptr, *WH = nullptr, *FT = nullptr, *TT = nullptr;
    int rec_steps_cap = 0;
    std::vector<uint8_t*> flb, fla;
    std::vector<double4*> sold;
    double4* tmpx = nullptr;
    // the steps at which the forward pass built its list: a replay, which
    // repeats the forward run bit for bit, builds at exactly these
    std::vector<int> build_steps;
    // each replayed step's control block after its servo (device), read back
    // once per segment
    Ctl* ctl_rec = nullptr;
    std::vector<Ctl> ctl_host;
    // adjoint state
    double4 *Xb = nullptr, *Vb = nullptr, *Wb = nullptr, *Fb = nullptr, *Tb = nullptr;
    double4 *xhb = nullptr, *vhb = nullptr, *whb = nullptr, *Fcb = nullptr, *Tcb = nullptr;
    double *mbar = nullptr, *rbar = nullptr;
    double4 *Sb = nullptr, *Sbo = nullptr;
    size_t sb_cap = 0;
    AdjCtl* actl = nullptr;
    double* part = nullptr;  // [2][blocks][8]
    int n_alloc = 0;
};

namespace {

template <typename T>
int dmalloc(T** p, size_t count) {
    *p = nullptr;
    HIP_TRY(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
    return 0;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <typename T>
T* arena_copy(Arena& a, const T* src, size_t count, hipStream_t st, int* rc) {
    T* d = static_cast<T*>(a.get(std::max<size_t>(count, 1) * sizeof(T)));
    if (!d) {
        *rc = tmps_fail(TMPS_ERR_HIP, "adjoint: out of device memory");
        return nullptr;
    }
    if (count > 0 &&
        hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyDeviceToDevice, st) != hipSuccess)
        *rc = tmps_fail(TMPS_ERR_HIP, "adjoint: device copy failed");
    return d;
}

// ------------------------------------------------------------------ kernels

// per new slot: the old slot whose history the restore copied, or -1
// (the merge of k_hist_restore, contact_history.py:91-139)
__global__ void k_adj_remap_src(int n, const int* nb_start, const int* nb_j, const int* onb_start,
                                const int* onb_j, const uint8_t* onb_flag, int old_n, int* src) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = 0, o1 = 0;
    if (i < old_n) {
        o = onb_start[i];
        o1 = onb_start[i + 1];
    }
    for (int s = nb_start[i]; s < nb_start[i + 1]; ++s) {
        const int j = nb_j[s];
        while (o < o1 && onb_j[o] < j) ++o;
        src[s] = (o < o1 && onb_j[o] == j && onb_flag[o]) ? o : -1;
    }
}

// image shift of the rebuild's PBC wrap: after = before + k prd
__global__ void k_adj_wrap(int n, const double4* before, const double4* after, BoxRec B, int* wrap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 a = before[i], b = after[i];
    const double d[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
    for (int k = 0; k < 3; ++k) wrap[4 * i + k] = B.per[k] == 1 ? (int)llround(d[k] / B.prd[k]) : 0;
    wrap[4 * i + 3] = 0;
}

// pairwise block sums of NC leaves into part[block][0..NC)
template <int NC>
__device__ __forceinline__ void leaves_out(double leaf[NC], double* part) {
    double l8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < NC; ++k) l8[k] = leaf[k];
    block_tree_sum<8>(l8, part + (size_t)blockIdx.x * 8);
}

// sums part[0..nparts)[k] in one block (deterministic order); thread 0 gets it
__device__ void block_sum8(const double* part, int nparts, double out[8]) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += ADJ_THREADS)
        for (int k = 0; k < 8; ++k) acc[k] += part[(size_t)b * 8 + k];
    double tmp[8];
    __shared__ double res[8];
    __shared__ double ws[ADJ_THREADS / 64][8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 8; ++k) {
        tmp[k] = wave_tree_sum(acc[k]);
        if (lane == 0) ws[wave][k] = tmp[k];
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) res[k] = (ws[0][k] + ws[1][k]) + (ws[2][k] + ws[3][k]);
    __syncthreads();
    for (int k = 0; k < 8; ++k) out[k] = res[k];
}

// remap x of the step's end: xh_bar = X_bar s; leaves sum X_bar and
// X_bar (xh - lo) per axis
__global__ void k_adj_remap(int n, const double4* Xb, const double4* XH, BoxRec B, double4* xhb,
                            double* part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double4 xb = Xb[i], xh = XH[i];
        const double xbv[3] = {xb.x, xb.y, xb.z}, xhv[3] = {xh.x, xh.y, xh.z};
        double o[3];
        for (int k = 0; k < 3; ++k) {
            const double s = B.prd_n[k] / B.prd[k];
            o[k] = xbv[k] * s;
            leaf[k] = xbv[k];
            leaf[3 + k] = xbv[k] * (xhv[k] - B.lo[k]);
        }
        xhb[i] = make_double4(o[0], o[1], o[2], 0.0);
    }
    leaves_out<6>(leaf, part);
}

// rate = limit_rate(pgain (p - ptarget)) and its derivative in p
__device__ double limit_rate_d(const tmps_servo_params& S, double raw, double ptarget,
                               double* drate_draw) {
    double out = raw, d = 1.0;
    if (S.normalize == 1) {
        if (ptarget == 0.0) {
            d = 0.0;
            if (out != 0.0) out = S.max_rate * (out > 0.0 ? 1.0 : -1.0);
        } else {
            out = out / fabs(ptarget);
            d = 1.0 / fabs(ptarget);
        }
    }
    if (S.max_rate != 0.0 && fabs(out) > S.max_rate) {
        out = S.max_rate * (out > 0.0 ? 1.0 : -1.0);
        d = 0.0;
    }
    *drate_draw = d;
    return out;
}

// The servo and remap adjoint of one step (one block): from the remap sums
// and the incoming cum adjoint to the adjoint of the reduced virial(+kinetic)
// sum, and the pending adjoints of the step's own box.
__global__ void __launch_bounds__(ADJ_THREADS) k_adj_servo(AdjCtl* A, const double* part,
                                                           int nparts, BoxRec B,
                                                           tmps_servo_params S, double dt,
                                                           double dl0, double dl1, double dl2,
                                                           double dl3, double dl4, double dl5) {
    double sums[8];
    block_sum8(part, nparts, sums);
    if (threadIdx.x != 0) return;
    const double dL[6] = {dl0, dl1, dl2, dl3, dl4, dl5};
    double lob[3], prdb[3], cumb[3];
    for (int k = 0; k < 3; ++k) {
        const double s = B.prd_n[k] / B.prd[k];
        const double lob_n = sums[k];
        const double prdb_n = sums[3 + k] / B.prd[k];
        lob[k] = -sums[k] * s;
        prdb[k] = -sums[3 + k] * s / B.prd[k];
        cumb[k] = A->cumb[k];
        if (S.style[k] == 8 || S.style[k] == 9) {
            cumb[k] += -0.5 * lob_n + prdb_n;
        } else {
            // the axis keeps its box: the new box is the old one
            lob[k] += lob_n;
            prdb[k] += prdb_n;
        }
    }
    // p_i from the tensor, as servo_update (deform_pressure.py:1113-1219)
    const double* t = B.tensor;
    const double scalar = (t[0] + t[1] + t[2]) / 3.0;
    double p[3];
    for (int i = 0; i < 3; ++i) p[i] = t[i];
    int a = 0, b = 1;
    if (S.couple == 1) {
        p[0] = p[1] = p[2] = scalar;
    } else if (S.couple >= 2) {
        if (S.couple == 3) { a = 1; b = 2; }
        if (S.couple == 4) { a = 0; b = 2; }
        p[a] = p[b] = (t[a] + t[b]) / 2.0;
    } else {
        for (int i = 0; i < 3; ++i)
            if (S.style[i] == 9) p[i] = scalar;
    }
    double tb[6];
    for (int k = 0; k < 6; ++k) tb[k] = dL[k];
    for (int i = 0; i < 3; ++i) {
        if (!(S.style[i] == 8 || S.style[i] == 9)) continue;
        double dr = 0.0;
        const double rate = limit_rate_d(S, S.pgain[i] * (p[i] - S.ptarget[i]), S.ptarget[i], &dr);
        // cum' = cum + dt rate prd
        const double rateb = cumb[i] * dt * B.prd[i];
        prdb[i] += cumb[i] * dt * rate;
        const double pb = rateb * dr * S.pgain[i];
        if (S.couple == 1 || (S.couple == 0 && S.style[i] == 9)) {
            for (int k = 0; k < 3; ++k) tb[k] += pb / 3.0;
        } else if (S.couple >= 2 && (i == a || i == b)) {
            tb[a] += pb / 2.0;
            tb[b] += pb / 2.0;
        } else {
            tb[i] += pb;
        }
    }
    // t = vs / vol
    const double vol = B.prd[0] * B.prd[1] * B.prd[2];
    double volb = 0.0;
    for (int k = 0; k < 6; ++k) {
        A->vsb[k] = tb[k] / vol;
        volb += -tb[k] * t[k] / vol;
    }
    prdb[0] += volb * B.prd[1] * B.prd[2];
    prdb[1] += volb * B.prd[0] * B.prd[2];
    prdb[2] += volb * B.prd[0] * B.prd[1];
    for (int k = 0; k < 3; ++k) {
        A->cumb[k] = cumb[k];
        A->lob[k] = lob[k];
        A->prdb[k] = prdb[k];
    }
}

// final_integrate and the kinetic part of the pressure, reversed
__global__ void k_adj_final(int n, const double4* XH, const double4* VH, const double4* FT,
                            const double4* TT, const double4* Vb, const double4* Wb,
                            const double4* Fb, const double4* Tb, const AdjCtl* A, double dt,
                            double4* vhb, double4* whb, double4* Fcb, double4* Tcb, double* mbar,
                            double* rbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dtf = 0.5 * dt;
    const double4 vh = VH[i], F = FT[i], T = TT[i];
    const double m = vh.w, r = XH[i].w;
    const double inertia = 0.4 * m * r * r;
    const double vp[3] = {vh.x + dtf * (F.x / m), vh.y + dtf * (F.y / m), vh.z + dtf * (F.z / m)};
    const double* vs = A->vsb;
    const double4 vb4 = Vb[i];
    double vb[3] = {vb4.x, vb4.y, vb4.z};
    // vs_ab += m V'_a V'_b  (xx yy zz xy xz yz)
    vb[0] += m * (2.0 * vp[0] * vs[0] + vp[1] * vs[3] + vp[2] * vs[4]);
    vb[1] += m * (2.0 * vp[1] * vs[1] + vp[0] * vs[3] + vp[2] * vs[5]);
    vb[2] += m * (2.0 * vp[2] * vs[2] + vp[0] * vs[4] + vp[1] * vs[5]);
    double mb = vp[0] * vp[0] * vs[0] + vp[1] * vp[1] * vs[1] + vp[2] * vp[2] * vs[2] +
                vp[0] * vp[1] * vs[3] + vp[0] * vp[2] * vs[4] + vp[1] * vp[2] * vs[5];
    // V' = vh + dtf F / m
    const double4 fb4 = Fb[i];
    const double Fv[3] = {F.x, F.y, F.z};
    double fcb[3] = {fb4.x, fb4.y, fb4.z};
    double vdotf = 0.0;
    for (int k = 0; k < 3; ++k) {
        fcb[k] += dtf * vb[k] / m;
        vdotf += vb[k] * Fv[k];
    }
    mb += -dtf * vdotf / (m * m);
    // W' = wh + dtf T / I
    const double4 wb4 = Wb[i], tb4 = Tb[i];
    const double wb[3] = {wb4.x, wb4.y, wb4.z}, Tv[3] = {T.x, T.y, T.z};
    double tcb[3] = {tb4.x, tb4.y, tb4.z};
    double wdott = 0.0;
    for (int k = 0; k < 3; ++k) {
        tcb[k] += dtf * wb[k] / inertia;
        wdott += wb[k] * Tv[k];
    }
    const double ib = -dtf * wdott / (inertia * inertia);
    mb += ib * 0.4 * r * r;
    rbar[i] += ib * 0.8 * m * r;
    mbar[i] += mb;
    vhb[i] = make_double4(vb[0], vb[1], vb[2], 0.0);
    whb[i] = make_double4(wb[0], wb[1], wb[2], 0.0);
    Fcb[i] = make_double4(fcb[0], fcb[1], fcb[2], 0.0);
    Tcb[i] = make_double4(tcb[0], tcb[1], tcb[2], 0.0);
}

__device__ __forceinline__ int find_slot(const int* start, const int* jarr, int row, int j) {
    int lo = start[row], hi = start[row + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (jarr[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The contacts of the step, reversed: each atom gathers the pair adjoints of
// its row (every pair evaluated in canonical orientation, lower index first).
// Leaves: kn, kt, gamman, gammat, xmu, and the adjoint of the box lengths
// through periodic images (owner side only).
template <bool HERTZ>
__global__ void k_adj_contacts(int n, const int* start, const int* jarr, const double4* XH,
                               const double4* VH, const double4* WH, const uint8_t* FLB,
                               const uint8_t* FLA, const double4* SOLD, const double4* Sb,
                               const double4* Fcb, const double4* Tcb, const double* vsb,
                               BoxRec B, PairP pp, double dt, int shearupdate, double4* xhb,
                               double4* vhb, double4* whb, double* mbar, double* rbar,
                               double4* Sbo, double* part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        double4 xa4 = xhb[i], va4 = vhb[i], wa4 = whb[i];
        double xacc[3] = {xa4.x, xa4.y, xa4.z}, vacc[3] = {va4.x, va4.y, va4.z};
        double wacc[3] = {wa4.x, wa4.y, wa4.z};
        double macc = 0.0, racc = 0.0;
        const double4 xi = XH[i], vi = VH[i], wi = WH[i];
        for (int s = start[i]; s < start[i + 1]; ++s) {
            const int j = jarr[s];
            if (!FLA[s]) {
                Sbo[s] = make_double4(0.0, 0.0, 0.0, 0.0);
                continue;
            }
            const bool first = i < j;
            const int a = first ? i : j, b = first ? j : i;
            const double4 xj = XH[j], vj = VH[j], wj = WH[j];
            const double4 xa = first ? xi : xj, xb = first ? xj : xi;
            const double4 va = first ? vi : vj, vbv = first ? vj : vi;
            const double4 oa = first ? wi : wj, ob = first ? wj : wi;
            double d[3] = {xa.x - xb.x, xa.y - xb.y, xa.z - xb.z};
            double nimg[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < 3; ++k) {
                if (B.per[k] == 1)
                    nimg[k] = round(B.fast[k] ? d[k] * (1.0 / B.prd[k]) : d[k] / B.prd[k]);
                d[k] -= nimg[k] * B.prd[k];
            }
            const double vav[3] = {va.x, va.y, va.z}, vbvv[3] = {vbv.x, vbv.y, vbv.z};
            const double oav[3] = {oa.x, oa.y, oa.z}, obv[3] = {ob.x, ob.y, ob.z};
            const bool before = FLB[s] != 0;
            const double4 h4 = SOLD[s];
            const double hist[3] = {h4.x, h4.y, h4.z};
            PairVjpIn in;
            const double4 fa = Fcb[a], fb = Fcb[b], ta = Tcb[a], tbq = Tcb[b];
            in.Fb[0] = fa.x - fb.x; in.Fb[1] = fa.y - fb.y; in.Fb[2] = fa.z - fb.z;
            in.tab[0] = ta.x; in.tab[1] = ta.y; in.tab[2] = ta.z;
            in.tbb[0] = tbq.x; in.tbb[1] = tbq.y; in.tbb[2] = tbq.z;
            const int sm = find_slot(start, jarr, j, i);
            const double4 s1 = Sb[s], s2 = Sb[sm];
            in.shb[0] = s1.x + s2.x; in.shb[1] = s1.y + s2.y; in.shb[2] = s1.z + s2.z;
            // the pair's virial is added to both atoms' sums
            for (int k = 0; k < 6; ++k) in.virb[k] = vsb ? 2.0 * vsb[k] : 0.0;
            PairVjpOut o{};
            legacy_history_vjp<HERTZ>(d, xa.w, xb.w, va.w, vbv.w, vav, vbvv, oav, obv, pp.kn, pp.kt,
                                      pp.gamman, pp.gammat, pp.xmu, pp.limit_damping, dt,
                                      shearupdate, before, hist, in, o);
            if (first) {
                for (int k = 0; k < 3; ++k) {
                    xacc[k] += o.db[k];
                    vacc[k] += o.vab[k];
                    wacc[k] += o.oab[k];
                }
                racc += o.rab;
                macc += o.mab;
                Sbo[s] = make_double4(o.shob[0], o.shob[1], o.shob[2], 0.0);
                for (int k = 0; k < 5; ++k) leaf[k] += o.pb[k];
                for (int k = 0; k < 3; ++k) leaf[5 + k] += -nimg[k] * o.db[k];
            } else {
                for (int k = 0; k < 3; ++k) {
                    xacc[k] -= o.db[k];
                    vacc[k] += o.vbb[k];
                    wacc[k] += o.obb[k];
                }
                racc += o.rbb;
                macc += o.mbb;
                Sbo[s] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
            (void)a;
            (void)b;
        }
        xhb[i] = make_double4(xacc[0], xacc[1], xacc[2], 0.0);
        vhb[i] = make_double4(vacc[0], vacc[1], vacc[2], 0.0);
        whb[i] = make_double4(wacc[0], wacc[1], wacc[2], 0.0);
        mbar[i] += macc;
        rbar[i] += racc;
    }
    leaves_out<8>(leaf, part);
}

// initial_integrate reversed (and the PBC wrap of a rebuild): the adjoint of
// the state before the step; leaves the wrap's box-length adjoint
__global__ void k_adj_initial(int n, const double4* XH, const double4* VH, const double4* Fp,
                              const double4* Tp, const double4* xhb, const double4* vhb,
                              const double4* whb, const int* wrap, double dt, double4* Xb,
                              double4* Vb, double4* Wb, double4* Fb, double4* Tb, double* mbar,
                              double* rbar, double* part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double leaf[3] = {0, 0, 0};
    if (i < n) {
        const double dtf = 0.5 * dt;
        const double m = VH[i].w, r = XH[i].w;
        const double inertia = 0.4 * m * r * r;
        const double4 xb4 = xhb[i], vb4 = vhb[i], wb4 = whb[i], F = Fp[i], T = Tp[i];
        const double xb[3] = {xb4.x, xb4.y, xb4.z};
        if (wrap)
            for (int k = 0; k < 3; ++k) leaf[k] = (double)wrap[4 * i + k] * xb[k];
        // xh = X + dt vh
        const double vb[3] = {vb4.x + dt * xb[0], vb4.y + dt * xb[1], vb4.z + dt * xb[2]};
        // vh = V + dtf F / m, wh = W + dtf T / I
        const double Fv[3] = {F.x, F.y, F.z}, Tv[3] = {T.x, T.y, T.z};
        const double wb[3] = {wb4.x, wb4.y, wb4.z};
        double fb[3], tb[3], vdotf = 0.0, wdott = 0.0;
        for (int k = 0; k < 3; ++k) {
            fb[k] = dtf * vb[k] / m;
            tb[k] = dtf * wb[k] / inertia;
            vdotf += vb[k] * Fv[k];
            wdott += wb[k] * Tv[k];
        }
        const double ib = -dtf * wdott / (inertia * inertia);
        mbar[i] += -dtf * vdotf / (m * m) + ib * 0.4 * r * r;
        rbar[i] += ib * 0.8 * m * r;
        Xb[i] = make_double4(xb[0], xb[1], xb[2], 0.0);
        Vb[i] = make_double4(vb[0], vb[1], vb[2], 0.0);
        Wb[i] = make_double4(wb[0], wb[1], wb[2], 0.0);
        Fb[i] = make_double4(fb[0], fb[1], fb[2], 0.0);
        Tb[i] = make_double4(tb[0], tb[1], tb[2], 0.0);
    }
    leaves_out<3>(leaf, part);
}

// the step's coefficient gradients, and its box adjoints folded into cum's
__global__ void __launch_bounds__(ADJ_THREADS) k_adj_close(AdjCtl* A, const double* part_c,
                                                           const double* part_w, int nparts,
                                                           int have_wrap, tmps_servo_params S,
                                                           int fold_box) {
    double c[8], w[8];
    block_sum8(part_c, nparts, c);
    if (have_wrap) block_sum8(part_w, nparts, w);
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 5; ++k) A->grad[k] += c[k];
    for (int k = 0; k < 3; ++k) {
        double prdb = A->prdb[k] + c[5 + k] + (have_wrap ? w[k] : 0.0);
        if (fold_box && (S.style[k] == 8 || S.style[k] == 9))
            A->cumb[k] += -0.5 * A->lob[k] + prdb;
        A->lob[k] = 0.0;
        A->prdb[k] = 0.0;
    }
}

// slot adjoints carried back across a rebuild: old slot o <- new slot s
__global__ void k_adj_slots_back(int nslots_new, const int* src, const double4* Sbo, double4* Sb) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots_new) return;
    const int o = src[s];
    if (o >= 0) Sb[o] = Sbo[s];
}

__global__ void k_adj_add4(int n, const double4* a, double4* b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 x = a[i];
    double4 y = b[i];
    y.x += x.x; y.y += x.y; y.z += x.z;
    b[i] = y;
}

// r_bar += m_bar dm/dr at fixed density (m proportional to r^3)
__global__ void k_adj_radius(int n, const double4* xr, const double4* vm, const double* mbar,
                             double* rbar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = xr[i].w, m = vm[i].w;
    rbar[i] += mbar[i] * 3.0 * m / r;
}

__global__ void k_zero4(size_t n, double4* a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

__global__ void k_zero(size_t n, double* a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// ------------------------------------------------------------------ host

inline dim3 grid1(size_t n) { return dim3((unsigned)((std::max<size_t>(n, 1) + 255) / 256)); }

int zero4(System& s, double4* p, size_t n) {
    hipLaunchKernelGGL(k_zero4, grid1(n), dim3(256), 0, s.stream, n, p);
    HIP_TRY(hipGetLastError());
    return 0;
}

int zero(System& s, double* p, size_t n) {
    hipLaunchKernelGGL(k_zero, grid1(n), dim3(256), 0, s.stream, n, p);
    HIP_TRY(hipGetLastError());
    return 0;
}

int check_supported(System& s) {
    if (s.precision != TMPS_FP64)
        return tmps_fail(TMPS_ERR_UNSUPPORTED, "the adjoint runs on FP64 systems");
    if (!s.have_pair || (s.pair.style != TMPS_PAIR_HOOKE_HISTORY &&
                         s.pair.style != TMPS_PAIR_HERTZ_HISTORY))
        return tmps_fail(TMPS_ERR_UNSUPPORTED,
                         "the adjoint covers gran/hooke/history and gran/hertz/history");
    if (!(s.have_servo && s.servo.enabled))
        return tmps_fail(TMPS_ERR_UNSUPPORTED,
                         "the adjoint differentiates the pressure of fix deform/pressure: enable it");
    {
        const tmps_servo_params& S = s.servo;
        if (S.nevery != 1)
            return tmps_fail(TMPS_ERR_UNSUPPORTED, "the adjoint needs deform/pressure nevery 1");
        if (S.remap != 0)
            return tmps_fail(TMPS_ERR_UNSUPPORTED, "the adjoint needs deform/pressure remap x");
        for (int k = 0; k < 3; ++k)
            if (S.style[k] != 0 && S.style[k] != 8 && S.style[k] != 9)
                return tmps_fail(TMPS_ERR_UNSUPPORTED,
                                 "the adjoint covers the pressure and pressure/mean styles");
    }
    return 0;
}

int save_checkpoint(System& s, Checkpoint& c) {
    int rc = 0;
    c.mem.reset();
    const size_t n = (size_t)s.n;
    c.xr = arena_copy(c.mem, s.xr, n, s.stream, &rc);
    c.vm = arena_copy(c.mem, s.vm, n, s.stream, &rc);
    c.om = arena_copy(c.mem, s.om, n, s.stream, &rc);
    c.f = arena_copy(c.mem, s.f, n, s.stream, &rc);
    c.tq = arena_copy(c.mem, s.tq, n, s.stream, &rc);
    c.x0 = arena_copy(c.mem, s.x0, n, s.stream, &rc);
    c.nslots = s.nslots;
    c.nb_start = arena_copy(c.mem, s.nb_start, n + 1, s.stream, &rc);
    c.nb_j = arena_copy(c.mem, s.nb_j, (size_t)s.nslots, s.stream, &rc);
    c.nb_flag = arena_copy(c.mem, s.nb_flag, (size_t)s.nslots, s.stream, &rc);
    c.nb_hist = arena_copy(c.mem, static_cast<const double4*>(s.nb_hist), (size_t)s.nslots,
                           s.stream, &rc);
    if (rc) return rc;
    rc = tmps_api::ctl_pull(s);
    if (rc) return rc;
    c.ctl = *s.hctl;
    c.timestep = s.timestep;
    c.current_step = s.current_step;
    c.last_build_step = s.last_build_step;
    c.has_built = s.has_built;
    c.build_count = s.build_count;
    c.npairs = s.npairs;
    c.setup_done = s.setup_done;
    c.servo_step_count = s.servo_step_count;
    c.servo_nsteps = s.servo_nsteps;
    c.grid = s.grid;
    c.ref_grid = s.ref_grid;
    return 0;
}

int restore_checkpoint(System& s, const Checkpoint& c) {
    int rc = ensure_list_capacity(s, (size_t)c.nslots);
    if (rc) return rc;
    const size_t n = (size_t)s.n;
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return 0;
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s.stream));
        return 0;
    };
    rc = cp(s.xr, c.xr, n * sizeof(double4));
    if (!rc) rc = cp(s.vm, c.vm, n * sizeof(double4));
    if (!rc) rc = cp(s.om, c.om, n * sizeof(double4));
    if (!rc) rc = cp(s.f, c.f, n * sizeof(double4));
    if (!rc) rc = cp(s.tq, c.tq, n * sizeof(double4));
    if (!rc) rc = cp(s.x0, c.x0, n * sizeof(double4));
    if (!rc) rc = cp(s.nb_start, c.nb_start, (n + 1) * sizeof(int));
    if (!rc) rc = cp(s.nb_j, c.nb_j, (size_t)c.nslots * sizeof(int));
    if (!rc) rc = cp(s.nb_flag, c.nb_flag, (size_t)c.nslots);
    if (!rc) rc = cp(s.nb_hist, c.nb_hist, (size_t)c.nslots * sizeof(double4));
    if (rc) return rc;
    s.nslots = c.nslots;
    s.have_old = false;
    *s.hctl = c.ctl;
    rc = tmps_api::ctl_push(s);
    if (rc) return rc;
    s.timestep = c.timestep;
    s.current_step = c.current_step;
    s.last_build_step = c.last_build_step;
    s.has_built = c.has_built;
    s.build_count = c.build_count;
    s.npairs = c.npairs;
    s.setup_done = c.setup_done;
    s.servo_step_count = c.servo_step_count;
    s.servo_nsteps = c.servo_nsteps;
    s.grid = c.grid;
    s.ref_grid = c.ref_grid;
    drop_graphs(s);
    return 0;
}

void box_of(const Ctl& c, double lo[3], double prd[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = c.box.lo[k];
        prd[k] = c.box.prd[k];
    }
}

int ensure_state(System& s, AdjContext& A) {
    if (A.n_alloc >= s.n && A.Xb) return 0;
    const size_t n = (size_t)std::max(s.n, 1);
    dfree(A.Xb); dfree(A.Vb); dfree(A.Wb); dfree(A.Fb); dfree(A.Tb);
    dfree(A.xhb); dfree(A.vhb); dfree(A.whb); dfree(A.Fcb); dfree(A.Tcb);
    dfree(A.mbar); dfree(A.rbar); dfree(A.part); dfree(A.tmpx);
    int rc = dmalloc(&A.Xb, n);
    if (!rc) rc = dmalloc(&A.Vb, n);
    if (!rc) rc = dmalloc(&A.Wb, n);
    if (!rc) rc = dmalloc(&A.Fb, n);
    if (!rc) rc = dmalloc(&A.Tb, n);
    if (!rc) rc = dmalloc(&A.xhb, n);
    if (!rc) rc = dmalloc(&A.vhb, n);
    if (!rc) rc = dmalloc(&A.whb, n);
    if (!rc) rc = dmalloc(&A.Fcb, n);
    if (!rc) rc = dmalloc(&A.Tcb, n);
    if (!rc) rc = dmalloc(&A.mbar, n);
    if (!rc) rc = dmalloc(&A.rbar, n);
    if (!rc) rc = dmalloc(&A.tmpx, n);
    if (!rc) rc = dmalloc(&A.part, (size_t)2 * nblocks_for((int)n) * 8);
    if (!rc && !A.actl) rc = dmalloc(&A.actl, 1);
    if (rc) return rc;
    A.n_alloc = (int)n;
    return 0;
}

int ensure_slots(AdjContext& A, size_t need) {
    if (need <= A.sb_cap && A.Sb) return 0;
    dfree(A.Sb);
    dfree(A.Sbo);
    const size_t cap = need + need / 4 + 1024;
    int rc = dmalloc(&A.Sb, cap);
    if (!rc) rc = dmalloc(&A.Sbo, cap);
    if (rc) return rc;
    A.sb_cap = cap;
    return 0;
}

// Replays steps [cp.timestep + 1, cp.timestep + K] through the stage path from
// checkpoint `cp`, recording each step's inputs.
int replay_segment(System& s, AdjContext& A, const Checkpoint& cp, int K) {
    int rc = restore_checkpoint(s, cp);
    if (rc) return rc;
    const size_t n = (size_t)s.n;
    A.rec.reset();
    A.steps.assign(K, StepRec());
    A.lists.clear();
    ListVer v0;
    v0.start = cp.nb_start;
    v0.j = cp.nb_j;
    v0.nslots = cp.nslots;
    A.lists.push_back(v0);
    A.flb.assign(K, nullptr);
    A.fla.assign(K, nullptr);
    A.sold.assign(K, nullptr);
    if (K > A.rec_steps_cap) {
        dfree(A.XH); dfree(A.VH); dfree(A.WH); dfree(A.FT); dfree(A.TT); dfree(A.ctl_rec);
        rc = dmalloc(&A.XH, (size_t)K * n);
        if (!rc) rc = dmalloc(&A.VH, (size_t)K * n);
        if (!rc) rc = dmalloc(&A.WH, (size_t)K * n);
        if (!rc) rc = dmalloc(&A.FT, (size_t)K * n);
        if (!rc) rc = dmalloc(&A.TT, (size_t)K * n);
        if (!rc) rc = dmalloc(&A.ctl_rec, (size_t)K);
        if (rc) return rc;
        A.rec_steps_cap = K;
    }
    const double dt = A.dt;
    const bool servo = s.have_servo && s.servo.enabled;
    // The steps run on the device back to back: the list is rebuilt at the
    // steps the forward pass rebuilt at (no displacement check, whose result
    // is known), and each step's control block after its servo is recorded
    // on the device; the host reads the records once, after the segment, for
    // the boxes the reverse steps need. Only a rebuild waits for the device
    // (its build reads the box and the slot count).
    for (int k = 0; k < K; ++k) {
        StepRec& R = A.steps[k];
        for (int q = 0; q < 3; ++q) R.box.per[q] = cp.ctl.box.per[q];
        s.timestep += 1;
        rc = launch_nve_initial_f64(s, dt);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(A.tmpx, s.xr, n * sizeof(double4), hipMemcpyDeviceToDevice,
                               s.stream));
        s.current_step = s.timestep;
        const int rebuilt =
            std::binary_search(A.build_steps.begin(), A.build_steps.end(), s.timestep) ? 1 : 0;
        if (rebuilt) {
            rc = tmps_api::reneighbor_now(s);
            if (rc) return rc;
            // the build read the box of the step into the host mirror
            box_of(*s.hctl, R.box.lo, R.box.prd);
        }
        R.list = (int)A.lists.size() - 1;
        R.rebuilt = rebuilt;
        if (rebuilt) {
            ListVer v;
            v.nslots = s.nslots;
            v.start = arena_copy(A.rec, s.nb_start, n + 1, s.stream, &rc);
            v.j = arena_copy(A.rec, s.nb_j, (size_t)s.nslots, s.stream, &rc);
            int* src = static_cast<int*>(A.rec.get(std::max<size_t>(s.nslots, 1) * sizeof(int)));
            int* wrap = static_cast<int*>(A.rec.get(n * 4 * sizeof(int)));
            if (rc || !src || !wrap) return rc ? rc : tmps_fail(TMPS_ERR_HIP, "adjoint: out of memory");
            hipLaunchKernelGGL(k_adj_remap_src, grid1(n), dim3(256), 0, s.stream, s.n, s.nb_start,
                               s.nb_j, s.onb_start, s.onb_j, s.onb_flag, s.onb_n, src);
            HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(k_adj_wrap, grid1(n), dim3(256), 0, s.stream, s.n, A.tmpx, s.xr,
                               R.box, wrap);
            HIP_TRY(hipGetLastError());
            v.src = src;
            v.wrap = wrap;
            A.lists.push_back(v);
            R.list = (int)A.lists.size() - 1;
        }
        for (int q = 0; q < 3; ++q) R.box.fast[q] = s.hctl->fast_img[q];
        HIP_TRY(hipMemcpyAsync(A.XH + (size_t)k * n, s.xr, n * sizeof(double4),
                               hipMemcpyDeviceToDevice, s.stream));
        HIP_TRY(hipMemcpyAsync(A.VH + (size_t)k * n, s.vm, n * sizeof(double4),
                               hipMemcpyDeviceToDevice, s.stream));
        HIP_TRY(hipMemcpyAsync(A.WH + (size_t)k * n, s.om, n * sizeof(double4),
                               hipMemcpyDeviceToDevice, s.stream));
        A.flb[k] = arena_copy(A.rec, s.nb_flag, (size_t)s.nslots, s.stream, &rc);
        A.sold[k] = arena_copy(A.rec, static_cast<const double4*>(s.nb_hist), (size_t)s.nslots,
                               s.stream, &rc);
        if (rc) return rc;
        rc = tmps_api::clear_forces(s);
        if (!rc) rc = tmps_api::pair_compute(s, dt, 1, 0);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(A.FT + (size_t)k * n, s.f, n * sizeof(double4),
                               hipMemcpyDeviceToDevice, s.stream));
        HIP_TRY(hipMemcpyAsync(A.TT + (size_t)k * n, s.tq, n * sizeof(double4),
                               hipMemcpyDeviceToDevice, s.stream));
        A.fla[k] = arena_copy(A.rec, s.nb_flag, (size_t)s.nslots, s.stream, &rc);
        if (rc) return rc;
        rc = launch_nve_final_f64(s, dt);
        if (!rc) rc = tmps_api::servo_end_of_step_launch(s, dt);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(A.ctl_rec + k, s.ctl, sizeof(Ctl), hipMemcpyDeviceToDevice,
                               s.stream));
    }
    A.ctl_host.resize((size_t)K);
    HIP_TRY(copy_sync(s.stream, A.ctl_host.data(), A.ctl_rec, (size_t)K * sizeof(Ctl),
                      hipMemcpyDeviceToHost));
    for (int k = 0; k < K; ++k) {
        StepRec& R = A.steps[k];
        const Ctl& before = k == 0 ? cp.ctl : A.ctl_host[(size_t)k - 1];
        const Ctl& after = A.ctl_host[(size_t)k];
        if (after.error)
            return tmps_fail(after.error, "adjoint replay: fix deform/pressure drove a box "
                                          "dimension to non-positive");
        box_of(before, R.box.lo, R.box.prd);
        box_of(after, R.box.lo_n, R.box.prd_n);
        for (int q = 0; q < 6; ++q) R.box.tensor[q] = after.tensor[q];
        if (!servo) {
            // no servo: the box stays and nothing reads the tensor
            box_of(before, R.box.lo_n, R.box.prd_n);
            for (int q = 0; q < 6; ++q) R.box.tensor[q] = 0.0;
        }
    }
    // the host mirror follows the device to the segment's end
    *s.hctl = A.ctl_host[(size_t)K - 1];
    s.servo_step_count = s.hctl->servo.step_count;
    s.servo_nsteps = s.hctl->servo.nsteps;
    return 0;
}

template <bool HERTZ>
int launch_contacts(System& s, AdjContext& A, const ListVer& L, const double4* XH,
                    const double4* VH, const double4* WH, const uint8_t* FLB, const uint8_t* FLA,
                    const double4* SOLD, const double* vsb, const BoxRec& B, int shearupdate) {
    const int nb = nblocks_for(s.n);
    hipLaunchKernelGGL(k_adj_contacts<HERTZ>, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                       L.start, L.j, XH, VH, WH, FLB, FLA, SOLD, A.Sb, A.Fcb, A.Tcb, vsb, B,
                       s.pair, A.dt, shearupdate, A.xhb, A.vhb, A.whb, A.mbar, A.rbar, A.Sbo,
                       A.part);
    HIP_TRY(hipGetLastError());
    return 0;
}

int contacts(System& s, AdjContext& A, const ListVer& L, const double4* XH, const double4* VH,
             const double4* WH, const uint8_t* FLB, const uint8_t* FLA, const double4* SOLD,
             const double* vsb, const BoxRec& B, int shearupdate) {
    if (s.pair.style == TMPS_PAIR_HERTZ_HISTORY)
        return launch_contacts<true>(s, A, L, XH, VH, WH, FLB, FLA, SOLD, vsb, B, shearupdate);
    return launch_contacts<false>(s, A, L, XH, VH, WH, FLB, FLA, SOLD, vsb, B, shearupdate);
}

// One step reversed (the step recorded at index k of the replayed segment).
// dL: adjoint of the step's pressure tensor from the loss (final step only).
int reverse_step(System& s, AdjContext& A, const Checkpoint& cp, int k, const double dL[6]) {
    const size_t n = (size_t)s.n;
    const int nb = nblocks_for(s.n);
    const StepRec& R = A.steps[k];
    const ListVer& L = A.lists[R.list];
    const double4* XH = A.XH + (size_t)k * n;
    const double4* VH = A.VH + (size_t)k * n;
    const double4* WH = A.WH + (size_t)k * n;
    const double4* FT = A.FT + (size_t)k * n;
    const double4* TT = A.TT + (size_t)k * n;
    double* part_c = A.part;
    double* part_w = A.part + (size_t)nb * 8;
    const bool servo = s.have_servo && s.servo.enabled;
    // end of step: remap and servo
    if (servo) {
        hipLaunchKernelGGL(k_adj_remap, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.n, A.Xb, XH,
                           R.box, A.xhb, part_c);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_adj_servo, dim3(1), dim3(ADJ_THREADS), 0, s.stream, A.actl, part_c, nb,
                           R.box, s.servo, A.dt, dL[0], dL[1], dL[2], dL[3], dL[4], dL[5]);
        HIP_TRY(hipGetLastError());
    } else {
        HIP_TRY(hipMemcpyAsync(A.xhb, A.Xb, n * sizeof(double4), hipMemcpyDeviceToDevice,
                               s.stream));
    }
    hipLaunchKernelGGL(k_adj_final, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.n, XH, VH, FT, TT,
                       A.Vb, A.Wb, A.Fb, A.Tb, A.actl, A.dt, A.vhb, A.whb, A.Fcb, A.Tcb, A.mbar,
                       A.rbar);
    HIP_TRY(hipGetLastError());
    // forces
    const double* vsb = reinterpret_cast<const double*>(reinterpret_cast<const char*>(A.actl) +
                                                        offsetof(AdjCtl, vsb));
    int rc = contacts(s, A, L, XH, VH, WH, A.flb[k], A.fla[k], A.sold[k], servo ? vsb : nullptr,
                      R.box, 1);
    if (rc) return rc;
    // start of step
    const double4* Fp = k > 0 ? A.FT + (size_t)(k - 1) * n : cp.f;
    const double4* Tp = k > 0 ? A.TT + (size_t)(k - 1) * n : cp.tq;
    hipLaunchKernelGGL(k_adj_initial, dim3(nb), dim3(TMPS_BLOCK), 0, s.stream, s.n, XH, VH, Fp, Tp,
                       A.xhb, A.vhb, A.whb, R.rebuilt ? L.wrap : nullptr, A.dt, A.Xb, A.Vb, A.Wb,
                       A.Fb, A.Tb, A.mbar, A.rbar, part_w);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_adj_close, dim3(1), dim3(ADJ_THREADS), 0, s.stream, A.actl, part_c, part_w,
                       nb, R.rebuilt, s.servo, servo ? 1 : 0);
    HIP_TRY(hipGetLastError());
    // slot adjoints of the previous step's shear
    if (R.rebuilt) {
        const ListVer& P = A.lists[R.list - 1];
        rc = zero4(s, A.Sb, (size_t)P.nslots);
        if (rc) return rc;
        hipLaunchKernelGGL(k_adj_slots_back, grid1(L.nslots), dim3(256), 0, s.stream, L.nslots,
                           L.src, A.Sbo, A.Sb);
        HIP_TRY(hipGetLastError());
    } else {
        std::swap(A.Sb, A.Sbo);
    }
    return 0;
}

}  // namespace

void adjoint_free(System& s) {
    AdjContext* A = s.adj;
    if (!A) return;
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    for (auto* c : A->cps) {
        c->mem.release();
        delete c;
    }
    A->cps.clear();
    A->setup_mem.release();
    A->rec.release();
    dfree(A->XH); dfree(A->VH); dfree(A->WH); dfree(A->FT); dfree(A->TT);
    dfree(A->Xb); dfree(A->Vb); dfree(A->Wb); dfree(A->Fb); dfree(A->Tb);
    dfree(A->xhb); dfree(A->vhb); dfree(A->whb); dfree(A->Fcb); dfree(A->Tcb);
    dfree(A->mbar); dfree(A->rbar); dfree(A->Sb); dfree(A->Sbo); dfree(A->actl);
    dfree(A->part); dfree(A->tmpx); dfree(A->ctl_rec);
    delete A;
    s.adj = nullptr;
}

extern "C" {

int tmps_adjoint_forward(tmps_system* h, int nsteps, int every, double dt, double tensor_out[6],
                         double* coordination) {
    System& s = *reinterpret_cast<System*>(h);
    s.hctl_fresh = false;
    if (nsteps < 1 || every < 1) return tmps_fail(TMPS_ERR_ARG, "nsteps and every must be >= 1");
    if (s.n <= 0) return tmps_fail(TMPS_ERR_ARG, "the adjoint needs atoms");
    int rc = check_supported(s);
    if (rc) return rc;
    adjoint_free(s);
    s.adj = new AdjContext();
    AdjContext& A = *s.adj;
    A.nsteps = nsteps;
    A.every = every;
    A.dt = dt;
    // Verlet::setup (simulation.py:189-203): the setup force, with the
    // touching flags and stored shear it starts from recorded
    s.setup_done = 1;
    rc = tmps_api::clear_forces(s);
    int rebuilt = 0;
    if (!rc) rc = tmps_api::reneighbor(s, true, &rebuilt);
    if (rc) return rc;
    s.rebuilds += rebuilt;
    A.setup_mem.reset();
    A.setup_flb = arena_copy(A.setup_mem, s.nb_flag, (size_t)s.nslots, s.stream, &rc);
    A.setup_sold = arena_copy(A.setup_mem, static_cast<const double4*>(s.nb_hist),
                              (size_t)s.nslots, s.stream, &rc);
    if (!rc) rc = tmps_api::pair_compute(s, dt, 0, 1);
    if (rc) return rc;
    s.build_log.clear();
    s.log_builds = true;
    for (int t0 = 0; t0 < nsteps && !rc; t0 += every) {
        Checkpoint* c = new Checkpoint();
        A.cps.push_back(c);
        rc = save_checkpoint(s, *c);
        if (!rc) rc = tmps_api::run(s, std::min(every, nsteps - t0), dt, 0, 0, 0.0, 0.0);
    }
    s.log_builds = false;
    A.build_steps = s.build_log;
    if (rc) return rc;
    Checkpoint* end = new Checkpoint();
    A.cps.push_back(end);
    rc = save_checkpoint(s, *end);
    if (rc) return rc;
    if (tensor_out)
        for (int k = 0; k < 6; ++k) tensor_out[k] = s.hctl->tensor[k];
    if (coordination) *coordination = (double)s.hctl->ntouch / (double)s.n;
    A.done_forward = 1;
    return 0;
}

int tmps_adjoint_backward(tmps_system* h, const double dL_dtensor[6], double grad_coeffs[5],
                          double* grad_radius, double* grad_x, double* grad_v,
                          double* grad_omega) {
    System& s = *reinterpret_cast<System*>(h);
    s.hctl_fresh = false;
    AdjContext* Ap = s.adj;
    if (!Ap || !Ap->done_forward)
        return tmps_fail(TMPS_ERR_ARG, "tmps_adjoint_backward needs tmps_adjoint_forward first");
    AdjContext& A = *Ap;
    int rc = ensure_state(s, A);
    if (rc) return rc;
    const size_t n = (size_t)s.n;
    const Checkpoint& endc = *A.cps.back();
    size_t maxslots = (size_t)endc.nslots;
    for (auto* c : A.cps) maxslots = std::max(maxslots, (size_t)c->nslots);
    rc = ensure_slots(A, maxslots);
    for (double4* p : {A.Xb, A.Vb, A.Wb, A.Fb, A.Tb})
        if (!rc) rc = zero4(s, p, n);
    if (!rc) rc = zero(s, A.mbar, n);
    if (!rc) rc = zero(s, A.rbar, n);
    if (!rc) rc = zero4(s, A.Sb, A.sb_cap);
    if (!rc) rc = zero(s, reinterpret_cast<double*>(A.actl), sizeof(AdjCtl) / sizeof(double));
    if (rc) return rc;
    const double zero6[6] = {0, 0, 0, 0, 0, 0};
    const int nseg = (int)A.cps.size() - 1;
    for (int g = nseg - 1; g >= 0; --g) {
        const Checkpoint& cp = *A.cps[g];
        const int K = std::min(A.every, A.nsteps - g * A.every);
        rc = replay_segment(s, A, cp, K);
        if (rc) return rc;
        // the slot adjoints must fit the lists of this segment
        for (const ListVer& L : A.lists)
            if ((size_t)L.nslots > A.sb_cap)
                return tmps_fail(TMPS_ERR_CAPACITY, "adjoint slot buffer too small");
        for (int k = K - 1; k >= 0; --k) {
            const bool last = g == nseg - 1 && k == K - 1;
            rc = reverse_step(s, A, cp, k, last ? dL_dtensor : zero6);
            if (rc) return rc;
        }
    }
    // the setup force (shearupdate off, its virial unused), from checkpoint 0
    {
        const Checkpoint& c0 = *A.cps[0];
        ListVer L;
        L.start = c0.nb_start;
        L.j = c0.nb_j;
        L.nslots = c0.nslots;
        BoxRec B{};
        box_of(c0.ctl, B.lo, B.prd);
        for (int q = 0; q < 3; ++q) {
            B.per[q] = c0.ctl.box.per[q];
            B.fast[q] = c0.ctl.fast_img[q];
        }
        HIP_TRY(hipMemcpyAsync(A.Fcb, A.Fb, n * sizeof(double4), hipMemcpyDeviceToDevice,
                               s.stream));
        HIP_TRY(hipMemcpyAsync(A.Tcb, A.Tb, n * sizeof(double4), hipMemcpyDeviceToDevice,
                               s.stream));
        for (double4* p : {A.xhb, A.vhb, A.whb})
            if (!rc) rc = zero4(s, p, n);
        if (!rc)
            rc = contacts(s, A, L, c0.xr, c0.vm, c0.om, A.setup_flb, c0.nb_flag, A.setup_sold,
                          nullptr, B, 0);
        if (rc) return rc;
        const int nb = nblocks_for(s.n);
        hipLaunchKernelGGL(k_adj_close, dim3(1), dim3(ADJ_THREADS), 0, s.stream, A.actl, A.part,
                           A.part, nb, 0, s.servo, 0);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_adj_add4, grid1(n), dim3(256), 0, s.stream, s.n, A.xhb, A.Xb);
        hipLaunchKernelGGL(k_adj_add4, grid1(n), dim3(256), 0, s.stream, s.n, A.vhb, A.Vb);
        hipLaunchKernelGGL(k_adj_add4, grid1(n), dim3(256), 0, s.stream, s.n, A.whb, A.Wb);
        hipLaunchKernelGGL(k_adj_radius, grid1(n), dim3(256), 0, s.stream, s.n, c0.xr, c0.vm,
                           A.mbar, A.rbar);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(s.stream));
    AdjCtl hc;
    HIP_TRY(copy_sync(s.stream, &hc, A.actl, sizeof(AdjCtl), hipMemcpyDeviceToHost));
    if (grad_coeffs)
        for (int k = 0; k < 5; ++k) grad_coeffs[k] = hc.grad[k];
    if (grad_radius) HIP_TRY(copy_sync(s.stream, grad_radius, A.rbar, n * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double4> tmp(grad_x || grad_v || grad_omega ? n : 0);
    const double4* srcs[3] = {A.Xb, A.Vb, A.Wb};
    double* dsts[3] = {grad_x, grad_v, grad_omega};
    for (int q = 0; q < 3; ++q) {
        if (!dsts[q]) continue;
        HIP_TRY(copy_sync(s.stream, tmp.data(), srcs[q], n * sizeof(double4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            dsts[q][3 * i + 0] = tmp[i].x;
            dsts[q][3 * i + 1] = tmp[i].y;
            dsts[q][3 * i + 2] = tmp[i].z;
        }
    }
    // leave the system where the forward run ended
    return restore_checkpoint(s, endc);
}

}  // extern "C"
