// Neighbor-list rebuild: PBC wrap, binning (a radix sort of the cells; the
// counting sort for ensemble members), exclusive scans, the full per-atom
// neighbor list and the contact-history carry-over.
//
// Built with -ffp-contract=off: the bin of every atom and the cutoff test of
// every pair round exactly as the CPU oracle's, so bins and neighbor sets are
// bitwise identical to it.
#include "pbc.h"
#include "tmps_internal.h"

using tmps::pbc_atom;

namespace {

// ------------------------------------------------------------------ PBC

__global__ void k_pbc(double4* xr, double4* vm, int n, const Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pbc_atom(i, xr, vm, ctl);
}

// ------------------------------------------------------------------ bins

// NeighborList.get_cell_coord, neighbor.py:191-206 (lamda with zero tilt)
__device__ __forceinline__ void cell_coord(const DevBox& B, const GridP& G, const double4& p,
                                           int c[3]) {
    const double d0 = p.x - B.lo[0], d1 = p.y - B.lo[1], d2 = p.z - B.lo[2];
    double lam[3];
    lam[0] = d0 / B.prd[0] + 0.0 * d1 + 0.0 * d2;
    lam[1] = d1 / B.prd[1] + 0.0 * d2;
    lam[2] = d2 / B.prd[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ck = (int)floor(lam[k] / G.csize[k]);
        c[k] = max(0, min(ck, G.g[k] - 1));
    }
}

__global__ void k_zero_int(int* a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// a[0, n) = b[0, n) = 0, and c[0, nc) = 0
__global__ void k_zero_int3(int* a, int* b, int n, int* c, int nc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = 0;
        b[i] = 0;
    }
    if (i < nc) c[i] = 0;
}

// The current list into the copy the history remap merges with (the
// reference's save_state, contact_history.py:141-145): row starts, partners,
// touch flags and histories, the slot count read on the device.
// Items first, first + stride, ... of the copy.
template <typename H>
__device__ __forceinline__ void list_save_items(int first, int stride, int n, const int* nb_start,
                                                const int* nb_j, const uint8_t* nb_flag,
                                                const H* nb_hist, int* onb_start, int* onb_j,
                                                uint8_t* onb_flag, H* onb_hist, int cap) {
    const int nslots = nb_start[n];
    for (int k = first; k <= n || k < nslots; k += stride) {
        if (k <= n) onb_start[k] = nb_start[k];
        if (k < nslots && k < cap) {
            onb_j[k] = nb_j[k];
            const uint8_t fl = nb_flag[k];
            onb_flag[k] = fl;
            // the restore reads a saved history only where its flag is set
            if (fl) onb_hist[k] = nb_hist[k];
        }
    }
}

template <typename H>
__global__ void k_list_save(int n, const int* nb_start, const int* nb_j, const uint8_t* nb_flag,
                            const H* nb_hist, int* onb_start, int* onb_j, uint8_t* onb_flag,
                            H* onb_hist, int cap) {
    list_save_items<H>(blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, n, nb_start,
                       nb_j, nb_flag, nb_hist, onb_start, onb_j, onb_flag, onb_hist, cap);
}

// build_grid count phase, neighbor.py:264-268
__device__ __forceinline__ void bin_count_atom(int i, const double4* xr, const Ctl* ctl,
                                               const GridP& G, int* cell_of, int* count) {
    int c[3];
    cell_coord(ctl->box, G, xr[i], c);
    const int idx = c[0] + c[1] * G.g[0] + c[2] * G.g[0] * G.g[1];
    cell_of[i] = idx;
    atomicAdd(&count[idx], 1);
}

// Gd: the grid of a build decided on the device (ctl->lgrid), else G
__global__ void k_bin_count(const double4* xr, int n, const Ctl* ctl, GridP G, const GridP* Gd,
                            int* cell_of, int* count) {
    if (Gd && ctl->error) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bin_count_atom(i, xr, ctl, Gd ? *Gd : G, cell_of, count);
}

// build_grid scatter phase, neighbor.py:276-282
__device__ __forceinline__ void bin_fill_atom(int i, const int* cell_of, const int* start,
                                              int* fill, int* particles) {
    const int c = cell_of[i];
    const int slot = atomicAdd(&fill[c], 1);
    particles[start[c] + slot] = i;
}

__global__ void k_bin_fill(int n, const int* cell_of, const int* start, int* fill, int* particles) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bin_fill_atom(i, cell_of, start, fill, particles);
}

// Bin members in ascending atom index (the atomic scatter leaves them in
// arrival order).
__device__ __forceinline__ void bin_sort_cell(int c, const int* start, int* particles) {
    const int a = start[c], b = start[c + 1];
    for (int p = a + 1; p < b; ++p) {
        const int v = particles[p];
        int q = p - 1;
        while (q >= a && particles[q] > v) {
            particles[q + 1] = particles[q];
            --q;
        }
        particles[q + 1] = v;
    }
}

__global__ void k_bin_sort(int ncell, const int* start, int* particles) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell) bin_sort_cell(c, start, particles);
}

// ------------------------------------------------------------------ scan

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Exclusive scan of a block's TMPS_BLOCK values; returns the block total.
__device__ __forceinline__ int block_excl_scan(int v, int* excl) {
    __shared__ int wtot[TMPS_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int inc = wave_incl_scan(v);
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    int woff = 0, total = 0;
    for (int w = 0; w < TMPS_BLOCK / 64; ++w) {
        if (w < wave) woff += wtot[w];
        total += wtot[w];
    }
    *excl = woff + inc - v;
    __syncthreads();
    return total;
}

#define SCAN_PER_THREAD 4
#define SCAN_TILE (TMPS_BLOCK * SCAN_PER_THREAD)
static_assert(SCAN_TILE == TMPS_SCAN_TILE, "scan tile of the batched redo's maps");

// Tile `tile` of the exclusive scan; its sum into tile_sums[tile].
__device__ __forceinline__ void scan_tile(int tile, const int* in, int* out, int n,
                                          int* tile_sums) {
    const int base = tile * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    int v[SCAN_PER_THREAD];
    int sum = 0;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        sum += v[k];
    }
    int excl;
    const int total = block_excl_scan(sum, &excl);
    int run = excl;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tile_sums[tile] = total;
}

__global__ void k_scan_tiles(const int* in, int* out, int n, int* tile_sums) {
    scan_tile(blockIdx.x, in, out, n, tile_sums);
}

// One block: exclusive scan of the tile sums in place; writes the grand total.
__device__ __forceinline__ void scan_sums(int* tile_sums, int ntiles, int* total_out,
                                          int* total_copy) {
    const int per = (ntiles + TMPS_BLOCK - 1) / TMPS_BLOCK;
    const int a = threadIdx.x * per;
    int sum = 0;
    for (int k = 0; k < per; ++k)
        if (a + k < ntiles) sum += tile_sums[a + k];
    int excl;
    const int total = block_excl_scan(sum, &excl);
    int run = excl;
    for (int k = 0; k < per; ++k)
        if (a + k < ntiles) {
            const int t = tile_sums[a + k];
            tile_sums[a + k] = run;
            run += t;
        }
    if (threadIdx.x == 0) {
        *total_out = total;
        if (total_copy) *total_copy = total;
    }
}

__global__ void k_scan_sums(int* tile_sums, int ntiles, int* total_out, int* total_copy) {
    scan_sums(tile_sums, ntiles, total_out, total_copy);
}

__global__ void k_scan_add(int* out, int n, const int* tile_sums) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] += tile_sums[i / SCAN_TILE];
}

// ------------------------------------------------------------------ list

// Minimum image of one component for the list build. Where the dimension
// holds at least 3 bins (so the box is at least two cutoffs wide) the image
// count comes from a reciprocal multiply: it can differ from round(d / prd)
// only when d is within rounding of half a box, and then both images are
// beyond the cutoff, so the neighbor set and every accepted rsq are the
// reference's.
__device__ __forceinline__ double list_image(double d, double prd, double inv_prd, bool fast) {
    return d - prd * round(fast ? d * inv_prd : d / prd);
}

// Scratch row length of the single-pass list build (longer rows take the
// two-pass count-then-fill path).
#define LIST_ROW_CAP 32

// The stencil walk of NeighborList.build_neighbor_list (neighbor.py:315-369),
// visiting bins in the reference's (dx, dy, dz) order, writing each atom's
// full row at nb_start[i] and sorting it ascending in j: the path for lists
// with a row longer than k_list_rows' scratch (whose counts gave nb_start).
__device__ __forceinline__ void list_fill_atom(int i, const double4* xr, const Ctl* ctl,
                                               const GridP& G, const int* cell_start,
                                               const int* particles, double skin,
                                               const int* nb_start, int* nb_j) {
    const DevBox& B = ctl->box;
    const double4 pi = xr[i];
    int c[3];
    cell_coord(B, G, pi, c);
    const int gx = G.g[0], gy = G.g[1], gz = G.g[2], gxy = gx * gy;
    const int px = B.per[0], py = B.per[1], pz = B.per[2];
    const double prd[3] = {B.prd[0], B.prd[1], B.prd[2]};
    const double inv[3] = {1.0 / prd[0], 1.0 / prd[1], 1.0 / prd[2]};
    const bool fast[3] = {G.g[0] >= 3, G.g[1] >= 3, G.g[2] >= 3};
    int nfull = 0;
    int* row = nb_j + nb_start[i];
    for (int dx = -1; dx <= 1; ++dx) {
        if (px == 1 && gx == 1 && dx != 0) continue;
        int nx = c[0] + dx;
        if (px == 1) nx = nx < 0 ? nx + gx : (nx >= gx ? nx - gx : nx);
        if (nx < 0 || nx >= gx) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            if (py == 1 && gy == 1 && dy != 0) continue;
            int ny = c[1] + dy;
            if (py == 1) ny = ny < 0 ? ny + gy : (ny >= gy ? ny - gy : ny);
            if (ny < 0 || ny >= gy) continue;
            for (int dz = -1; dz <= 1; ++dz) {
                if (pz == 1 && gz == 1 && dz != 0) continue;
                int nz = c[2] + dz;
                if (pz == 1) nz = nz < 0 ? nz + gz : (nz >= gz ? nz - gz : nz);
                if (nz < 0 || nz >= gz) continue;
                const int cidx = nx + ny * gx + nz * gxy;
                const int pa = cell_start[cidx], pb = cell_start[cidx + 1];
                for (int p = pa; p < pb; ++p) {
                    const int j = particles[p];
                    if (j == i) continue;
                    const double4 pj = xr[j];
                    double d0 = pi.x - pj.x, d1 = pi.y - pj.y, d2 = pi.z - pj.z;
                    if (px == 1) d0 = list_image(d0, prd[0], inv[0], fast[0]);
                    if (py == 1) d1 = list_image(d1, prd[1], inv[1], fast[1]);
                    if (pz == 1) d2 = list_image(d2, prd[2], inv[2], fast[2]);
                    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
                    const double rad_sum = pi.w + pj.w + skin;
                    if (rsq < rad_sum * rad_sum) row[nfull++] = j;
                }
            }
        }
    }
    for (int p = 1; p < nfull; ++p) {
        const int v = row[p];
        int q = p - 1;
        while (q >= 0 && row[q] > v) {
            row[q + 1] = row[q];
            --q;
        }
        row[q + 1] = v;
    }
}

// Gd / flags: a build decided on the device, which launches this kernel
// whatever the row lengths: it writes only where a row exceeded the scratch
// (flags[1]) and the list fits its buffers (no error recorded)
__global__ void k_list_fill(const double4* xr, int n, const Ctl* ctl, GridP G, const GridP* Gd,
                            const int* flags, const int* cell_start, const int* particles,
                            double skin, const int* nb_start, int* nb_j) {
    if (flags && (!flags[1] || ctl->error)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) list_fill_atom(i, xr, ctl, Gd ? *Gd : G, cell_start, particles, skin, nb_start, nb_j);
}

// Positions in bin order (after k_bin_sort): bx[p] = xr[particles[p]], so the
// list build reads a bin's candidates and their positions as two contiguous
// streams instead of a dependent gather per candidate.
__global__ void k_bin_gather(int n, const int* particles, const double4* xr, double4* bx) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) bx[p] = xr[particles[p]];
}

// Whether the lean candidate walk applies (see list_rows_item): positions
// wrapped into the box and at least 3 bins on every periodic axis.
__device__ __forceinline__ bool list_lean(const DevBox& B, const GridP& G, int wrapped) {
    return wrapped && (B.per[0] != 1 || G.g[0] >= 3) && (B.per[1] != 1 || G.g[1] >= 3) &&
           (B.per[2] != 1 || G.g[2] >= 3);
}

// A row entry v inserted in ascending order into a thread's LDS column of k
// sorted entries (column stride TMPS_BLOCK).
__device__ __forceinline__ void column_insert(int* row, int k, int v) {
    int q2 = k - 1;
    while (q2 >= 0 && row[q2 * TMPS_BLOCK] > v) {
        row[(q2 + 1) * TMPS_BLOCK] = row[q2 * TMPS_BLOCK];
        --q2;
    }
    row[(q2 + 1) * TMPS_BLOCK] = v;
}

// Accepted neighbour j of atom i, inserted in ascending order into the
// thread's LDS row (column of a [LIST_ROW_CAP][TMPS_BLOCK] array); counted
// past the cap.
__device__ __forceinline__ void list_insert(int i, int j, int* row, int& nfull, int& nhalf) {
    if (nfull < LIST_ROW_CAP) column_insert(row, nfull, j);
    ++nfull;
    if (j > i) ++nhalf;
}

// Candidate j at pj for atom i at pi: the cutoff test of k_list_fill, each
// accepted j inserted in ascending order into the thread's LDS row (column
// of a [LIST_ROW_CAP][TMPS_BLOCK] array). `img` selects the axes that take
// the nearest image (bit k: axis k); the others use the plain difference.
// On a periodic axis of at least 3 bins, a candidate in a bin reached
// without wrapping differs from i by less than two bin widths: if that is
// under half the box, the nearest image is the plain difference itself
// (round gives 0, and d - prd * 0 is d); if not, both it and the nearest
// image exceed a bin width, so the axis alone rules the pair out either way.
// Only wrapped bins, and periodic axes of fewer than 3 bins, need the image,
// and the accepted set is the one the image on every periodic axis gives.
// That needs every position inside the box on its periodic axes (a build
// after the PBC wrap); other builds take the image on every periodic axis.
__device__ __forceinline__ void list_test(int i, int j, const double4& pi, const double4& pj,
                                          int img, const double prd[3], const double inv[3],
                                          const bool fast[3], double skin, int* row, int& nfull,
                                          int& nhalf) {
    if (j == i) return;
    double d0 = pi.x - pj.x, d1 = pi.y - pj.y, d2 = pi.z - pj.z;
    if (img & 1) d0 = list_image(d0, prd[0], inv[0], fast[0]);
    if (img & 2) d1 = list_image(d1, prd[1], inv[1], fast[1]);
    if (img & 4) d2 = list_image(d2, prd[2], inv[2], fast[2]);
    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    const double rad_sum = pi.w + pj.w + skin;
    if (rsq < rad_sum * rad_sum) list_insert(i, j, row, nfull, nhalf);
}

// A candidate of list_lean_range: the separation on each axis is the plain
// difference minus that axis's shift s (0, or +-prd where the candidate's bin
// was reached across the periodic boundary), which is the nearest image
// wherever the image decides the test (see list_rows_item).
__device__ __forceinline__ void list_lean_test(int i, int j, const double4& pi, const double4& pj,
                                               double s0, double s1, double s2, double skin,
                                               int* row, int& nfull, int& nhalf) {
    const double d0 = (pi.x - pj.x) - s0, d1 = (pi.y - pj.y) - s1, d2 = (pi.z - pj.z) - s2;
    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    const double rad_sum = pi.w + pj.w + skin;
    if (rsq < rad_sum * rad_sum && j != i) list_insert(i, j, row, nfull, nhalf);
}

// The candidates [a, b) of one bin piece, two per iteration.
__device__ __forceinline__ void list_lean_range(int a, int b, int i, const double4& pi, double s0,
                                                double s1, double s2, const int* particles,
                                                const double4* bx, double skin, int* row,
                                                int& nfull, int& nhalf) {
    int p = a;
    for (; p + 1 < b; p += 2) {
        const int j0 = particles[p], j1 = particles[p + 1];
        const double4 x0 = bx[p], x1 = bx[p + 1];
        list_lean_test(i, j0, pi, x0, s0, s1, s2, skin, row, nfull, nhalf);
        list_lean_test(i, j1, pi, x1, s0, s1, s2, skin, row, nfull, nhalf);
    }
    if (p < b) list_lean_test(i, particles[p], pi, bx[p], s0, s1, s2, skin, row, nfull, nhalf);
}

// One bin-order range [a, b) of candidates for atom i, read four at a time,
// each tested by list_test with image mask `img`.
__device__ __forceinline__ void list_range(int a, int b, int i, const double4& pi,
                                           const int* particles, const double4* bx, int img,
                                           const double prd[3], const double inv[3],
                                           const bool fast[3], double skin, int* row, int& nfull,
                                           int& nhalf) {
    for (int p0 = a; p0 < b; p0 += 4) {
        int jq[4];
        double4 xq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int p = p0 + q < b ? p0 + q : b - 1;
            jq[q] = particles[p];
            xq[q] = bx[p];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (p0 + q < b)
                list_test(i, jq[q], pi, xq[q], img, prd, inv, fast, skin, row, nfull, nhalf);
    }
}

// The stencil walk of NeighborList.build_neighbor_list (neighbor.py:315-369)
// as a set: for each of the 9 (dy, dz) bin rows, the x-adjacent bins are one
// contiguous range of bin-ordered candidates (two where the row wraps a
// periodic x boundary). The row is kept sorted in LDS and written, slot-major,
// to rows[k * n + i] (k < LIST_ROW_CAP); full and half counts go to
// full_count / half_count, flags[0] is raised past max_nb half entries and
// flags[1] when a row exceeds LIST_ROW_CAP (k_list_fill then writes the rows
// directly).
//
// Threads walk the atoms in bin order: a wave's atoms then sit in
// neighbouring bins of one bin row, share their 9 candidate rows and read
// overlapping candidate ranges (about 2.5x faster than atom order at 1M
// atoms). The row written for atom i does not depend on the order. `row` is
// the thread's column of the block's LDS rows.
__device__ __forceinline__ void list_rows_item(int p, int* row, const double4* xr, int n,
                                               const Ctl* ctl, const GridP& G,
                                               const int* cell_start, const int* particles,
                                               const double4* bx, double skin, int max_nb,
                                               int* full_count, int* half_count, int* flags,
                                               int* rows, int wrapped) {
    const int i = particles[p];
    const DevBox& B = ctl->box;
    const double4 pi = bx[p];  // = xr[i]
    int c[3];
    cell_coord(B, G, pi, c);
    const int gx = G.g[0], gy = G.g[1], gz = G.g[2], gxy = gx * gy;
    const int per[3] = {B.per[0], B.per[1], B.per[2]};
    const double prd[3] = {B.prd[0], B.prd[1], B.prd[2]};
    const double inv[3] = {1.0 / prd[0], 1.0 / prd[1], 1.0 / prd[2]};
    const bool fast[3] = {G.g[0] >= 3, G.g[1] >= 3, G.g[2] >= 3};
    // x-adjacent bins [x0, x1] of every row, plus a wrapped piece [w, w]
    int x0 = c[0] - 1, x1 = c[0] + 1, w = -1;
    if (gx == 1) {
        x0 = 0; x1 = 0;
    } else if (per[0] == 1) {
        // periodic with at least 3 bins (setup_grid collapses a periodic
        // 2-bin axis to 1)
        if (x0 < 0) { x0 = 0; w = gx - 1; }
        if (x1 >= gx) { x1 = gx - 1; w = 0; }
    } else {
        x0 = max(x0, 0);
        x1 = min(x1, gx - 1);
    }
    // per axis: whether an unwrapped bin still needs the nearest image (every
    // periodic axis where the positions may lie outside the box)
    const int img_base = (per[0] == 1 && (gx < 3 || !wrapped) ? 1 : 0) |
                         (per[1] == 1 && (gy < 3 || !wrapped) ? 2 : 0) |
                         (per[2] == 1 && (gz < 3 || !wrapped) ? 4 : 0);
    // The 9 (dy, dz) rows in the stencil's order: the bin-row base of each
    // (-1: no such row) and which axes its bins reached by wrapping, then the
    // main x-piece bounds of all of them, requested before any candidate is
    // read, so a row costs one memory round trip instead of two.
    int rbase[9], rimg[9], ra[9], rbnd[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int dz = r / 3 - 1, dy = r % 3 - 1;
        rimg[r] = img_base;
        bool skip = (per[2] == 1 && gz == 1 && dz != 0) || (per[1] == 1 && gy == 1 && dy != 0);
        int nz = c[2] + dz, ny = c[1] + dy;
        if (per[2] == 1 && (nz < 0 || nz >= gz)) {
            nz = nz < 0 ? nz + gz : nz - gz;
            rimg[r] |= 4;
        }
        if (per[1] == 1 && (ny < 0 || ny >= gy)) {
            ny = ny < 0 ? ny + gy : ny - gy;
            rimg[r] |= 2;
        }
        if (nz < 0 || nz >= gz || ny < 0 || ny >= gy) skip = true;
        rbase[r] = skip ? -1 : ny * gx + nz * gxy;
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        ra[r] = rbase[r] < 0 ? 0 : cell_start[rbase[r] + x0];
        rbnd[r] = rbase[r] < 0 ? 0 : cell_start[rbase[r] + x1 + 1];
    }
    int nfull = 0, nhalf = 0;
    // With every position inside the box and at least 3 bins on every
    // periodic axis, a candidate's nearest image on an axis is known from how
    // its bin was reached. Without crossing the boundary it is the plain
    // difference (argued at list_test). Across it, |d| < prd and the sign of
    // d is fixed, so round(d / prd) is k = +-1 (the crossing's direction) or
    // 0; with k the reference's d - prd * k is exactly d - (+-prd), and where
    // it is 0 both the image and d - (+-prd) exceed a bin width, so the pair
    // is out either way. One subtraction per axis then replaces the rounding.
    if (list_lean(B, G, wrapped)) {
        const double sw = w == gx - 1 ? -prd[0] : prd[0];  // the x piece across the boundary
#pragma unroll 1
        for (int r = 0; r < 9; ++r) {
            if (rbase[r] < 0) continue;
            const int ny = c[1] + (r % 3 - 1), nz = c[2] + (r / 3 - 1);
            const double s1 = per[1] != 1 ? 0.0 : (ny < 0 ? -prd[1] : (ny >= gy ? prd[1] : 0.0));
            const double s2 = per[2] != 1 ? 0.0 : (nz < 0 ? -prd[2] : (nz >= gz ? prd[2] : 0.0));
            list_lean_range(ra[r], rbnd[r], i, pi, 0.0, s1, s2, particles, bx, skin, row, nfull,
                            nhalf);
            if (w >= 0)
                list_lean_range(cell_start[rbase[r] + w], cell_start[rbase[r] + w + 1], i, pi, sw,
                                s1, s2, particles, bx, skin, row, nfull, nhalf);
        }
    } else {
        const int ximg = per[0] == 1 ? 1 : 0;
#pragma unroll 1
        for (int r = 0; r < 9; ++r) {
            if (rbase[r] < 0) continue;
            list_range(ra[r], rbnd[r], i, pi, particles, bx, rimg[r], prd, inv, fast, skin, row,
                       nfull, nhalf);
            if (w >= 0)
                list_range(cell_start[rbase[r] + w], cell_start[rbase[r] + w + 1], i, pi,
                           particles, bx, rimg[r] | ximg, prd, inv, fast, skin, row, nfull,
                           nhalf);
        }
    }
    full_count[i] = nfull;
    half_count[i] = nhalf;
    if (nhalf > max_nb) flags[0] = 1;
    if (nfull > LIST_ROW_CAP) {
        flags[1] = 1;
        return;
    }
    for (int k = 0; k < nfull; ++k) rows[(size_t)k * n + i] = row[k * TMPS_BLOCK];
}

// lean_skip: launched after k_list_rows_tile, which builds the list itself
// where the lean walk applies
__global__ void __launch_bounds__(TMPS_BLOCK) k_list_rows(const double4* xr, int n, const Ctl* ctl,
                                                          GridP G, const GridP* Gd,
                                                          const int* cell_start,
                                                          const int* particles, const double4* bx,
                                                          double skin, int max_nb, int* full_count,
                                                          int* half_count, int* flags, int* rows,
                                                          int wrapped, int lean_skip) {
    __shared__ int lrow[LIST_ROW_CAP * TMPS_BLOCK];
    if (Gd && ctl->error) return;
    if (lean_skip && list_lean(ctl->box, Gd ? *Gd : G, wrapped)) return;
    // consecutive blocks read overlapping candidate ranges: one XCD (and its
    // L2) takes a contiguous run of them
    const int p = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p < n)
        list_rows_item(p, lrow + threadIdx.x, xr, n, ctl, Gd ? *Gd : G, cell_start, particles,
                       bx, skin, max_nb, full_count, half_count, flags, rows, wrapped);
}

// Candidates of k_list_rows_tile: the lean test, an accepted j appended to
// atom i's scratch row rows[k * n + i] (unsorted; k_list_compact sorts).
__device__ __forceinline__ void tile_test(int i, int j, const double4& pi, const double4& pj,
                                          double s0, double s1, double s2, double skin, int* rows,
                                          int n, int& nfull, int& nhalf) {
    const double d0 = (pi.x - pj.x) - s0, d1 = (pi.y - pj.y) - s1, d2 = (pi.z - pj.z) - s2;
    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    const double rad_sum = pi.w + pj.w + skin;
    if (rsq < rad_sum * rad_sum && j != i) {
        if (nfull < LIST_ROW_CAP) rows[(size_t)nfull * n + i] = j;
        ++nfull;
        if (j > i) ++nhalf;
    }
}

// The candidates [a, b) of a bin piece, from the LDS copy (sj, sx) or from
// global memory (particles, bx).
__device__ __forceinline__ void tile_range(int a, int b, int i, const double4& pi, double s0,
                                           double s1, double s2, const int* js,
                                           const double4* xs, double skin, int* rows, int n,
                                           int& nfull, int& nhalf) {
    int q = a;
    for (; q + 1 < b; q += 2) {
        const int j0 = js[q], j1 = js[q + 1];
        const double4 x0 = xs[q], x1 = xs[q + 1];
        tile_test(i, j0, pi, x0, s0, s1, s2, skin, rows, n, nfull, nhalf);
        tile_test(i, j1, pi, x1, s0, s1, s2, skin, rows, n, nfull, nhalf);
    }
    if (q < b) tile_test(i, js[q], pi, xs[q], s0, s1, s2, skin, rows, n, nfull, nhalf);
}

// bin-order candidates of one stencil plane a block stages in LDS
#define LIST_TILE_CAP 1024

// k_list_rows where the lean walk applies (list_lean; otherwise it returns
// and k_list_rows, launched after it, builds the list). The block's atoms
// are consecutive in bin order, so for each plane of the stencil (dz) the
// candidates of all their rows lie in one contiguous bin-order range: the
// block copies it into LDS once, with coalesced loads, and its threads test
// their rows against the copy instead of gathering every candidate from
// memory themselves. A plane whose range exceeds LIST_TILE_CAP (a block
// whose rows wrap a periodic boundary in y) is read from memory, as are the
// single bins reached across the x boundary. Rows go to the scratch
// unsorted; the accepted set is list_rows_item's.
__global__ void __launch_bounds__(TMPS_BLOCK) k_list_rows_tile(
    int n, const Ctl* ctl, GridP G, const GridP* Gd, const int* cell_start, const int* particles,
    const double4* bx, double skin, int max_nb, int* full_count, int* half_count, int* flags,
    int* rows, int wrapped) {
    __shared__ int sj[LIST_TILE_CAP];
    __shared__ double4 sx[LIST_TILE_CAP];
    __shared__ int wred[TMPS_BLOCK / 64][6];
    __shared__ int hull[6];
    if (Gd && ctl->error) return;
    const GridP g = Gd ? *Gd : G;
    const DevBox& B = ctl->box;
    if (!list_lean(B, g, wrapped)) return;
    const int p = xcd_block(blockIdx.x, gridDim.x) * TMPS_BLOCK + threadIdx.x;
    const bool act = p < n;
    const int i = particles[act ? p : n - 1];
    const double4 pi = bx[act ? p : n - 1];  // = xr[i]
    int c[3];
    cell_coord(B, g, pi, c);
    const int gx = g.g[0], gy = g.g[1], gz = g.g[2], gxy = gx * gy;
    const int per[3] = {B.per[0], B.per[1], B.per[2]};
    const double prd[3] = {B.prd[0], B.prd[1], B.prd[2]};
    // x-adjacent bins [x0, x1] of every row, plus a wrapped piece [w, w]
    int x0 = c[0] - 1, x1 = c[0] + 1, w = -1;
    if (per[0] == 1) {
        if (x0 < 0) { x0 = 0; w = gx - 1; }
        if (x1 >= gx) { x1 = gx - 1; w = 0; }
    } else {
        x0 = max(x0, 0);
        x1 = min(x1, gx - 1);
    }
    int rbase[9], ra[9], rb[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        int nz = c[2] + r / 3 - 1, ny = c[1] + r % 3 - 1;
        if (per[2] == 1) nz = nz < 0 ? nz + gz : (nz >= gz ? nz - gz : nz);
        if (per[1] == 1) ny = ny < 0 ? ny + gy : (ny >= gy ? ny - gy : ny);
        rbase[r] = (nz < 0 || nz >= gz || ny < 0 || ny >= gy) ? -1 : ny * gx + nz * gxy;
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        ra[r] = rbase[r] < 0 ? 0 : cell_start[rbase[r] + x0];
        rb[r] = rbase[r] < 0 ? 0 : cell_start[rbase[r] + x1 + 1];
    }
    // each plane's candidate range over the block: [min ra, max rb)
    int h[6];
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
        h[pl] = INT_MAX;
        h[3 + pl] = INT_MIN;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const int r = pl * 3 + dy;
            if (act && rbase[r] >= 0 && rb[r] > ra[r]) {
                h[pl] = min(h[pl], ra[r]);
                h[3 + pl] = max(h[3 + pl], rb[r]);
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int o = __shfl_xor(h[k], off);
            h[k] = k < 3 ? min(h[k], o) : max(h[k], o);
        }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) wred[threadIdx.x >> 6][k] = h[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        int v = wred[0][k];
        for (int wv = 1; wv < TMPS_BLOCK / 64; ++wv) v = k < 3 ? min(v, wred[wv][k]) : max(v, wred[wv][k]);
        hull[k] = v;
    }
    __syncthreads();
    int nfull = 0, nhalf = 0;
    const double sw = w == gx - 1 ? -prd[0] : prd[0];  // the x piece across the boundary
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
        const int lo = hull[pl], hi = hull[3 + pl];
        const bool staged = hi > lo && hi - lo <= LIST_TILE_CAP;
        if (staged)
            for (int q = threadIdx.x; q < hi - lo; q += TMPS_BLOCK) {
                sj[q] = particles[lo + q];
                sx[q] = bx[lo + q];
            }
        __syncthreads();
        if (act) {
            const int nz = c[2] + pl - 1;
            const double s2 = per[2] != 1 ? 0.0 : (nz < 0 ? -prd[2] : (nz >= gz ? prd[2] : 0.0));
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int r = pl * 3 + dy;
                if (rbase[r] < 0) continue;
                const int ny = c[1] + dy - 1;
                const double s1 =
                    per[1] != 1 ? 0.0 : (ny < 0 ? -prd[1] : (ny >= gy ? prd[1] : 0.0));
                if (staged)
                    tile_range(ra[r] - lo, rb[r] - lo, i, pi, 0.0, s1, s2, sj, sx, skin, rows, n,
                               nfull, nhalf);
                else
                    tile_range(ra[r], rb[r], i, pi, 0.0, s1, s2, particles, bx, skin, rows, n,
                               nfull, nhalf);
                if (w >= 0)
                    tile_range(cell_start[rbase[r] + w], cell_start[rbase[r] + w + 1], i, pi, sw,
                               s1, s2, particles, bx, skin, rows, n, nfull, nhalf);
            }
        }
        __syncthreads();  // the next plane overwrites the copy
    }
    if (!act) return;
    full_count[i] = nfull;
    half_count[i] = nhalf;
    if (nhalf > max_nb) flags[0] = 1;
    if (nfull > LIST_ROW_CAP) flags[1] = 1;
}

// Rows from the slot-major scratch of k_list_rows into the CSR list.
__device__ __forceinline__ void list_compact_atom(int i, int n, const int* rows,
                                                  const int* nb_start, int* nb_j) {
    const int a = nb_start[i], m = nb_start[i + 1] - a;
    for (int k = 0; k < m; ++k) nb_j[a + k] = rows[(size_t)k * n + i];
}

// The same, each row sorted ascending in j on the way (k_list_rows_tile
// appends a row's entries in the order it meets them; k_list_rows' rows are
// sorted already). flags: a build decided on the device, which launches this
// kernel whatever the row lengths: nothing to do where a row exceeded the
// scratch (k_list_fill writes the list then)
__global__ void __launch_bounds__(TMPS_BLOCK) k_list_compact(int n, const int* rows,
                                                             const int* nb_start, int* nb_j,
                                                             const int* flags) {
    __shared__ int lrow[LIST_ROW_CAP * TMPS_BLOCK];
    if (flags && flags[1]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = nb_start[i], m = nb_start[i + 1] - a;  // at most LIST_ROW_CAP here
    int* row = lrow + threadIdx.x;
    for (int k = 0; k < m; ++k) column_insert(row, k, rows[(size_t)k * n + i]);
    for (int k = 0; k < m; ++k) nb_j[a + k] = row[k * TMPS_BLOCK];
}

// ContactHistory save/restore (contact_history.py:91-139) as a merge of the
// atom's old and new rows, both ascending in j: a new slot inherits the
// touching flag and shear of the same partner in the old row, if it was
// touching there.
template <typename H>
__device__ __forceinline__ void hist_restore_atom(int i, const int* nb_start, const int* nb_j,
                                                  uint8_t* nb_flag, H* nb_hist,
                                                  const int* onb_start, const int* onb_j,
                                                  const uint8_t* onb_flag, const H* onb_hist,
                                                  int old_n) {
    const int s0 = nb_start[i], s1 = nb_start[i + 1];
    int o = 0, o1 = 0;
    if (i < old_n) {
        o = onb_start[i];
        o1 = onb_start[i + 1];
    }
    for (int s = s0; s < s1; ++s) {
        const int j = nb_j[s];
        while (o < o1 && onb_j[o] < j) ++o;
        if (o < o1 && onb_j[o] == j && onb_flag[o]) {
            nb_flag[s] = 1;
            nb_hist[s] = onb_hist[o];
        } else {
            nb_flag[s] = 0;
        }
    }
}

template <typename H>
__global__ void k_hist_restore(int n, const int* nb_start, const int* nb_j, uint8_t* nb_flag,
                               H* nb_hist, const int* onb_start, const int* onb_j,
                               const uint8_t* onb_flag, const H* onb_hist, int old_n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        hist_restore_atom<H>(i, nb_start, nb_j, nb_flag, nb_hist, onb_start, onb_j, onb_flag,
                             onb_hist, old_n);
}

template <typename H>
__global__ void k_hist_clear(int nslots, uint8_t* nb_flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslots) nb_flag[s] = 0;
}

__global__ void k_store_x0(const double4* xr, double4* x0, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x0[i] = xr[i];
}

// Rows' half entries (j > i) in the reference's flat pair layout.
template <typename H>
__global__ void k_half_pairs(int n, const int* nb_start, const int* nb_j, const uint8_t* nb_flag,
                             const H* nb_hist, const int* tag, const int* half_start, int* pi,
                             int* pj, int* partner, double* shear) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int h = half_start[i];
    for (int s = nb_start[i]; s < nb_start[i + 1]; ++s) {
        const int j = nb_j[s];
        if (j <= i) continue;
        pi[h] = i;
        pj[h] = j;
        const bool on = nb_flag[s] != 0;
        partner[h] = on ? tag[j] : -1;
        const H hv = nb_hist[s];
        shear[3 * h + 0] = on ? (double)hv.x : 0.0;
        shear[3 * h + 1] = on ? (double)hv.y : 0.0;
        shear[3 * h + 2] = on ? (double)hv.z : 0.0;
        ++h;
    }
}

// Seed both copies of every slot from a flat half-list history
// (partner tag or -1, shear), as the pair kernels would read it.
template <typename H>
__global__ void k_set_half(int n, const int* nb_start, const int* nb_j, uint8_t* nb_flag, H* nb_hist,
                           const int* tag, const int* half_start, const int* partner,
                           const double* shear) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int s = nb_start[i]; s < nb_start[i + 1]; ++s) {
        const int j = nb_j[s];
        const int a = i < j ? i : j, b = i < j ? j : i;
        // position of b among a's entries greater than a
        const int r0 = nb_start[a], r1 = nb_start[a + 1];
        int lo = r0, hi = r1;
        while (lo < hi) {  // first entry > a
            const int mid = (lo + hi) >> 1;
            if (nb_j[mid] > a) hi = mid; else lo = mid + 1;
        }
        const int first_gt = lo;
        lo = first_gt; hi = r1;
        while (lo < hi) {  // entry == b
            const int mid = (lo + hi) >> 1;
            if (nb_j[mid] < b) lo = mid + 1; else hi = mid;
        }
        const int h = half_start[a] + (lo - first_gt);
        const bool on = partner[h] == tag[b];
        nb_flag[s] = on ? 1 : 0;
        H hv;
        hv.x = shear[3 * h + 0];
        hv.y = shear[3 * h + 1];
        hv.z = shear[3 * h + 2];
        hv.w = 0;
        nb_hist[s] = hv;
    }
}

__global__ void k_half_count(int n, const int* nb_start, const int* nb_j, int* half_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = 0;
    for (int s = nb_start[i]; s < nb_start[i + 1]; ++s)
        if (nb_j[s] > i) ++c;
    half_count[i] = c;
}

// ------------------------------------------------ batched redo (batch.hip)
//
// The rebuild of many ensemble members at once: each kernel below runs the
// per-item body of its single-system counterpart above for every entry in a
// block map of (entry, local block) pairs, so each member's bins, list and
// history are exactly those of its own rebuild.

#define REDO_ITEM(map)                         \
    const int2 mb = (map)[blockIdx.x];         \
    const RedoEntry& e = T[mb.x];              \
    const int i = mb.y * TMPS_BLOCK + threadIdx.x;

__global__ void kb_list_save(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    list_save_items<float4>(i, e.save_blocks * TMPS_BLOCK, e.n, e.nb_start, e.nb_j, e.nb_flag,
                            e.nb_hist, e.onb_start, e.onb_j, e.onb_flag, e.onb_hist, e.onb_cap);
}

__global__ void kb_pbc(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i < e.n) pbc_atom(i, e.xr, e.vm, e.ctl);
}

// bin counts and fill cursors, and the list build's flags (k_zero_int3)
__global__ void kb_bins_zero(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i < e.ncell + 1) {
        e.cell_count[i] = 0;
        e.cell_fill[i] = 0;
    }
    if (i < 4) e.flags[i] = 0;
}

__global__ void kb_bin_count(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i < e.n) bin_count_atom(i, e.xr, e.ctl, e.grid, e.cell_of, e.cell_count);
}

// The scans of the build: bin counts -> bin starts (ATOMS 0), the list's row
// lengths -> row starts, total into flags[2] (ATOMS 1).
template <int ATOMS>
__device__ __forceinline__ void redo_scan_arrays(const RedoEntry& e, const int*& in, int*& out,
                                                 int& n) {
    in = ATOMS ? e.full_count : e.cell_count;
    out = ATOMS ? e.nb_start : e.cell_start;
    n = ATOMS ? e.n : e.ncell;
}

template <int ATOMS>
__global__ void kb_scan_tiles(const RedoEntry* T, const int2* map) {
    const int2 mb = map[blockIdx.x];
    const RedoEntry& e = T[mb.x];
    const int* in;
    int* out;
    int n;
    redo_scan_arrays<ATOMS>(e, in, out, n);
    scan_tile(mb.y, in, out, n, e.scan_tmp);
}

// one block per entry
template <int ATOMS>
__global__ void kb_scan_sums(const RedoEntry* T) {
    const RedoEntry& e = T[blockIdx.x];
    const int* in;
    int* out;
    int n;
    redo_scan_arrays<ATOMS>(e, in, out, n);
    scan_sums(e.scan_tmp, (n + SCAN_TILE - 1) / SCAN_TILE, out + n, ATOMS ? e.flags + 2 : nullptr);
}

template <int ATOMS>
__global__ void kb_scan_add(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    const int* in;
    int* out;
    int n;
    redo_scan_arrays<ATOMS>(e, in, out, n);
    if (i < n) out[i] += e.scan_tmp[i / SCAN_TILE];
}

__global__ void kb_bin_fill(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i < e.n) bin_fill_atom(i, e.cell_of, e.cell_start, e.cell_fill, e.cell_particles);
}

__global__ void kb_bin_sort(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i < e.ncell) bin_sort_cell(i, e.cell_start, e.cell_particles);
}

__global__ void kb_bin_gather(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i < e.n) e.bin_x[i] = e.xr[e.cell_particles[i]];
}

__global__ void __launch_bounds__(TMPS_BLOCK) kb_list_rows(const RedoEntry* T, const int2* map) {
    __shared__ int lrow[LIST_ROW_CAP * TMPS_BLOCK];
    REDO_ITEM(map)
    if (i < e.n)
        // the batched redo wraps the positions (kb_pbc) before its build
        list_rows_item(i, lrow + threadIdx.x, e.xr, e.n, e.ctl, e.grid, e.cell_start,
                       e.cell_particles, e.bin_x, e.skin, e.max_nb, e.full_count, e.half_count,
                       e.flags, e.row_tmp, 1);
}

// every entry's build flags (overflow, long row, slot count) into one array
__global__ void kb_list_flags(const RedoEntry* T, int entries, int4* out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= entries) return;
    const int* f = T[q].flags;
    out[q] = make_int4(f[0], f[1], f[2], 0);
}

// the CSR rows (from the scratch rows, or by the two-pass fill), and x0
__global__ void kb_list_finish(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i >= e.n) return;
    if (e.long_rows)
        list_fill_atom(i, e.xr, e.ctl, e.grid, e.cell_start, e.cell_particles, e.skin, e.nb_start,
                       e.nb_j);
    else
        list_compact_atom(i, e.n, e.row_tmp, e.nb_start, e.nb_j);
    e.x0[i] = e.xr[i];
}

__global__ void kb_hist_restore(const RedoEntry* T, const int2* map) {
    REDO_ITEM(map)
    if (i < e.n)
        hist_restore_atom<float4>(i, e.nb_start, e.nb_j, e.nb_flag, e.nb_hist, e.onb_start,
                                  e.onb_j, e.onb_flag, e.onb_hist, e.old_n);
}

}  // namespace

// ================================================================ launchers

// ------------------------------------- builds decided on the device
//
// The list build of a device-resident run's rebuild redo reads everything it
// needs from the control block and the list's buffers, so the host enqueues
// it without a round trip: the grid from the box of the step, the cell
// counts and the slot count stay on the device. Where the device cannot
// serve the build (a grid beyond the bin arrays, a list beyond its buffers)
// the control block records TMPS_NEED_HOST and the abort, the new list is
// left empty so that every later kernel of the redo stays in bounds, and the
// host redoes the build once it reads the run's status.

// NeighborList.setup_grid (neighbor.py:143-189) for the list's own grid, on
// the device from the control block's box: the host's setup_grid arithmetic
// in the same order (this unit rounds as written too).
__global__ void k_list_grid(Ctl* ctl, double max_cutoff, double skin, int max_cells, int cell_cap,
                            int split) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || ctl->error) return;
    const DevBox& B = ctl->box;
    double cut = max_cutoff + skin;
    if (cut <= 0) cut = 1.0;
    const double xprd = B.prd[0], yprd = B.prd[1], zprd = B.prd[2];
    const double vol = xprd * yprd * zprd;
    const double c0 = yprd * zprd - 0.0 * 0.0;
    const double ca = zprd * xprd - 0.0 * 0.0;
    const double c2 = xprd * yprd - 0.0 * 0.0;
    const double w0 = vol / sqrt(c0 * c0);
    const double w1 = vol / sqrt(ca * ca);
    const double w2 = vol / sqrt(c2 * c2);
    int gx = max(1, (int)floor(w0 / cut));
    int gy = max(1, (int)floor(w1 / cut));
    int gz = max(1, (int)floor(w2 / cut));
    if (B.per[0] == 1 && gx == 2) gx = 1;
    if (B.per[1] == 1 && gy == 2) gy = 1;
    if (B.per[2] == 1 && gz == 2) gz = 1;
    const long long total = (long long)gx * gy * gz;
    if (total > max_cells) {
        const double factor = pow((double)max_cells / (double)total, 1.0 / 3.0);
        gx = max(1, (int)(gx * factor));
        gy = max(1, (int)(gy * factor));
        gz = max(1, (int)(gz * factor));
        if (B.per[0] == 1 && gx == 2) gx = 1;
        if (B.per[1] == 1 && gy == 2) gy = 1;
        if (B.per[2] == 1 && gz == 2) gz = 1;
    }
    if (split && (gx > 2047 || gy > 2047 || gz > 1023)) {
        ctl->error = TMPS_ERR_UNSUPPORTED;
        ctl->abort = 1;
        return;
    }
    if ((long long)gx * gy * gz + 1 > cell_cap) {
        ctl->error = TMPS_NEED_HOST;
        ctl->abort = 1;
        return;
    }
    GridP G;
    G.g[0] = gx;
    G.g[1] = gy;
    G.g[2] = gz;
    G.csize[0] = 1.0 / gx;
    G.csize[1] = 1.0 / gy;
    G.csize[2] = 1.0 / gz;
    G.ncell = gx * gy * gz;
    ctl->lgrid = G;
}

// Bin counts and fill cursors over the device grid's cells (and one past),
// and the list build's flags.
__global__ void k_zero_cells_dev(const Ctl* ctl, int* count, int* fill, int* flags) {
    if (ctl->error) return;
    const int m = ctl->lgrid.ncell + 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        count[i] = 0;
        fill[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) flags[threadIdx.x] = 0;
}

// The exclusive scan of the bin counts over the device grid's cells: tiles
// launched for the bin arrays' capacity, those past the cell count idle.
__global__ void k_scan_tiles_dev(const Ctl* ctl, const int* in, int* out, int* tile_sums) {
    if (ctl->error) return;
    const int m = ctl->lgrid.ncell;
    if (blockIdx.x * SCAN_TILE >= m) {
        if (threadIdx.x == 0) tile_sums[blockIdx.x] = 0;
        return;
    }
    scan_tile(blockIdx.x, in, out, m, tile_sums);
}

__global__ void k_scan_sums_dev(const Ctl* ctl, int* tile_sums, int ntiles, int* out) {
    if (ctl->error) return;
    const int m = ctl->lgrid.ncell;
    scan_sums(tile_sums, min(ntiles, (m + SCAN_TILE - 1) / SCAN_TILE), out + m, nullptr);
}

__global__ void k_scan_add_dev(const Ctl* ctl, int* out, const int* tile_sums) {
    if (ctl->error) return;
    const int m = ctl->lgrid.ncell;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        out[i] += tile_sums[i / SCAN_TILE];
}

__global__ void k_bin_fill_dev(const Ctl* ctl, int n, const int* cell_of, const int* start,
                               int* fill, int* particles) {
    if (ctl->error) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bin_fill_atom(i, cell_of, start, fill, particles);
}

__global__ void k_bin_sort_dev(const Ctl* ctl, const int* start, int* particles) {
    if (ctl->error) return;
    const int m = ctl->lgrid.ncell;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x)
        bin_sort_cell(c, start, particles);
}

// After the rows' scan: NeighborList.build's overflow (neighbor.py:426-431)
// and the list's buffers checked; on success the slot count goes into the
// control block, on failure the list is emptied and the error recorded. The
// decision is the same in every thread: a thread that reads the error
// another one recorded decides to empty the list as that one did.
__global__ void k_list_check(Ctl* ctl, const int* flags, int* nb_start, int n, long long cap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int err = ctl->error;
    if (!err && flags[0]) err = TMPS_ERR_OVERFLOW;
    if (!err && (long long)flags[2] > cap) err = TMPS_NEED_HOST;
    if (err) {
        if (i <= n) nb_start[i] = 0;
        if (i == 0) {
            ctl->error = err;
            ctl->abort = 1;
        }
    } else if (i == 0) {
        ctl->nslots = flags[2];
    }
}

// ------------------------------------------------ bins by a radix sort
//
// NeighborList.build_grid's counting sort (neighbor.py:247-282) as a stable
// least-significant-digit radix sort of the atoms' cells, 8 bits a pass:
// per tile of RS_TILE atoms a digit histogram in LDS, an exclusive scan over
// (digit, tile), and a stable scatter whose ranks come from wavefront ballots.
// Stability keeps each bin's atoms in ascending index, the order the atomic
// counting sort reaches with its per-bin insertion sort, so the bins are the
// same; no global atomics are needed. Gd: a build decided on the device
// (its grid in the control block, nothing done on a recorded error).

#define RS_DIGITS 256
#define RS_ROUNDS 16
#define RS_TILE (TMPS_BLOCK * RS_ROUNDS)
static_assert(TMPS_BLOCK == RS_DIGITS, "one thread per digit of a tile's histogram");

// The cell of every atom (the sort key, into keys) and each tile's
// histogram of the first digit, digit-major: hist[d * ntiles + tile].
__global__ void k_rs_keys(const double4* xr, int n, const Ctl* ctl, GridP G, const GridP* Gd,
                          int* keys, int* hist, int ntiles) {
    if (Gd && ctl->error) return;
    __shared__ int h[RS_DIGITS];
    const GridP g = Gd ? *Gd : G;
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int i = base + r * TMPS_BLOCK + threadIdx.x;
        if (i < n) {
            int c[3];
            cell_coord(ctl->box, g, xr[i], c);
            const int key = c[0] + c[1] * g.g[0] + c[2] * g.g[0] * g.g[1];
            keys[i] = key;
            atomicAdd(&h[key & (RS_DIGITS - 1)], 1);
        }
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Each tile's histogram of the digit at `shift`.
__global__ void k_rs_hist(const Ctl* ctl, const GridP* Gd, const int* keys, int n, int shift,
                          int* hist, int ntiles) {
    if (Gd && ctl->error) return;
    __shared__ int h[RS_DIGITS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int i = base + r * TMPS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (RS_DIGITS - 1)], 1);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// The stable scatter of one pass. A tile goes in RS_ROUNDS rounds of
// TMPS_BLOCK consecutive keys; in a round each lane finds the lanes of its
// wavefront holding the same digit (eight ballots), its rank among them, and
// the counts of the earlier wavefronts, so keys of one digit keep their order
// within the round, rounds follow each other, and offs (the scanned
// histograms) orders the tiles. vin null: the values are the keys' indices.
__global__ void k_rs_scatter(const Ctl* ctl, const GridP* Gd, const int* kin, const int* vin,
                             int* kout, int* vout, int n, int shift, const int* offs,
                             int ntiles) {
    if (Gd && ctl->error) return;
    constexpr int NW = TMPS_BLOCK / 64;
    __shared__ int base[RS_DIGITS];
    __shared__ int wcnt[NW][RS_DIGITS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    base[threadIdx.x] = offs[threadIdx.x * ntiles + blockIdx.x];
    for (int r = 0; r < RS_ROUNDS; ++r) {
#pragma unroll
        for (int w = 0; w < NW; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const int i = blockIdx.x * RS_TILE + r * TMPS_BLOCK + threadIdx.x;
        const bool valid = i < n;
        int key = 0, val = 0, d = 0;
        if (valid) {
            key = kin[i];
            val = vin ? vin[i] : i;
            d = (key >> shift) & (RS_DIGITS - 1);
        }
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = ((d >> b) & 1) != 0;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = base[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) add += wcnt[w][threadIdx.x];
        base[threadIdx.x] += add;
    }
}

// Bin starts from the sorted keys: the atoms of bin c sit at [start[c],
// start[c + 1]); empty bins take the next bin's start.
__global__ void k_rs_starts(const Ctl* ctl, GridP G, const GridP* Gd, const int* skeys, int n,
                            int* start) {
    if (Gd && ctl->error) return;
    const int ncell = Gd ? Gd->ncell : G.ncell;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= n; p += gridDim.x * blockDim.x) {
        const int k = p < n ? skeys[p] : ncell;
        const int kp = p > 0 ? skeys[p - 1] : -1;
        for (int c = kp + 1; c <= k; ++c) start[c] = p;
    }
}

__global__ void k_zero_flags(const Ctl* ctl, int guard, int* flags) {
    if (guard && ctl->error) return;
    if (threadIdx.x < 4) flags[threadIdx.x] = 0;
}

static inline dim3 grid_for(int n) { return dim3(nblocks_for(n > 0 ? n : 1)); }

int launch_pbc(System& s) {
    if (s.n <= 0) return 0;
    hipLaunchKernelGGL(k_pbc, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.xr, s.vm, s.n, s.ctl);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Exclusive scan of in[0, n) into out[0, n] (out[n] = total) in one block,
// tile after tile with a running carry: one launch for small arrays.
__global__ void k_scan_one_block(const int* in, int* out, int n, int* total_copy) {
    __shared__ int carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int t0 = 0; t0 < n; t0 += SCAN_TILE) {
        int v[SCAN_PER_THREAD];
        int sum = 0;
#pragma unroll
        for (int q = 0; q < SCAN_PER_THREAD; ++q) {
            const int idx = t0 + threadIdx.x * SCAN_PER_THREAD + q;
            v[q] = idx < n ? in[idx] : 0;
            sum += v[q];
        }
        int excl = 0;
        const int tile_total = block_excl_scan(sum, &excl);
        int run = carry_s + excl;
#pragma unroll
        for (int q = 0; q < SCAN_PER_THREAD; ++q) {
            const int idx = t0 + threadIdx.x * SCAN_PER_THREAD + q;
            if (idx < n) out[idx] = run;
            run += v[q];
        }
        __syncthreads();
        if (threadIdx.x == 0) carry_s += tile_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[n] = carry_s;
        if (total_copy) *total_copy = carry_s;
    }
}

// arrays up to this many entries are scanned by one block
#define SCAN_ONE_BLOCK_MAX 65536

int launch_scan(System& s, const int* in, int* out, int n, int* total_copy) {
    if (n <= SCAN_ONE_BLOCK_MAX) {
        hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(TMPS_BLOCK), 0, s.stream, in, out, n,
                           total_copy);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    const int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    const int nt = ntiles > 0 ? ntiles : 1;
    if (s.scan_tmp_cap < nt + 1) return tmps_fail(TMPS_ERR_ARG, "scan workspace too small");
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(TMPS_BLOCK), 0, s.stream, in, out, n, s.scan_tmp);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(TMPS_BLOCK), 0, s.stream, s.scan_tmp, nt, out + n,
                       total_copy);
    HIP_TRY(hipGetLastError());
    if (n > 0) {
        hipLaunchKernelGGL(k_scan_add, grid_for(n), dim3(TMPS_BLOCK), 0, s.stream, out, n, s.scan_tmp);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

int launch_bins_radix(System& s, const double4* xr, const GridP& G, const GridP* Gd,
                      int max_ncell);

int launch_bins(System& s, const double4* xr) {
    // the list build's flags
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s.stream, s.ctl, 0, s.flags);
    HIP_TRY(hipGetLastError());
    return launch_bins_radix(s, xr, s.grid, nullptr, s.grid.ncell);
}


// The radix binning (k_rs_*): cell_of holds each atom's cell, cell_start /
// cell_particles the bins. max_ncell bounds the grid's cells (the passes
// cover its bits); G / Gd as for k_bin_count.
static int rs_reserve(System& s, int ntiles) {
    const size_t nh = (size_t)RS_DIGITS * ntiles + 1;
    const size_t na = (size_t)std::max(s.max_atoms, 1);
    if (!s.rs_k1 || s.rs_na < na) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        for (int** p : {&s.rs_k1, &s.rs_k2, &s.rs_v1, &s.rs_v2}) {
            if (*p) HIP_TRY(hipFree(*p));
            HIP_TRY(hipMalloc(p, na * sizeof(int)));
        }
        s.rs_na = na;
    }
    if (!s.rs_hist || s.rs_nh < nh) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        for (int** p : {&s.rs_hist, &s.rs_offs}) {
            if (*p) HIP_TRY(hipFree(*p));
            HIP_TRY(hipMalloc(p, nh * sizeof(int)));
        }
        s.rs_nh = nh;
    }
    const int need_scan = (int)((nh + SCAN_TILE - 1) / SCAN_TILE) + 1;
    if (s.scan_tmp_cap < need_scan) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        if (s.scan_tmp) HIP_TRY(hipFree(s.scan_tmp));
        HIP_TRY(hipMalloc(&s.scan_tmp, (size_t)need_scan * sizeof(int)));
        s.scan_tmp_cap = need_scan;
    }
    return 0;
}

int bins_reserve(System& s) {
    return s.n > 0 ? rs_reserve(s, (s.n + RS_TILE - 1) / RS_TILE) : 0;
}

int launch_bins_radix(System& s, const double4* xr, const GridP& G, const GridP* Gd,
                      int max_ncell) {
    if (s.n <= 0) return 0;
    const int ntiles = (s.n + RS_TILE - 1) / RS_TILE;
    int rc = rs_reserve(s, ntiles);
    if (rc) return rc;
    int bits = 1;
    while (bits < 31 && (1LL << bits) < (long long)max_ncell) ++bits;
    const int passes = (bits + 7) / 8;
    const int nh = RS_DIGITS * ntiles;
    const int scan_tiles = (nh + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_rs_keys, dim3(ntiles), dim3(TMPS_BLOCK), 0, s.stream, xr, s.n, s.ctl, G,
                       Gd, s.cell_of, s.rs_hist, ntiles);
    HIP_TRY(hipGetLastError());
    // pass p reads (kin, vin) and writes (kout, vout); the last writes the bins
    const int* kin = s.cell_of;
    const int* vin = nullptr;
    int* kbuf[2] = {s.rs_k2, s.rs_k1};
    int* vbuf[2] = {s.rs_v1, s.rs_v2};
    for (int p = 0; p < passes; ++p) {
        if (p > 0) {
            hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(TMPS_BLOCK), 0, s.stream, s.ctl, Gd,
                               kin, s.n, 8 * p, s.rs_hist, ntiles);
            HIP_TRY(hipGetLastError());
        }
        // exclusive scan of the (digit, tile) counts, several blocks always
        hipLaunchKernelGGL(k_scan_tiles, dim3(scan_tiles), dim3(TMPS_BLOCK), 0, s.stream,
                           (const int*)s.rs_hist, s.rs_offs, nh, s.scan_tmp);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(TMPS_BLOCK), 0, s.stream, s.scan_tmp,
                           scan_tiles, s.rs_offs + nh, (int*)nullptr);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_scan_add, grid_for(nh), dim3(TMPS_BLOCK), 0, s.stream, s.rs_offs, nh,
                           (const int*)s.scan_tmp);
        HIP_TRY(hipGetLastError());
        int* kout = kbuf[p & 1];
        int* vout = p == passes - 1 ? s.cell_particles : vbuf[p & 1];
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(TMPS_BLOCK), 0, s.stream, s.ctl, Gd,
                           kin, vin, kout, vout, s.n, 8 * p, (const int*)s.rs_offs, ntiles);
        HIP_TRY(hipGetLastError());
        kin = kout;
        vin = vout;
    }
    const int sblocks = std::min(nblocks_for(s.n + 1), 2048);
    hipLaunchKernelGGL(k_rs_starts, dim3(sblocks), dim3(TMPS_BLOCK), 0, s.stream, s.ctl, G, Gd,
                       kin, s.n, s.cell_start);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int grow_list(System& s, size_t need) {
    if (need <= s.nb_cap) return 0;
    // the old list keeps its own buffers; graphs hold the current pointers
    HIP_TRY(hipStreamSynchronize(s.stream));
    size_t cap = need + need / 4 + 1024;
    const size_t hsz = s.precision == TMPS_FP64 ? sizeof(double4) : sizeof(float4);
    if (s.nb_j) HIP_TRY(hipFree(s.nb_j));
    if (s.nb_flag) HIP_TRY(hipFree(s.nb_flag));
    if (s.nb_hist) HIP_TRY(hipFree(s.nb_hist));
    if (s.nb_flag2) HIP_TRY(hipFree(s.nb_flag2));
    if (s.nb_hist2) HIP_TRY(hipFree(s.nb_hist2));
    if (s.nb_code) HIP_TRY(hipFree(s.nb_code));
    HIP_TRY(hipMalloc(&s.nb_j, cap * sizeof(int)));
    HIP_TRY(hipMalloc(&s.nb_flag, cap));
    HIP_TRY(hipMalloc(&s.nb_hist, cap * hsz));
    HIP_TRY(hipMalloc(&s.nb_flag2, cap));
    HIP_TRY(hipMalloc(&s.nb_hist2, cap * hsz));
    HIP_TRY(hipMalloc(&s.nb_code, cap * sizeof(uint16_t)));
    s.nb_cap = cap;
    drop_graphs(s);
    return 0;
}

int ensure_list_capacity(System& s, size_t need) { return grow_list(s, need); }

// Builds the full list: one pass into per-atom scratch rows, a scan, and a
// compaction (two passes, count then fill, if some row exceeds the scratch).
// *total_out receives the slot count (a host synchronisation per rebuild).
int list_reserve_rows(System& s) {
    const size_t rows_need = (size_t)std::max(s.n, 1) * LIST_ROW_CAP;
    if (rows_need > s.row_tmp_cap) {
        HIP_TRY(hipStreamSynchronize(s.stream));
        if (s.row_tmp) HIP_TRY(hipFree(s.row_tmp));
        HIP_TRY(hipMalloc(&s.row_tmp, rows_need * sizeof(int)));
        s.row_tmp_cap = rows_need;
    }
    return 0;
}

int launch_list(System& s, const double4* xr, int* overflow_out, int* total_out, bool wrapped) {
    // s.flags was cleared by launch_bins (flags[2] receives the slot count)
    int rc = list_reserve_rows(s);
    if (rc) return rc;
    if (s.n > 0) {
        hipLaunchKernelGGL(k_bin_gather, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.cell_particles, xr, s.bin_x);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_list_rows_tile, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.ctl, s.grid, nullptr, s.cell_start, s.cell_particles, s.bin_x, s.skin,
                           s.max_nb, s.full_count, s.half_count, s.flags, s.row_tmp,
                           wrapped ? 1 : 0);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_list_rows, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, xr, s.n,
                           s.ctl, s.grid, nullptr, s.cell_start, s.cell_particles, s.bin_x, s.skin,
                           s.max_nb, s.full_count, s.half_count, s.flags, s.row_tmp,
                           wrapped ? 1 : 0, 1);
        HIP_TRY(hipGetLastError());
    }
    rc = launch_scan(s, s.full_count, s.nb_start, s.n, s.flags + 2);
    if (rc) return rc;
    int hv[3];
    HIP_TRY(hipMemcpyAsync(hv, s.flags, 3 * sizeof(int), hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    *overflow_out = hv[0];
    *total_out = hv[2];
    if (hv[0]) return 0;
    rc = grow_list(s, (size_t)hv[2]);
    if (rc) return rc;
    if (s.n > 0) {
        if (!hv[1]) {
            hipLaunchKernelGGL(k_list_compact, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                               (const int*)s.row_tmp, (const int*)s.nb_start, s.nb_j, nullptr);
        } else {
            hipLaunchKernelGGL(k_list_fill, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, xr, s.n,
                               s.ctl, s.grid, nullptr, nullptr, s.cell_start, s.cell_particles,
                               s.skin, (const int*)s.nb_start, s.nb_j);
        }
        HIP_TRY(hipGetLastError());
    }
    s.nslots = hv[2];
    return 0;
}

// The list build of a rebuild redo decided on the device (see the kernels'
// comment): grid, bins, rows, the CSR list and x0, with no host round trip.
// max_cutoff: twice the largest radius; max_cells: the list grid's cap. The
// bin arrays must hold max_cells + 1 cells, the row scratch n rows.
int launch_build_dev(System& s, double max_cutoff, int max_cells, bool store_x0) {
    if (s.n <= 0) return 0;
    const int split = s.precision == TMPS_FP32_SPLIT ? 1 : 0;
    const GridP* Gd = &s.ctl->lgrid;
    const GridP none{};
    hipLaunchKernelGGL(k_list_grid, dim3(1), dim3(64), 0, s.stream, s.ctl, max_cutoff, s.skin,
                       max_cells, s.cell_cap, split);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s.stream, s.ctl, 1, s.flags);
    HIP_TRY(hipGetLastError());
    int rc = launch_bins_radix(s, s.xr, none, Gd, max_cells);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bin_gather, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                       s.cell_particles, s.xr, s.bin_x);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_list_rows_tile, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                       s.ctl, none, Gd, s.cell_start, s.cell_particles, s.bin_x, s.skin, s.max_nb,
                       s.full_count, s.half_count, s.flags, s.row_tmp, 1);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_list_rows, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.xr, s.n,
                       s.ctl, none, Gd, s.cell_start, s.cell_particles, s.bin_x, s.skin, s.max_nb,
                       s.full_count, s.half_count, s.flags, s.row_tmp, 1, 1);
    HIP_TRY(hipGetLastError());
    rc = launch_scan(s, s.full_count, s.nb_start, s.n, s.flags + 2);
    if (rc) return rc;
    hipLaunchKernelGGL(k_list_check, grid_for(s.n + 1), dim3(TMPS_BLOCK), 0, s.stream, s.ctl,
                       (const int*)s.flags, s.nb_start, s.n, (long long)s.nb_cap);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_list_compact, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                       (const int*)s.row_tmp, (const int*)s.nb_start, s.nb_j,
                       (const int*)s.flags);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_list_fill, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.xr, s.n,
                       s.ctl, none, Gd, (const int*)s.flags, s.cell_start, s.cell_particles,
                       s.skin, (const int*)s.nb_start, s.nb_j);
    HIP_TRY(hipGetLastError());
    return store_x0 ? launch_store_x0(s) : 0;
}

int launch_list_save(System& s) {
    // grid-stride over the slot count read on the device: sized for the
    // list's capacity, so the launch stays valid in a replayed graph
    const int items = (int)std::min<size_t>(std::max<size_t>((size_t)s.n + 1, s.nb_cap),
                                            (size_t)INT32_MAX);
    const int blocks = std::min((items + TMPS_BLOCK - 1) / TMPS_BLOCK, 4096);
    const int cap = (int)std::min<size_t>(s.onb_cap, (size_t)INT32_MAX);
    if (s.precision == TMPS_FP64)
        hipLaunchKernelGGL(k_list_save<double4>, dim3(blocks), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (const double4*)s.nb_hist, s.onb_start,
                           s.onb_j, s.onb_flag, (double4*)s.onb_hist, cap);
    else
        hipLaunchKernelGGL(k_list_save<float4>, dim3(blocks), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (const float4*)s.nb_hist, s.onb_start,
                           s.onb_j, s.onb_flag, (float4*)s.onb_hist, cap);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_hist_restore(System& s) {
    if (s.n <= 0) return 0;
    if (!s.have_old) return launch_hist_clear(s);
    const int old_n = s.onb_start ? s.n : 0;
    if (s.precision == TMPS_FP64)
        hipLaunchKernelGGL(k_hist_restore<double4>, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (double4*)s.nb_hist, s.onb_start, s.onb_j,
                           s.onb_flag, (const double4*)s.onb_hist, old_n);
    else
        hipLaunchKernelGGL(k_hist_restore<float4>, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (float4*)s.nb_hist, s.onb_start, s.onb_j,
                           s.onb_flag, (const float4*)s.onb_hist, old_n);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_hist_clear(System& s) {
    if (s.nslots <= 0) return 0;
    hipLaunchKernelGGL(k_hist_clear<int>, grid_for(s.nslots), dim3(TMPS_BLOCK), 0, s.stream, s.nslots,
                       s.nb_flag);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_store_x0(System& s) {
    if (s.n <= 0) return 0;
    hipLaunchKernelGGL(k_store_x0, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.xr, s.x0, s.n);
    HIP_TRY(hipGetLastError());
    return 0;
}

// d_counts: n+1 ints of workspace (half counts, scanned in place to offsets).
int launch_half_pairs(System& s, int* d_pair_i, int* d_pair_j, int* d_partner, double* d_shear,
                      int* d_counts) {
    if (s.n <= 0) return 0;
    hipLaunchKernelGGL(k_half_count, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n, s.nb_start,
                       s.nb_j, s.half_count);
    HIP_TRY(hipGetLastError());
    int rc = launch_scan(s, s.half_count, d_counts, s.n, nullptr);
    if (rc) return rc;
    if (d_pair_i == nullptr) return 0;
    if (s.precision == TMPS_FP64)
        hipLaunchKernelGGL(k_half_pairs<double4>, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (const double4*)s.nb_hist, s.tag, d_counts,
                           d_pair_i, d_pair_j, d_partner, d_shear);
    else
        hipLaunchKernelGGL(k_half_pairs<float4>, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (const float4*)s.nb_hist, s.tag, d_counts,
                           d_pair_i, d_pair_j, d_partner, d_shear);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_set_half_history(System& s, const int* d_partner, const double* d_shear,
                            const int* d_half_start) {
    if (s.n <= 0) return 0;
    if (s.precision == TMPS_FP64)
        hipLaunchKernelGGL(k_set_half<double4>, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (double4*)s.nb_hist, s.tag, d_half_start,
                           d_partner, d_shear);
    else
        hipLaunchKernelGGL(k_set_half<float4>, grid_for(s.n), dim3(TMPS_BLOCK), 0, s.stream, s.n,
                           s.nb_start, s.nb_j, s.nb_flag, (float4*)s.nb_hist, s.tag, d_half_start,
                           d_partner, d_shear);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ------------------------------------------------ batched redo (batch.hip)

#define REDO_LAUNCH(kernel, nb, ...)                                                      \
    do {                                                                                  \
        if ((nb) > 0) {                                                                   \
            hipLaunchKernelGGL(kernel, dim3(nb), dim3(TMPS_BLOCK), 0, st, __VA_ARGS__);   \
            HIP_TRY(hipGetLastError());                                                   \
        }                                                                                 \
    } while (0)

int launch_redo_save(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    REDO_LAUNCH(kb_list_save, M.n_save, T, M.save);
    return 0;
}

int launch_redo_pbc(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    REDO_LAUNCH(kb_pbc, M.n_atoms, T, M.atoms);
    return 0;
}

int launch_redo_bins(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    REDO_LAUNCH(kb_bins_zero, M.n_cells, T, M.cells);
    REDO_LAUNCH(kb_bin_count, M.n_atoms, T, M.atoms);
    REDO_LAUNCH(kb_scan_tiles<0>, M.n_cell_tiles, T, M.cell_tiles);
    REDO_LAUNCH(kb_scan_sums<0>, M.entries, T);
    REDO_LAUNCH(kb_scan_add<0>, M.n_cells, T, M.cells);
    REDO_LAUNCH(kb_bin_fill, M.n_atoms, T, M.atoms);
    REDO_LAUNCH(kb_bin_sort, M.n_cells, T, M.cells);
    return 0;
}

int launch_redo_rows(const RedoEntry* T, const RedoMaps& M, int4* flags_out, hipStream_t st) {
    REDO_LAUNCH(kb_bin_gather, M.n_atoms, T, M.atoms);
    REDO_LAUNCH(kb_list_rows, M.n_atoms, T, M.atoms);
    REDO_LAUNCH(kb_scan_tiles<1>, M.n_atom_tiles, T, M.atom_tiles);
    REDO_LAUNCH(kb_scan_sums<1>, M.entries, T);
    REDO_LAUNCH(kb_scan_add<1>, M.n_atoms, T, M.atoms);
    REDO_LAUNCH(kb_list_flags, (M.entries + TMPS_BLOCK - 1) / TMPS_BLOCK, T, M.entries, flags_out);
    return 0;
}

int launch_redo_list(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    REDO_LAUNCH(kb_list_finish, M.n_atoms, T, M.atoms);
    return 0;
}

int launch_redo_restore(const RedoEntry* T, const RedoMaps& M, hipStream_t st) {
    REDO_LAUNCH(kb_hist_restore, M.n_atoms, T, M.atoms);
    return 0;
}
