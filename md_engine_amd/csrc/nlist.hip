// Cell binning, deterministic cell-order sort and full neighbor-list build.
//
// Replaces GridGPU::periodicBoundaryConditions (src/GridGPU.cu:824-1103) and
// its kernels setBuildFlag (755-774), periodicWrap (138-161),
// countNumInGridCells (190-203), sortPerAtomArrays (253-322),
// countNumNeighbors / assignNeighbors (326-734).
//
// Design (gfx950):
//   * cells have edge >= r_list / reach (reach = 2), so a (2*reach+1)^3 stencil
//     holds every neighbor; columns of cells are ranked along a 2-D Morton
//     curve and z runs contiguously inside a column, so each stencil column is
//     one or two contiguous ranges of the sorted arrays (coalesced loads);
//   * the sort is a counting sort whose slots inside a cell are re-ordered by
//     atom id, so the layout depends only on positions, never on atomics order;
//   * the list is built in one pass: a team of T lanes per atom tests T
//     candidates per step with the canonical FP64 distance (no FMA contraction:
//     this file is compiled with -ffp-contract=off) and compacts hits with a
//     wave64 ballot; rows have a fixed capacity and the true count is kept so
//     overflow is detected and the host grows the rows and rebuilds;
//   * entries are index | tag << 30 with the 1-2/1-3/1-4 tag of the atom pair
//     (GridGPU.cu:471-520, globalDefs.h:40).
#include <algorithm>
#include <cstring>

#include "mdx_internal.h"

#pragma clang fp contract(off)

namespace mdx {

namespace {


__device__ __forceinline__ double mi_d(double d, double L, double invL, int periodic) {
    if (periodic) {
        double img = rint(d * invL);
        d = d - img * L;
    }
    return d;
}

// canonical squared distance, evaluated exactly as orc_dist2 in the oracle
__device__ __forceinline__ double dist2(const double4 &a, const double4 &b, const Geometry &g) {
    double dx = mi_d(b.x - a.x, g.L[0], g.invL[0], g.periodic[0]);
    double dy = mi_d(b.y - a.y, g.L[1], g.invL[1], g.periodic[1]);
    double dz = mi_d(b.z - a.z, g.L[2], g.invL[2], g.periodic[2]);
    return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ int cell_coord(double x, double lo, double cellInv, int nc) {
    int c = (int)floor((x - lo) * cellInv);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

// Max displacement since the last build (setBuildFlag, GridGPU.cu:755-774):
// flag when any |minImage(x - x_last)|^2 > (skin/2)^2.
__device__ __forceinline__ void check_body(int blk, int nblk, int n, const double4 *__restrict__ pos, const double4 *__restrict__ last,
                        const BoxDev *__restrict__ bx, double thr2, int *__restrict__ flags) {
    const Geometry g = bx->geo;
    int i = blk * blockDim.x + threadIdx.x;
    bool move = false;
    if (i < n) move = dist2(pos[i], last[i], g) > thr2;
    if (__any(move) && (threadIdx.x & 63) == 0) flags[0] = 1;
}

// Wrap into the box (periodicWrap) and bin into cells (countNumInGridCells).
// Charged and charge-free atoms go to two copies of the grid (slots
// [0, ncell) and [ncell, 2 ncell)), so that after the permutation the charged
// atoms come first, each class sorted by cell: a neighbor row then lists its
// charged partners first, and the charge spread and gather see only charged
// atoms.
__device__ __forceinline__ void bin_atom(int i, double4 *__restrict__ pos, const float *__restrict__ q,
                                         const BoxDev *__restrict__ bx, int ncell, const int *__restrict__ colRank,
                                         int *__restrict__ cellOf, int *__restrict__ cellCount,
                                         const double4 *__restrict__ cen, const int *__restrict__ ids) {
    const Geometry g = bx->geo;
    const CellGrid cg = bx->cg;
    double4 p = pos[i];
    double *pc = &p.x;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (g.periodic[a]) {
            double img = floor((pc[a] - g.lo[a]) * g.invL[a]);
            pc[a] = pc[a] - img * g.L[a];
        }
    }
    pos[i] = p;
    // ring centroid lists: the atom goes to the cell of its ring's centroid
    // (already in the box)
    if (cen) p = cen[ids[i]];
    int cx = cell_coord(p.x, g.lo[0], cg.cellInv[0], cg.nc[0]);
    int cy = cell_coord(p.y, g.lo[1], cg.cellInv[1], cg.nc[1]);
    int cz = cell_coord(p.z, g.lo[2], cg.cellInv[2], cg.nc[2]);
    int rank = colRank[cx * cg.nc[1] + cy] * cg.nc[2] + cz + (q[i] == 0.0f ? ncell : 0);
    cellOf[i] = rank;
    atomicAdd(&cellCount[rank], 1);
}

__device__ __forceinline__ void bin_body(int blk, int nblk, const int *__restrict__ gate, int n, double4 *__restrict__ pos, const float *__restrict__ q,
                      const BoxDev *__restrict__ bx, int ncell, const int *__restrict__ colRank,
                      int *__restrict__ cellOf, int *__restrict__ cellCount,
                      const double4 *__restrict__ cen, const int *__restrict__ ids) {
    if (gate && *gate == 0) return;
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bin_atom(i, pos, q, bx, ncell, colRank, cellOf, cellCount, cen, ids);
}

// Exclusive scan of the cell counts by one workgroup (replaces the host
// cumulativeSum round trip, helpers.h:15-24).  Each thread sums a contiguous
// segment, the segment totals are scanned across the workgroup, and each
// thread then writes its segment's prefixes.  Also seeds the fill cursors.
// The counts are cleared as they are consumed, ready for the next binning
// (no separate clearing launch in every captured step).
// one workgroup of 1024 threads
__device__ __forceinline__ void scan_wg(int ncell, int *__restrict__ count, int *__restrict__ start, int *__restrict__ fill,
                                        const BoxDev *__restrict__ box, BoxDev *__restrict__ boxLast) {
    // the box this build uses, kept for rebuilding the same list later
    constexpr int kBoxWords = (int)(sizeof(BoxDev) / sizeof(int));
    static_assert(sizeof(BoxDev) % sizeof(int) == 0, "BoxDev copied as words");
    if (boxLast)
        for (int k = threadIdx.x; k < kBoxWords; k += blockDim.x)
            reinterpret_cast<int *>(boxLast)[k] = reinterpret_cast<const int *>(box)[k];
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int seg = (ncell + 1023) / 1024;
    const int b0 = threadIdx.x * seg, b1 = min(b0 + seg, ncell);
    // a thread's segment is read eight counts at a time with the loads in
    // flight together (one dependent load per count made this kernel ~17 us)
    constexpr int U = 8;
    int v[U];
    int local = 0;
    for (int k0 = b0; k0 < b1; k0 += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = k0 + u < b1 ? count[k0 + u] : 0;
#pragma unroll
        for (int u = 0; u < U; ++u) local += v[u];
    }
    int x = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int t = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < 16) wsum[lane] = t;
    }
    __syncthreads();
    int run = (wid > 0 ? wsum[wid - 1] : 0) + x - local;
    for (int k0 = b0; k0 < b1; k0 += U) {
        if (seg > U) {  // uniform; otherwise v still holds this segment
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = k0 + u < b1 ? count[k0 + u] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k0 + u < b1) {
                start[k0 + u] = run;
                fill[k0 + u] = run;
                count[k0 + u] = 0;
                run += v[u];
            }
        }
    }
    if (threadIdx.x == 1023) start[ncell] = wsum[15];
}

__device__ __forceinline__ void scan_body(int blk, int nblk, const int *__restrict__ gate, int ncell, int *__restrict__ count,
                                          int *__restrict__ start, int *__restrict__ fill,
                                          const BoxDev *__restrict__ box, BoxDev *__restrict__ boxLast) {
    if (gate && *gate == 0) return;
    scan_wg(ncell, count, start, fill, box, boxLast);
}

// Captured steps: the binning with the scan in its last workgroup (1024
// threads a workgroup), one launch fewer on every step's path.
__device__ __forceinline__ void bin_scan_body(int blk, int nblk, const int *__restrict__ gate, int n,
                                              double4 *__restrict__ pos, const float *__restrict__ q,
                                              const BoxDev *__restrict__ bx, int ncell, const int *__restrict__ colRank,
                                              int *__restrict__ cellOf, int *__restrict__ cellCount,
                                              const double4 *__restrict__ cen, const int *__restrict__ ids, int nslot,
                                              int *__restrict__ start, int *__restrict__ fill,
                                              BoxDev *__restrict__ boxLast, int *__restrict__ ticket) {
    if (gate && *gate == 0) return;
    const int i = blk * blockDim.x + threadIdx.x;
    if (i < n) bin_atom(i, pos, q, bx, ncell, colRank, cellOf, cellCount, cen, ids);
    if (last_block(ticket, nblk)) scan_wg(nslot, cellCount, start, fill, bx, boxLast);
}

__device__ __forceinline__ void place_body(int blk, int nblk, const int *__restrict__ gate, int n, const int *__restrict__ cellOf, int *__restrict__ fill,
                        int *__restrict__ perm) {
    if (gate && *gate == 0) return;
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int slot = atomicAdd(&fill[cellOf[i]], 1);
    perm[slot] = i;
}

// Make the order inside each cell independent of atomic slot order: one
// workgroup per cell slot ranks its entries by atom id in LDS (ids are unique,
// so rank = number of smaller keys) and writes them in rank order.
__device__ __forceinline__ void cellsort_body(int blk, int nblk, const int *__restrict__ gate, int ncell,
                                                 const int *__restrict__ start, const int *__restrict__ permIn,
                                                 int *__restrict__ permOut, const int *__restrict__ id) {
    if (gate && *gate == 0) return;
    constexpr int CAP = 1024;
    __shared__ int skey[CAP];
    __shared__ int sperm[CAP];
    const int c = blk;
    const int a = start[c], b = start[c + 1], m = b - a;
    if (m <= CAP) {
        for (int t = threadIdx.x; t < m; t += 64) {
            const int p = permIn[a + t];
            sperm[t] = p;
            skey[t] = id[p];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m; t += 64) {
            const int key = skey[t];
            int r = 0;
            for (int j = 0; j < m; ++j) r += skey[j] < key;
            permOut[a + r] = sperm[t];
        }
    } else if (threadIdx.x == 0) {
        // very crowded cell: serial insertion sort
        for (int s = a; s < b; ++s) permOut[s] = permIn[s];
        for (int s = a + 1; s < b; ++s) {
            const int v = permOut[s];
            const int key = id[v];
            int t = s - 1;
            while (t >= a && id[permOut[t]] > key) {
                permOut[t + 1] = permOut[t];
                --t;
            }
            permOut[t + 1] = v;
        }
    }
}

// Gather the per-atom records into cell order (sortPerAtomArrays,
// GridGPU.cu:253-293) and refresh idToIdx, the last-build positions and xq.
__device__ __forceinline__ void permute_body(int blk, int nblk, const int *__restrict__ gate, int n, const int *__restrict__ perm, const double4 *__restrict__ pos0,
                          const double4 *__restrict__ vel0, const float *__restrict__ q0,
                          const int *__restrict__ t0, const int *__restrict__ id0,
                          double4 *__restrict__ pos1, double4 *__restrict__ vel1,
                          float *__restrict__ q1, int *__restrict__ t1, int *__restrict__ id1,
                          int *__restrict__ idToIdx, double4 *__restrict__ last,
                          float4 *__restrict__ xq, const BoxDev *__restrict__ bx, float4 *__restrict__ rec,
                          const double4 *__restrict__ cen, double4 *__restrict__ bpos) {
    if (gate && *gate == 0) return;
    const Geometry g = bx->geo;
    int k = blk * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int o = perm[k];
    double4 p = pos0[o];
    pos1[k] = p;
    vel1[k] = vel0[o];
    float qq = q0[o];
    q1[k] = qq;
    t1[k] = t0[o];
    int myid = id0[o];
    id1[k] = myid;
    idToIdx[myid] = k;
    last[k] = p;
    // the P3M mesh records in the new order (captured steps whose first half
    // kick refreshed them before the rebuild)
    if (rec && qq != 0.0f) rec[k] = mesh_record(p, bx->mesh);
    // ring centroid lists: the build works on the centroids (xq is written
    // again from the positions before the forces)
    if (cen) {
        p = cen[myid];
        bpos[k] = p;
    }
    xq[k] = make_float4((float)(p.x - g.lo[0]), (float)(p.y - g.lo[1]), (float)(p.z - g.lo[2]), qq);
}

__device__ __forceinline__ void xq_body(int blk, int nblk, int n, const double4 *__restrict__ pos, const float *__restrict__ q,
                     float4 *__restrict__ xq, const BoxDev *__restrict__ bx) {
    const Geometry g = bx->geo;
    int k = blk * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 p = pos[k];
    xq[k] = make_float4((float)(p.x - g.lo[0]), (float)(p.y - g.lo[1]), (float)(p.z - g.lo[2]), q[k]);
}

// k_xq fused with the clearing of the fixed-point force rows and of the
// energy / virial accumulators that the force kernels add into; with P3M on
// (rec set) also each charged site's mesh point and offsets for the spread
// and the gather of the same force evaluation.
__device__ __forceinline__ void xq_zero_body(int blk, int nblk, int n, const double4 *__restrict__ pos, const float *__restrict__ q,
                          float4 *__restrict__ xq, const BoxDev *__restrict__ bx,
                          unsigned long long *__restrict__ force, int nacc, unsigned long long *__restrict__ acc,
                          float4 *__restrict__ rec, RowMaxArgs rm) {
    if (rm.on && blk == nblk - 1) {  // uniform over the workgroup
        if (!rm.gate || *rm.gate != 0) row_max_wg(rm.na, rm.a, rm.nb, rm.b, rm.flags);
        return;
    }
    const Geometry g = bx->geo;
    int k = blk * blockDim.x + threadIdx.x;
    if (k < nacc) acc[k] = 0ull;
    if (k >= n) return;
    double4 p = pos[k];
    const float qk = q[k];
    xq[k] = make_float4((float)(p.x - g.lo[0]), (float)(p.y - g.lo[1]), (float)(p.z - g.lo[2]), qk);
    force[3 * (size_t)k + 0] = 0ull;
    force[3 * (size_t)k + 1] = 0ull;
    force[3 * (size_t)k + 2] = 0ull;
    if (rec && qk != 0.0f) rec[k] = mesh_record(p, bx->mesh);
}

// exclusion row of atom i (id myid) in current indices, tags kept
__device__ __forceinline__ void excl_row(int i, int myid, const int *__restrict__ exclOff,
                                         const uint32_t *__restrict__ exclIds, const int *__restrict__ idToIdx,
                                         int maxExcl, uint32_t *__restrict__ exclIdx) {
    const int e0 = exclOff[myid], e1 = exclOff[myid + 1];
    for (int q = 0; q < maxExcl; ++q) {
        uint32_t v = ~0u;
        if (q < e1 - e0) {
            const uint32_t e = exclIds[e0 + q];
            v = (uint32_t)idToIdx[e & 0x3FFFFFFFu] | (e & 0xC0000000u);
        }
        exclIdx[(size_t)i * maxExcl + q] = v;
    }
}

__device__ __forceinline__ void copy_back_body(int blk, int nblk, const int *__restrict__ gate, int n, const double4 *__restrict__ pos1,
                            const double4 *__restrict__ vel1, const float *__restrict__ q1,
                            const int *__restrict__ t1, const int *__restrict__ id1, double4 *__restrict__ pos0,
                            double4 *__restrict__ vel0, float *__restrict__ q0, int *__restrict__ t0,
                            int *__restrict__ id0, const int *__restrict__ exclOff,
                            const uint32_t *__restrict__ exclIds, const int *__restrict__ idToIdx, int maxExcl,
                            uint32_t *__restrict__ exclIdx, int *__restrict__ atomMolO,
                            int *__restrict__ member) {
    if (gate && *gate == 0) return;
    int k = blk * blockDim.x + threadIdx.x;
    if (k >= n) return;
    pos0[k] = pos1[k];
    vel0[k] = vel1[k];
    q0[k] = q1[k];
    t0[k] = t1[k];
    const int myid = id1[k];
    id0[k] = myid;
    // the exclusion rows (k_excl_idx) and the cleared oxygen map of the new
    // order in the same pass: one launch fewer in every captured step
    if (maxExcl > 0) excl_row(k, myid, exclOff, exclIds, idToIdx, maxExcl, exclIdx);
    if (atomMolO) atomMolO[k] = -1;
    if (member) member[k] = 0;
}

// one wave: the maxima of the nb workgroups

// Device-side step decisions of a captured step (one thread): rebuild the
// list iff the largest displacement since the last build exceeds skin/2 (the
// reference's setBuildFlag test, GridGPU.cu:755-774), count it, and set the
// thermostat gate for the steps where the turn reaches a multiple of its
// interval.  Clears the displacement slot for the next step.
__device__ __forceinline__ void decide_body(int blk, int nblk, int nb, const float *__restrict__ part, unsigned *__restrict__ dmax,
                         int *__restrict__ flags, float half2, int every, int baroEvery,
                         unsigned long long *__restrict__ thermo) {
    DecideArgs d{};
    d.part = part;
    d.dmax = dmax;
    d.flags = flags;
    d.half2 = half2;
    d.every = every;
    d.baroEvery = baroEvery;
    d.thermo = thermo;
    decide_wg(nb, d);
}

// Stencil walk without local arrays (a dynamically indexed local array would
// live in scratch memory): along x and y, slot s of axis_slots() is either
// one of the (2 reach + 1) cells around c with periodic wrap, or every cell
// once when the axis has fewer cells; -1 marks a slot outside a non-periodic
// box.  Along z the candidates of a column are one or two contiguous runs.
__device__ __forceinline__ int axis_slots(int nc, int reach) { return nc < 2 * reach + 1 ? nc : 2 * reach + 1; }

__device__ __forceinline__ int axis_cell(int s, int c, int nc, int reach, int periodic) {
    if (nc < 2 * reach + 1) return s;
    int cc = c - reach + s;
    if (cc < 0 || cc >= nc) {
        if (!periodic) return -1;
        cc = cc < 0 ? cc + nc : cc - nc;
    }
    return cc;
}

// One or two contiguous runs of z-cells of a column; k0 / k1 is the image
// (-1, 0, +1 box lengths) in which a run's cells sit in the unwrapped stencil.
struct ZRuns {
    int n, lo0, hi0, lo1, hi1, k0, k1;
};

__device__ __forceinline__ ZRuns z_runs(int cz, int ncz, int reach, int periodic) {
    ZRuns z;
    z.k0 = z.k1 = 0;
    if (ncz < 2 * reach + 1) {
        z.n = 1;
        z.lo0 = 0;
        z.hi0 = ncz - 1;
        z.lo1 = z.hi1 = 0;
        return z;
    }
    int lo = cz - reach, hi = cz + reach;
    z.n = 1;
    z.lo1 = z.hi1 = 0;
    if (lo < 0) {
        if (periodic) {
            z.lo1 = lo + ncz;
            z.hi1 = ncz - 1;
            z.k1 = -1;
            z.n = 2;
        }
        lo = 0;
    } else if (hi >= ncz) {
        if (periodic) {
            z.lo1 = 0;
            z.hi1 = hi - ncz;
            z.k1 = 1;
            z.n = 2;
        }
        hi = ncz - 1;
    }
    z.lo0 = lo;
    z.hi0 = hi;
    return z;
}

__device__ __forceinline__ float dist2f(const float4 &a, const float4 &b, const Prefilter &p) {
    float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    dx -= p.L[0] * rintf(dx * p.invL[0]);
    dy -= p.L[1] * rintf(dy * p.invL[1]);
    dz -= p.L[2] * rintf(dz * p.invL[2]);
    return dx * dx + dy * dy + dz * dz;
}

// z-cells [ulo, uhi] (unwrapped, within one period of [0, nc)) as one or two
// contiguous runs of a periodic column
__device__ __forceinline__ ZRuns z_runs_cut(int ulo, int uhi, int ncz) {
    ZRuns z;
    z.n = 1;
    z.lo1 = z.hi1 = 0;
    z.k0 = z.k1 = 0;
    if (ulo > uhi) {
        z.n = 0;
        z.lo0 = z.hi0 = 0;
        return z;
    }
    if (ulo < 0) {
        if (uhi < 0) {
            z.lo0 = ulo + ncz;
            z.hi0 = uhi + ncz;
            z.k0 = -1;
        } else {
            z.lo0 = 0;
            z.hi0 = uhi;
            z.lo1 = ulo + ncz;
            z.hi1 = ncz - 1;
            z.k1 = -1;
            z.n = 2;
        }
    } else if (uhi >= ncz) {
        if (ulo >= ncz) {
            z.lo0 = ulo - ncz;
            z.hi0 = uhi - ncz;
            z.k0 = 1;
        } else {
            z.lo0 = ulo;
            z.hi0 = ncz - 1;
            z.lo1 = 0;
            z.hi1 = uhi - ncz;
            z.k1 = 1;
            z.n = 2;
        }
    } else {
        z.lo0 = ulo;
        z.hi0 = uhi;
    }
    return z;
}

__device__ __forceinline__ float axis_gap(int u, float w, float x) {
    const float a = (float)u * w;
    return fmaxf(0.0f, fmaxf(a - x, x - (a + w)));
}

// One team of T lanes per atom walks the culled stencil.  The team first
// lays out its table of candidate runs in LDS, one stencil column per lane
// (column bases and cell starts are then fetched in parallel, not as a chain
// of dependent loads per column): the runs of the charged atoms' grid, then
// those of the charge-free atoms' grid, each run padded to whole chunks of T
// so that one stream of chunks covers them all and a chunk never straddles two
// runs (the run of a chunk is uniform over the team; padding lanes test
// nothing).  The team then walks the stream U chunks at a time with all U
// loads in flight.  Hits are compacted with a ballot restricted to the team,
// so row order is (charge class, stencil column, z-run, index) -
// deterministic - and the row splits at nlCountA into charged and charge-free
// partners.
//
// SHIFT: every axis is open or holds >= 2 reach + 1 cells, so each candidate
// of a run sits in one known image of the unwrapped stencil (the run's image
// offsets are stored with it) and the FP32 pre-test is a plain difference
// against the shifted atom; otherwise it takes the minimum image by rint.
//
// Exclusion partners arrive as current indices (idx | tag << 30, rows of
// maxExcl padded with ~0u); the first four sit in registers and are compared
// only within runs flagged as holding one of them or the atom itself, further
// ones are read for hit lanes only.
constexpr int kMaxCols = 25;  // (2 reach + 1)^2 columns for an xy reach of up to 2
constexpr int kEndBits = 20;  // run ends in a team's candidate stream
constexpr uint32_t kEndMask = (1u << kEndBits) - 1u;

__device__ __forceinline__ bool in_run(uint32_t ex, int s0, int s1) {
    const int j = (int)(ex & 0x3FFFFFFFu);
    return ex != ~0u && j >= s0 && j < s1;
}

template <int T, bool SHIFT>
__device__ __forceinline__ void build_body(int blk, int nblk, const int *__restrict__ gate, 
    int n, const double4 *__restrict__ pos, const float4 *__restrict__ xq, const BoxDev *__restrict__ bx,
    const int *__restrict__ colRank, const int *__restrict__ cellStart, int ncell,
    double rlist2, const uint32_t *__restrict__ exclIdx, int maxExcl, int cap,
    uint32_t *__restrict__ nl, int *__restrict__ nlCount, int *__restrict__ nlCountA, int *__restrict__ flags,
    int *__restrict__ waveMax) {
    if (gate && *gate == 0) return;
    const Geometry g = bx->geo;
    const CellGrid cg = bx->cg;
    const Prefilter pf = bx->pf;
    const Cull cu = bx->cu;
    constexpr int U = T >= 32 ? 1 : 32 / T;  // chunks of T candidates per load block
    constexpr int R = 4 * kMaxCols;           // runs per table: 2 classes x 2 z-runs per column
    __shared__ int sOff[256 / T][R];          // atom index = stream position + sOff
    __shared__ uint32_t sEnd[256 / T][R];     // end of the run's candidates | code << kEndBits
    const int gtid = blk * blockDim.x + threadIdx.x;
    const int i = gtid / T;
    const int team = threadIdx.x / T;
    const int tl = threadIdx.x & (T - 1);
    const int lane = threadIdx.x & 63;
    const bool live = i < n;
    const int rx = cg.reach[0], ry = cg.reach[1], rz = cg.reach[2];
    const int sx = axis_slots(cg.nc[0], rx), sy = axis_slots(cg.nc[1], ry);
    const int ncol = sx * sy;
    const int ncz = cg.nc[2];
    // team-uniform stream lengths: runs are padded to whole chunks of T, so a
    // chunk never straddles two runs; [0, totalA) walks the charged atoms
    int total = 0, totalA = 0;
    if (live) {
        const double4 pi = pos[i];
        const float4 xi = xq[i];
        const int cx = cell_coord(pi.x, g.lo[0], cg.cellInv[0], cg.nc[0]);
        const int cy = cell_coord(pi.y, g.lo[1], cg.cellInv[1], cg.nc[1]);
        const int cz = cell_coord(pi.z, g.lo[2], cg.cellInv[2], cg.nc[2]);
        const uint32_t *exRow = exclIdx + (size_t)i * maxExcl;
        const uint32_t ex0 = maxExcl > 0 ? exRow[0] : ~0u, ex1 = maxExcl > 1 ? exRow[1] : ~0u;
        const uint32_t ex2 = maxExcl > 2 ? exRow[2] : ~0u, ex3 = maxExcl > 3 ? exRow[3] : ~0u;
        int nrun = 0;
#pragma unroll 1
        for (int cls = 0; cls < 2; ++cls) {
            const int *cs = cellStart + cls * ncell;
            for (int c0 = 0; c0 < ncol; c0 += T) {
                const int col = c0 + tl;
                int st0 = 0, len0 = 0, st1 = 0, len1 = 0;
                uint32_t code0 = 0, code1 = 0;
                if (col < ncol) {
                    const int a = col / sy, b = col - a * sy;
                    const int xx = axis_cell(a, cx, cg.nc[0], rx, g.periodic[0]);
                    const int yy = axis_cell(b, cy, cg.nc[1], ry, g.periodic[1]);
                    const float gx = cu.on[0] ? axis_gap(cx - rx + a, cu.w[0], xi.x) : 0.0f;
                    const float gy = cu.on[1] ? axis_gap(cy - ry + b, cu.w[1], xi.y) : 0.0f;
                    const float gxy2 = gx * gx + gy * gy;
                    if (xx >= 0 && yy >= 0 && gxy2 <= cu.r2) {
                        ZRuns zr;
                        if (cu.on[2]) {
                            const float h = sqrtf(cu.r2 - gxy2);
                            const int ulo = max((int)floorf((xi.z - h) / cu.w[2]), cz - rz);
                            const int uhi = min((int)floorf((xi.z + h) / cu.w[2]), cz + rz);
                            zr = z_runs_cut(ulo, uhi, ncz);
                        } else {
                            zr = z_runs(cz, ncz, rz, g.periodic[2]);
                        }
                        const int colBase = colRank[xx * cg.nc[1] + yy] * ncz;
                        if (zr.n > 0) {
                            st0 = cs[colBase + zr.lo0];
                            len0 = cs[colBase + zr.hi0 + 1] - st0;
                        }
                        if (zr.n > 1) {
                            st1 = cs[colBase + zr.lo1];
                            len1 = cs[colBase + zr.hi1 + 1] - st1;
                        }
                        const int ux = cx - rx + a, uy = cy - ry + b;
                        const uint32_t kx = ux < 0 ? 0u : (ux >= cg.nc[0] ? 2u : 1u);
                        const uint32_t ky = uy < 0 ? 0u : (uy >= cg.nc[1] ? 2u : 1u);
                        // flag: the run holds the atom itself or an exclusion partner
                        const uint32_t f0 =
                            (maxExcl > 4 || (i >= st0 && i < st0 + len0) || in_run(ex0, st0, st0 + len0) ||
                             in_run(ex1, st0, st0 + len0) || in_run(ex2, st0, st0 + len0) ||
                             in_run(ex3, st0, st0 + len0)) ? 1u : 0u;
                        const uint32_t f1 =
                            (maxExcl > 4 || (i >= st1 && i < st1 + len1) || in_run(ex0, st1, st1 + len1) ||
                             in_run(ex1, st1, st1 + len1) || in_run(ex2, st1, st1 + len1) ||
                             in_run(ex3, st1, st1 + len1)) ? 1u : 0u;
                        code0 = kx | (ky << 2) | ((uint32_t)(zr.k0 + 1) << 4) | (f0 << 6);
                        code1 = kx | (ky << 2) | ((uint32_t)(zr.k1 + 1) << 4) | (f1 << 6);
                    }
                }
                // team prefix sums place the non-empty runs in column order
                const int nr = (len0 > 0 ? 1 : 0) + (len1 > 0 ? 1 : 0);
                const int pl0 = (len0 + T - 1) & ~(T - 1), pl1 = (len1 + T - 1) & ~(T - 1);
                const int ln = pl0 + pl1;
                int snr = nr, sln = ln;
#pragma unroll
                for (int o = 1; o < T; o <<= 1) {
                    const int y0 = __shfl_up(snr, o, T), y1 = __shfl_up(sln, o, T);
                    if (tl >= o) {
                        snr += y0;
                        sln += y1;
                    }
                }
                int e = nrun + snr - nr, v = total + sln - ln;
                if (len0 > 0) {
                    sOff[team][e] = st0 - v;
                    sEnd[team][e] = (uint32_t)(v + len0) | (code0 << kEndBits);
                    v += pl0;
                    ++e;
                }
                if (len1 > 0) {
                    sOff[team][e] = st1 - v;
                    sEnd[team][e] = (uint32_t)(v + len1) | (code1 << kEndBits);
                }
                nrun += __shfl(snr, T - 1, T);
                total += __shfl(sln, T - 1, T);
            }
            if (cls == 0) totalA = total;
        }
        // stream positions are kEndBits wide
        if (tl == 0 && total >= (1 << kEndBits)) atomicOr(&flags[7], 1);
    }
    // a team lies inside one wave and reads only its own table
    wave_sync();
    int cnt = 0;
    if (live && total < (1 << kEndBits)) {
        const unsigned long long teamMask =
            (T == 64) ? ~0ull : (((1ull << T) - 1ull) << (lane & ~(T - 1)));
        const uint32_t tmLo = (uint32_t)teamMask, tmHi = (uint32_t)(teamMask >> 32);
        const double4 pi = pos[i];
        const float4 xi = xq[i];
        const uint32_t *exRow = exclIdx + (size_t)i * maxExcl;
        const uint32_t ex0 = maxExcl > 0 ? exRow[0] : ~0u, ex1 = maxExcl > 1 ? exRow[1] : ~0u;
        const uint32_t ex2 = maxExcl > 2 ? exRow[2] : ~0u, ex3 = maxExcl > 3 ? exRow[3] : ~0u;
        const uint32_t e0 = ex0 & 0x3FFFFFFFu, e1 = ex1 & 0x3FFFFFFFu, e2 = ex2 & 0x3FFFFFFFu, e3 = ex3 & 0x3FFFFFFFu;
        uint32_t *row = nl + (size_t)i * cap;
        int cntA = 0;
        // the team's current run (team-uniform): offset, end of its real
        // candidates, flag bit, and the atom moved into the run's image frame;
        // the next entry is read ahead
        int r = -1, off = 0, vend = 0;
        uint32_t fl = 0, img = 0x15u;  // image code of the interior, (0, 0, 0) + 1 per axis
        float xsx = xi.x, xsy = xi.y, xsz = xi.z;
        int nOff = sOff[team][0];
        uint32_t nEnd = sEnd[team][0];
        for (int v0 = 0; v0 < total; v0 += U * T) {
            uint32_t jv[U];  // atom index, bit 31: flagged run
            float px[U], py[U], pz[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = v0 + u * T;
                if (p < total && p >= vend) {
                    ++r;
                    off = nOff;
                    vend = (int)(nEnd & kEndMask);
                    const uint32_t code = nEnd >> kEndBits;
                    fl = (code >> 6) << 31;
                    // most runs share the image of the previous one
                    if (SHIFT && (code & 0x3Fu) != img) {
                        img = code & 0x3Fu;
                        xsx = xi.x - (float)((int)(code & 3u) - 1) * pf.L[0];
                        xsy = xi.y - (float)((int)((code >> 2) & 3u) - 1) * pf.L[1];
                        xsz = xi.z - (float)((int)((code >> 4) & 3u) - 1) * pf.L[2];
                    }
                    const int rn = min(r + 1, R - 1);
                    nOff = sOff[team][rn];
                    nEnd = sEnd[team][rn];
                }
                const int v = p + tl;
                // padding and positions past the stream: the atom itself in a
                // flagged slot (never a hit)
                jv[u] = (p < total && v < vend) ? ((uint32_t)(v + off) | fl) : ((uint32_t)i | 0x80000000u);
                px[u] = xsx;
                py[u] = xsy;
                pz[u] = xsz;
            }
            float4 blk[U];
#pragma unroll
            for (int u = 0; u < U; ++u) blk[u] = xq[jv[u] & 0x7FFFFFFFu];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (v0 + u * T == totalA) cntA = cnt;  // first charge-free chunk
                const int j = (int)(jv[u] & 0x7FFFFFFFu);
                const float4 cur = blk[u];
                float r2f;
                if (SHIFT) {
                    const float dx = cur.x - px[u], dy = cur.y - py[u], dz = cur.z - pz[u];
                    r2f = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                } else {
                    r2f = dist2f(xi, cur, pf);
                }
                bool hit = r2f < pf.lo2;
                if (!hit && r2f <= pf.hi2) hit = dist2(pi, pos[j], g) < rlist2;
                uint32_t tag = 0;
                if ((int)jv[u] < 0) {
                    hit = hit && j != i;
                    const uint32_t uj = (uint32_t)j;
                    tag = e0 == uj ? ex0 >> 30 : tag;
                    tag = e1 == uj ? ex1 >> 30 : tag;
                    tag = e2 == uj ? ex2 >> 30 : tag;
                    tag = e3 == uj ? ex3 >> 30 : tag;
                    if (maxExcl > 4 && hit)
                        for (int q = 4; q < maxExcl; ++q) {
                            const uint32_t e = exRow[q];
                            if ((e & 0x3FFFFFFFu) == uj) tag = e >> 30;
                        }
                }
                const unsigned long long m = __ballot(hit);
                const uint32_t mLo = (uint32_t)m & tmLo, mHi = (uint32_t)(m >> 32) & tmHi;
                if (hit) {
                    // team hits in lower lanes, counted by mbcnt on the masked ballot
                    const int k = (int)__builtin_amdgcn_mbcnt_hi(mHi, __builtin_amdgcn_mbcnt_lo(mLo, (uint32_t)cnt));
                    if (k < cap) row[k] = (uint32_t)j | (tag << 30);
                }
                cnt += __popc(mLo) + __popc(mHi);
            }
        }
        if (totalA >= total) cntA = cnt;  // no charge-free candidates
        if (tl == 0) {
            nlCount[i] = cnt;
            nlCountA[i] = cntA;
        }
    }
    // the longest row of the wave goes to the wave's own slot (k_row_max
    // reduces them): same-address device atomics serialise across the XCDs
    const int wmax = wave_max(cnt);
    if (lane == 0) waveMax[gtid >> 6] = wmax;
}

// One wave per atom (the team width 64 of the tuner).  The 64 lanes lay out
// the table in one step: lanes 0..31 take the columns of the charged atoms'
// grid, lanes 32..63 those of the charge-free atoms' grid, and a wave-wide
// prefix sum places the runs, unpadded, in one stream (class, column, z-run
// order).  Each table entry carries the run's end in the stream, the offset
// from stream position to atom index, its flag and the atom moved into the
// run's image frame (uniform over the wave, so computed once per run).  The
// wave then walks the stream 64 candidates per chunk and U chunks per block:
// a lane finds its run from the run of the chunk start (runs are rarely
// shorter than the few lanes they straddle), so the bookkeeping is a couple
// of LDS reads per lane and the compaction is one mbcnt on the wave ballot.
struct RunEntry {
    int vend;      // end of the run in the stream
    int off;       // atom index = stream position + off
    uint32_t fl;   // bit 31: the run holds the atom itself or an exclusion partner
    float xs[3];   // the atom in the run's image frame
};

template <bool SHIFT>
__device__ __forceinline__ void build_wave_body(int blk, int nblk, const int *__restrict__ gate,
    int n, const double4 *__restrict__ pos, const float4 *__restrict__ xq, const BoxDev *__restrict__ bx,
    const int *__restrict__ colRank, const int *__restrict__ cellStart, int ncell,
    double rlist2, const uint32_t *__restrict__ exclIdx, int maxExcl, int cap,
    uint32_t *__restrict__ nl, int *__restrict__ nlCount, int *__restrict__ nlCountA, int *__restrict__ flags,
    int *__restrict__ waveMax) {
    if (gate && *gate == 0) return;
    const Geometry g = bx->geo;
    const CellGrid cg = bx->cg;
    const Prefilter pf = bx->pf;
    const Cull cu = bx->cu;
    constexpr int U = 4;             // chunks of 64 candidates per load block
    constexpr int R = 4 * kMaxCols;  // 2 classes x 2 z-runs per column
    __shared__ RunEntry sRun[4][R + 1];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = __builtin_amdgcn_readfirstlane(blk * 4 + (int)(threadIdx.x >> 6));
    if (i >= n) {
        if (lane == 0) waveMax[i] = 0;
        return;
    }
    const int rx = cg.reach[0], ry = cg.reach[1], rz = cg.reach[2];
    const int sx = axis_slots(cg.nc[0], rx), sy = axis_slots(cg.nc[1], ry);
    const int ncol = sx * sy;
    const int ncz = cg.nc[2];
    const double4 pi = pos[i];
    const float4 xi = xq[i];
    const int cx = cell_coord(pi.x, g.lo[0], cg.cellInv[0], cg.nc[0]);
    const int cy = cell_coord(pi.y, g.lo[1], cg.cellInv[1], cg.nc[1]);
    const int cz = cell_coord(pi.z, g.lo[2], cg.cellInv[2], cg.nc[2]);
    const uint32_t *exRow = exclIdx + (size_t)i * maxExcl;
    const uint32_t ex0 = maxExcl > 0 ? exRow[0] : ~0u, ex1 = maxExcl > 1 ? exRow[1] : ~0u;
    const uint32_t ex2 = maxExcl > 2 ? exRow[2] : ~0u, ex3 = maxExcl > 3 ? exRow[3] : ~0u;
    // ---- table: one (class, column) per lane
    const int cls = lane >> 5, col = lane & 31;
    const int *cs = cellStart + cls * ncell;
    int st0 = 0, len0 = 0, st1 = 0, len1 = 0;
    uint32_t f0 = 0, f1 = 0;
    float s0x = xi.x, s0y = xi.y, s0z = xi.z, s1x = xi.x, s1y = xi.y, s1z = xi.z;
    if (col < ncol) {
        const int a = col / sy, b = col - a * sy;
        const int xx = axis_cell(a, cx, cg.nc[0], rx, g.periodic[0]);
        const int yy = axis_cell(b, cy, cg.nc[1], ry, g.periodic[1]);
        const float gx = cu.on[0] ? axis_gap(cx - rx + a, cu.w[0], xi.x) : 0.0f;
        const float gy = cu.on[1] ? axis_gap(cy - ry + b, cu.w[1], xi.y) : 0.0f;
        const float gxy2 = gx * gx + gy * gy;
        if (xx >= 0 && yy >= 0 && gxy2 <= cu.r2) {
            ZRuns zr;
            if (cu.on[2]) {
                const float h = sqrtf(cu.r2 - gxy2);
                const int ulo = max((int)floorf((xi.z - h) / cu.w[2]), cz - rz);
                const int uhi = min((int)floorf((xi.z + h) / cu.w[2]), cz + rz);
                zr = z_runs_cut(ulo, uhi, ncz);
            } else {
                zr = z_runs(cz, ncz, rz, g.periodic[2]);
            }
            const int colBase = colRank[xx * cg.nc[1] + yy] * ncz;
            if (zr.n > 0) {
                st0 = cs[colBase + zr.lo0];
                len0 = cs[colBase + zr.hi0 + 1] - st0;
            }
            if (zr.n > 1) {
                st1 = cs[colBase + zr.lo1];
                len1 = cs[colBase + zr.hi1 + 1] - st1;
            }
            f0 = (maxExcl > 4 || (i >= st0 && i < st0 + len0) || in_run(ex0, st0, st0 + len0) ||
                  in_run(ex1, st0, st0 + len0) || in_run(ex2, st0, st0 + len0) || in_run(ex3, st0, st0 + len0))
                     ? 0x80000000u : 0u;
            f1 = (maxExcl > 4 || (i >= st1 && i < st1 + len1) || in_run(ex0, st1, st1 + len1) ||
                  in_run(ex1, st1, st1 + len1) || in_run(ex2, st1, st1 + len1) || in_run(ex3, st1, st1 + len1))
                     ? 0x80000000u : 0u;
            if (SHIFT) {
                const int ux = cx - rx + a, uy = cy - ry + b;
                const float kx = ux < 0 ? -1.0f : (ux >= cg.nc[0] ? 1.0f : 0.0f);
                const float ky = uy < 0 ? -1.0f : (uy >= cg.nc[1] ? 1.0f : 0.0f);
                s0x = s1x = xi.x - kx * pf.L[0];
                s0y = s1y = xi.y - ky * pf.L[1];
                s0z = xi.z - (float)zr.k0 * pf.L[2];
                s1z = xi.z - (float)zr.k1 * pf.L[2];
            }
        }
    }
    const int nr = (len0 > 0 ? 1 : 0) + (len1 > 0 ? 1 : 0);
    const int ln = len0 + len1;
    int snr = nr, sln = ln;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y0 = __shfl_up(snr, o, 64), y1 = __shfl_up(sln, o, 64);
        if (lane >= o) {
            snr += y0;
            sln += y1;
        }
    }
    {
        int e = snr - nr, v = sln - ln;
        if (len0 > 0) {
            RunEntry &q = sRun[w][e];
            q.vend = v + len0;
            q.off = st0 - v;
            q.fl = f0;
            q.xs[0] = s0x;
            q.xs[1] = s0y;
            q.xs[2] = s0z;
            v += len0;
            ++e;
        }
        if (len1 > 0) {
            RunEntry &q = sRun[w][e];
            q.vend = v + len1;
            q.off = st1 - v;
            q.fl = f1;
            q.xs[0] = s1x;
            q.xs[1] = s1y;
            q.xs[2] = s1z;
        }
    }
    const int nrun = __shfl(snr, 63, 64);
    const int total = __shfl(sln, 63, 64);
    const int totalA = __shfl(sln, 31, 64);
    // a sentinel past the last run stops every search
    if (lane == 0) {
        RunEntry &q = sRun[w][nrun];
        q.vend = 0x7FFFFFFF;
        q.off = 0;
        q.fl = 0x80000000u;
        q.xs[0] = xi.x;
        q.xs[1] = xi.y;
        q.xs[2] = xi.z;
    }
    wave_sync();
    // ---- walk
    const uint32_t e0 = ex0 & 0x3FFFFFFFu, e1 = ex1 & 0x3FFFFFFFu, e2 = ex2 & 0x3FFFFFFFu, e3 = ex3 & 0x3FFFFFFFu;
    uint32_t *row = nl + (size_t)i * cap;
    int cnt = 0, cntA = 0;
    int ra = 0;  // run of the current chunk start (uniform)
    for (int p0 = 0; p0 < total; p0 += 64 * U) {
        uint32_t jv[U];
        float px[U], py[U], pz[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = p0 + 64 * u;
            // chunk start's run (uniform search)
            while (__builtin_amdgcn_readfirstlane(sRun[w][ra].vend) <= p) ++ra;
            const int v = p + lane;
            int r = ra;
            while (v >= sRun[w][r].vend) ++r;
            const RunEntry q = sRun[w][r];
            // past the stream: the atom itself in a flagged slot (never a hit)
            jv[u] = v < total ? ((uint32_t)(v + q.off) | q.fl) : ((uint32_t)i | 0x80000000u);
            px[u] = q.xs[0];
            py[u] = q.xs[1];
            pz[u] = q.xs[2];
        }
        float4 blk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) blk[u] = xq[jv[u] & 0x7FFFFFFFu];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = p0 + 64 * u;
            const int j = (int)(jv[u] & 0x7FFFFFFFu);
            const float4 cur = blk[u];
            float r2f;
            if (SHIFT) {
                const float dx = cur.x - px[u], dy = cur.y - py[u], dz = cur.z - pz[u];
                r2f = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            } else {
                r2f = dist2f(xi, cur, pf);
            }
            bool hit = r2f < pf.lo2;
            if (!hit && r2f <= pf.hi2) hit = dist2(pi, pos[j], g) < rlist2;
            uint32_t tag = 0;
            if ((int)jv[u] < 0) {
                hit = hit && j != i;
                const uint32_t uj = (uint32_t)j;
                tag = e0 == uj ? ex0 >> 30 : tag;
                tag = e1 == uj ? ex1 >> 30 : tag;
                tag = e2 == uj ? ex2 >> 30 : tag;
                tag = e3 == uj ? ex3 >> 30 : tag;
                if (maxExcl > 4 && hit)
                    for (int qq = 4; qq < maxExcl; ++qq) {
                        const uint32_t e = exRow[qq];
                        if ((e & 0x3FFFFFFFu) == uj) tag = e >> 30;
                    }
            }
            const unsigned long long m = __ballot(hit);
            if (hit) {
                const int k = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, (uint32_t)cnt));
                if (k < cap) row[k] = (uint32_t)j | (tag << 30);
            }
            // the charged/charge-free split falls inside this chunk
            if (p <= totalA && totalA < p + 64) {
                const int below = totalA - p;
                const unsigned long long lo = below >= 64 ? ~0ull : ((1ull << below) - 1ull);
                cntA = cnt + __popcll(m & lo);
            }
            cnt += __popcll(m);
        }
    }
    if (totalA >= total) cntA = cnt;
    if (lane == 0) {
        nlCount[i] = cnt;
        nlCountA[i] = cntA;
        waveMax[i] = cnt;
    }
}

// Exclusion partners of each atom as current indices (idx | tag << 30), one
// row of maxExcl per atom padded with ~0u, refreshed after each permutation.
__device__ __forceinline__ void excl_idx_body(int blk, int nblk, const int *__restrict__ gate, int n, const int *__restrict__ id, const int *__restrict__ exclOff,
                           const uint32_t *__restrict__ exclIds, const int *__restrict__ idToIdx,
                           int maxExcl, uint32_t *__restrict__ exclIdx) {
    if (gate && *gate == 0) return;
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    excl_row(i, id[i], exclOff, exclIds, idToIdx, maxExcl, exclIdx);
}

// E3B molecule list from the atom list: for each water, the oxygens among its
// oxygen's neighbors within (r2b + skin) at build time.  Replaces the private
// molecule GridGPU of FixE3B (FixE3B.cu:656-684) that the reference rebuilds
// every step; with the skin criterion the per-step r2b set is a subset.
template <int T>
__device__ __forceinline__ void e3b_list_body(int blk, int nblk, 
    const int *__restrict__ gate, int nmol, const int4 *__restrict__ molAtoms, const int *__restrict__ atomMolO,
    const double4 *__restrict__ pos, const BoxDev *__restrict__ bx, double rcut2, int cap,
    const uint32_t *__restrict__ nl,
    const int *__restrict__ nlCount, const int *__restrict__ nlCountA, int unchargedOnly, int mcap,
    int *__restrict__ mnl, int *__restrict__ mcount, int *__restrict__ waveMax) {
    if (gate && *gate == 0) return;
    const Geometry g = bx->geo;
    const int gtid = blk * blockDim.x + threadIdx.x;
    const int m = gtid / T;
    const int tl = threadIdx.x & (T - 1);
    const int lane = threadIdx.x & 63;
    int cnt = 0;
    if (m < nmol) {
        const unsigned long long teamMask =
            (T == 64) ? ~0ull : (((1ull << T) - 1ull) << (lane & ~(T - 1)));
        const unsigned long long below = (1ull << lane) - 1ull;
        const int oi = molAtoms[m].x;
        const double4 pi = pos[oi];
        const int total = min(nlCount[oi], cap);
        const int nA = min(nlCountA[oi], total);
        const uint32_t *row = nl + (size_t)oi * cap;
        // row = [charged partners | charge-free partners]; E3B oxygens of a
        // charge-free O site need only the second part
        const int first = unchargedOnly ? nA : 0;
        for (int base = first; base < total; base += T) {
            const int v = base + tl;
            bool hit = false;
            int mj = -1;
            if (v < total) {
                const int j = (int)(row[v] & 0x3FFFFFFFu);
                mj = atomMolO[j];
                if (mj >= 0 && mj != m) hit = dist2(pi, pos[j], g) < rcut2;
            }
            const unsigned long long msk = __ballot(hit) & teamMask;
            if (hit) {
                const int s = cnt + __popcll(msk & below);
                if (s < mcap) mnl[(size_t)m * mcap + s] = mj;
            }
            cnt += __popcll(msk);
        }
        if (tl == 0) mcount[m] = cnt;
    }
    const int wmax = wave_max(cnt);
    if (lane == 0) waveMax[gtid >> 6] = wmax;
}

// Captured steps also map the rigid molecules here (k_rigid_map's work, rigid
// molecules nmolR > 0, member cleared by k_copy_back): one launch for both.
__device__ __forceinline__ void mol_atoms_body(int blk, int nblk, const int *__restrict__ gate, int nmol, const int4 *__restrict__ molIds, const int *__restrict__ idToIdx,
                            int4 *__restrict__ molAtoms, int *__restrict__ atomMolO, int nmolR,
                            const int4 *__restrict__ rigidIds, int4 *__restrict__ rigidAtoms,
                            int *__restrict__ member) {
    if (gate && *gate == 0) return;
    int m = blk * blockDim.x + threadIdx.x;
    if (m < nmolR) {
        const int4 i = rigidIds[m];
        const int4 a = make_int4(idToIdx[i.x], idToIdx[i.y], idToIdx[i.z], i.w >= 0 ? idToIdx[i.w] : -1);
        rigidAtoms[m] = a;
        member[a.x] = 1;
        member[a.y] = 1;
        member[a.z] = 1;
        if (a.w >= 0) member[a.w] = 1;
    }
    if (m >= nmol) return;
    int4 ids = molIds[m];
    int4 a = make_int4(idToIdx[ids.x], idToIdx[ids.y], idToIdx[ids.z], idToIdx[ids.w]);
    molAtoms[m] = a;
    atomMolO[a.x] = m;
}

// The list build with the molecule map of a rebuild (mol_atoms_body) as
// extra workgroups of the same launch: the map only needs the new indices,
// which the permutation wrote before, and the E3B list that reads it comes
// after (one launch fewer per step).
struct MolMapArgs {
    int blocks;
    int nmol;
    const int4 *molIds;
    const int *idToIdx;
    int4 *molAtoms;
    int *atomMolO;
    int nmolR;
    const int4 *rigidIds;
    int4 *rigidAtoms;
    int *member;
};

template <bool SHIFT>
__device__ __forceinline__ void build_map_body(int blk, int nblk, const int *__restrict__ gate,
    int n, const double4 *__restrict__ pos, const float4 *__restrict__ xq, const BoxDev *__restrict__ bx,
    const int *__restrict__ colRank, const int *__restrict__ cellStart, int ncell,
    double rlist2, const uint32_t *__restrict__ exclIdx, int maxExcl, int cap,
    uint32_t *__restrict__ nl, int *__restrict__ nlCount, int *__restrict__ nlCountA, int *__restrict__ flags,
    int *__restrict__ waveMax, MolMapArgs mm) {
    const int nb = nblk - mm.blocks;
    if (blk >= nb) {  // uniform over the workgroup
        mol_atoms_body(blk - nb, mm.blocks, gate, mm.nmol, mm.molIds, mm.idToIdx, mm.molAtoms, mm.atomMolO, mm.nmolR,
                       mm.rigidIds, mm.rigidAtoms, mm.member);
        return;
    }
    build_wave_body<SHIFT>(blk, nb, gate, n, pos, xq, bx, colRank, cellStart, ncell, rlist2, exclIdx, maxExcl, cap, nl,
                           nlCount, nlCountA, flags, waveMax);
}

__device__ __forceinline__ void fill_int_body(int blk, int nblk, const int *__restrict__ gate, int n, int *__restrict__ p, int v) {
    if (gate && *gate == 0) return;
    int i = blk * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Longest pair row and molecule row of the build from the per-wave slots,
// folded into the running maxima flags[1], flags[2] (one workgroup).
__device__ __forceinline__ void row_max_body(int blk, int nblk, const int *__restrict__ gate, int na, const int *__restrict__ a,
                                                  int nb, const int *__restrict__ b, int *__restrict__ flags) {
    if (gate && *gate == 0) return;
    row_max_wg(na, a, nb, b, flags);
}

uint32_t morton2(uint32_t x, uint32_t y) {
    uint64_t r = 0;
    for (int b = 0; b < 16; ++b) {
        r |= (uint64_t)((x >> b) & 1u) << (2 * b);
        r |= (uint64_t)((y >> b) & 1u) << (2 * b + 1);
    }
    return (uint32_t)r;
}

// The FP32 pre-test band (box_derive) covers the FP32 error of r^2 for
// wrapped positions (x - lo in [0, L]) and r <= r_list.  Per component the
// FP32 difference is off by at most ~5 L 2^-24 (two rounded inputs, the
// subtraction, the rounded L and the image shift); r^2 is then off by at most
// 2 sqrt(3) r_list of that plus a few ulp of r^2 for the squares and the sum.
// The band takes 16 L 2^-24 and 32 r^2 2^-24, about three times both.  The
// culling margin 1e-5 (L + r_list) is far above the FP32 rounding of
// coordinates and cell bounds (~L 2^-24) and far below a cell.

// slots of the pair-list waves in c.waveMax; the molecule-list slots follow
size_t wave_slots_a(const Ctx &c) { return (size_t)c.n + 4; }

// mm.blocks > 0: the molecule map of the rebuild runs in the same launch
void launch_build_wave(Ctx &c, int cur, const MolMapArgs &mm) {
    MDX_REQUIRE(std::min(c.nc[0], 2 * c.reach[0] + 1) * std::min(c.nc[1], 2 * c.reach[1] + 1) <= kMaxCols, 1,
                "stencil wider than the build kernel's column table");
    const int blocks = cdiv(c.n, 4);
    const double4 *bpos = c.cenById ? c.buildPos : c.pos[cur];  // ring centroids or the positions
    bool shift = true;
    for (int a = 0; a < 3; ++a)
        if (c.geo.periodic[a] && c.nc[a] < 2 * c.reach[a] + 1) shift = false;
    if (shift)
        launch<build_map_body<true>, 256>(c, blocks + mm.blocks, 0, rebuild_gate(c), c.n, bpos, c.xq, c.box, c.colRank, c.cellStart, c.ncell, c.rlist * c.rlist, c.exclIdx, c.exclOff ? c.maxExcl : 0, c.cap, c.nl, c.nlCount, c.nlCountA, c.dFlags, c.waveMax, mm);
    else
        launch<build_map_body<false>, 256>(c, blocks + mm.blocks, 0, rebuild_gate(c), c.n, bpos, c.xq, c.box, c.colRank, c.cellStart, c.ncell, c.rlist * c.rlist, c.exclIdx, c.exclOff ? c.maxExcl : 0, c.cap, c.nl, c.nlCount, c.nlCountA, c.dFlags, c.waveMax, mm);
    c.waveMaxA = blocks * 4;
}

template <int T>
void launch_build(Ctx &c, int cur) {
    MDX_REQUIRE(std::min(c.nc[0], 2 * c.reach[0] + 1) * std::min(c.nc[1], 2 * c.reach[1] + 1) <= kMaxCols, 1,
                "stencil wider than the build kernel's column table");
    int threads = 256;
    int blocks = cdiv((long long)c.n * T, threads);
    const double4 *bpos = c.cenById ? c.buildPos : c.pos[cur];  // ring centroids or the positions
    // image shifts need every periodic axis to hold the full stencil
    bool shift = true;
    for (int a = 0; a < 3; ++a)
        if (c.geo.periodic[a] && c.nc[a] < 2 * c.reach[a] + 1) shift = false;
    if (shift)
        launch<build_body<T, true>, 256>(c, blocks, 0, rebuild_gate(c), c.n, bpos, c.xq, c.box, c.colRank, c.cellStart, c.ncell, c.rlist * c.rlist, c.exclIdx, c.exclOff ? c.maxExcl : 0, c.cap, c.nl, c.nlCount, c.nlCountA, c.dFlags, c.waveMax);
    else
        launch<build_body<T, false>, 256>(c, blocks, 0, rebuild_gate(c), c.n, bpos, c.xq, c.box, c.colRank, c.cellStart, c.ncell, c.rlist * c.rlist, c.exclIdx, c.exclOff ? c.maxExcl : 0, c.cap, c.nl, c.nlCount, c.nlCountA, c.dFlags, c.waveMax);
    c.waveMaxA = blocks * (threads / 64);
}

template <int T>
void launch_e3b_list(Ctx &c, int cur) {
    int threads = 256;
    int blocks = cdiv((long long)c.e3b.nmol * T, threads);
    double rc = (double)c.e3b.r2b + c.skin;
    launch<e3b_list_body<T>, 256>(c, blocks, 0, rebuild_gate(c), c.e3b.nmol, c.e3b.molAtoms, c.e3b.atomMolO, c.pos[cur], c.box, rc * rc, c.cap, c.nl, c.nlCount, c.nlCountA, c.e3b.oxUncharged ? 1 : 0, c.e3b.cap, c.e3b.nl, c.e3b.nlCount, c.waveMax + wave_slots_a(c));
    c.waveMaxB = blocks * (threads / 64);
}

}  // namespace

void nlist_setup(Ctx &c) {
    MDX_REQUIRE(c.n > 0, 3, "system not set");
    MDX_REQUIRE(c.rlist > 0, 1, "neighbor list cutoff must be positive");
    for (int a = 0; a < 3; ++a) {
        if (c.geo.periodic[a])
            MDX_REQUIRE(c.rlist < 0.5 * c.geo.L[a], 1,
                        "neighbor cutoff (rCut + padding) must be below half the box length");
        double edge = c.rlist / c.reach[a] * (1.0 + 1e-9);
        int nca = (int)std::floor(c.geo.L[a] * grid_headroom(c) / edge);
        c.nc[a] = std::max(1, nca);
    }
    int ncell = c.nc[0] * c.nc[1] * c.nc[2];
    if (ncell != c.ncell) {
        MDX_REQUIRE(!c.capture, 3, "the cell grid cannot change while a step is captured");
        if (c.ncell) {
            dfree(c, c.colXY);
            dfree(c, c.colRank);
            dfree(c, c.cellCount);
            dfree(c, c.cellStart);
            dfree(c, c.cellFill);
        }
        c.ncell = ncell;
        c.colRank = dalloc<int>(c, (size_t)c.nc[0] * c.nc[1]);
        c.colXY = dalloc<int2>(c, (size_t)c.nc[0] * c.nc[1]);
        // two copies of the grid: charged atoms, then charge-free ones
        c.cellCount = dalloc<int>(c, 2 * (size_t)ncell);
        // k_scan leaves the counts at zero for the next binning
        MDX_HIP(hipMemsetAsync(c.cellCount, 0, sizeof(int) * 2 * (size_t)ncell, c.stream));
        c.cellStart = dalloc<int>(c, 2 * (size_t)ncell + 1);
        c.cellFill = dalloc<int>(c, 2 * (size_t)ncell);
    }
    // column ranks along a 2-D Morton curve
    std::vector<std::pair<uint32_t, int>> keys;
    for (int x = 0; x < c.nc[0]; ++x)
        for (int y = 0; y < c.nc[1]; ++y) keys.push_back({morton2(x, y), x * c.nc[1] + y});
    std::sort(keys.begin(), keys.end());
    std::vector<int> rank(keys.size());
    std::vector<int2> xy(keys.size());
    for (size_t r = 0; r < keys.size(); ++r) {
        rank[keys[r].second] = (int)r;
        xy[r] = make_int2(keys[r].second / c.nc[1], keys[r].second % c.nc[1]);
    }
    MDX_HIP(hipMemcpyAsync(c.colRank, rank.data(), sizeof(int) * rank.size(), hipMemcpyHostToDevice,
                           c.stream));
    MDX_HIP(hipMemcpyAsync(c.colXY, xy.data(), sizeof(int2) * xy.size(), hipMemcpyHostToDevice,
                           c.stream));
    MDX_HIP(hipStreamSynchronize(c.stream));
    if (!c.cellOf) {
        c.cellOf = dalloc<int>(c, c.n);
        c.perm = dalloc<int>(c, c.n);
        c.perm2 = dalloc<int>(c, c.n);
    }
    // one wave per atom (k_build_wave) measured fastest for dense water and
    // LJ alike; mdx_tune revisits the width
    if (!c.buildTeamUser) c.buildTeam = 64;
    const double dens = c.n / (c.geo.L[0] * c.geo.L[1] * c.geo.L[2]);
    if (c.cap == 0) {
        double expect = dens * 4.18879020479 * c.rlist * c.rlist * c.rlist;
        c.cap = ((int)(expect * 1.3) + 32 + 15) & ~15;
        c.nl = dalloc<uint32_t>(c, (size_t)c.n * c.cap);
        c.nlCount = dalloc<int>(c, c.n);
        c.nlCountA = dalloc<int>(c, c.n);
    }
    c.nlistConfigured = true;
    ++c.layoutEpoch;  // cell counts and tables are kernel arguments of captured steps
    box_upload(c);  // cell counts changed
}

void refresh_xq(Ctx &c) {
    int threads = 256;
    launch<xq_body, 256>(c, cdiv(c.n, threads), 0, c.n, c.pos[c.cur], c.q[c.cur], c.xq, c.box);
}

// the pending row maxima of a captured build for the launch that takes them
RowMaxArgs row_max_args(Ctx &c) {
    RowMaxArgs rm{};
    if (c.rowMaxPending) {
        rm.on = 1;
        rm.gate = rebuild_gate(c);
        rm.na = c.waveMaxA;
        rm.a = c.waveMax;
        rm.nb = c.waveMaxB;
        rm.b = c.waveMax + wave_slots_a(c);
        rm.flags = c.dFlags;
        c.rowMaxPending = false;
    }
    return rm;
}

void row_max_flush(Ctx &c) {
    const RowMaxArgs rm = row_max_args(c);
    if (rm.on) launch<row_max_body, 1024>(c, 1, 0, rm.gate, rm.na, rm.a, rm.nb, rm.b, rm.flags);
}

void prepare_forces(Ctx &c) {
    const int threads = 256;
    const RowMaxArgs rm = row_max_args(c);
    launch<xq_zero_body, 256>(c, cdiv(std::max(c.n, kAccTotal), threads) + rm.on, 0, c.n, c.pos[c.cur], c.q[c.cur], c.xq, c.box, c.force, kAccTotal, c.acc,
                              c.pppm.on && c.hasCharges ? c.meshRec : nullptr, rm);
}

void nlist_check(Ctx &c) {
    int threads = 256;
    double thr = 0.5 * c.skin;
    launch<check_body, 256>(c, cdiv(c.n, threads), 0, c.n, c.pos[c.cur], c.posLast, c.box, thr * thr, c.dFlags);
}

static void run_build(Ctx &c, bool sync) {
    // per-wave row maxima: n + 4 slots cover the build at up to 64 lanes per
    // atom (workgroups of 4 waves), the rest the molecule list likewise
    const size_t needSlots = wave_slots_a(c) + (size_t)(c.e3b.on ? c.e3b.nmol : 0) + 4;
    if ((size_t)c.waveMaxN < needSlots) {
        MDX_REQUIRE(!c.capture, 3, "row-maximum slots must exist before a step is captured");
        if (c.waveMax) dfree(c, c.waveMax);
        c.waveMax = dalloc<int>(c, needSlots);
        c.waveMaxN = (int)needSlots;
    }
    int threads = 256;
    int cur = c.cur, nxt = 1 - c.cur;
    const int nslot = 2 * c.ncell;
    MolMapArgs mm{};
    bool mapFused = false;
    const double4 *cen = c.cenById;
    MDX_REQUIRE(!cen || (c.buildPos && !c.e3b.on), 3,
                "a centroid list needs its build positions and takes no E3B molecule list");
    c.centroidList = cen != nullptr;
    if (c.capture && c.fuseStep) {
        launch<bin_scan_body, 1024>(c, cdiv(c.n, 1024), 0, rebuild_gate(c), c.n, c.pos[cur], c.q[cur], c.box, c.ncell, c.colRank, c.cellOf, c.cellCount, cen, c.id[cur], nslot, c.cellStart, c.cellFill, c.boxLast, c.dFlags + 11);
    } else {
        launch<bin_body, 256>(c, cdiv(c.n, threads), 0, rebuild_gate(c), c.n, c.pos[cur], c.q[cur], c.box, c.ncell, c.colRank, c.cellOf, c.cellCount, cen, c.id[cur]);
        launch<scan_body, 1024>(c, 1, 0, rebuild_gate(c), nslot, c.cellCount, c.cellStart, c.cellFill, c.box, c.boxLast);
    }
    launch<place_body, 256>(c, cdiv(c.n, threads), 0, rebuild_gate(c), c.n, c.cellOf, c.cellFill, c.perm);
    launch<cellsort_body, 64>(c, nslot, 0, rebuild_gate(c), nslot, c.cellStart, c.perm, c.perm2, c.id[cur]);
    launch<permute_body, 256>(c, cdiv(c.n, threads), 0, rebuild_gate(c), c.n, c.perm2, c.pos[cur], c.vel[cur], c.q[cur], c.type[cur], c.id[cur], c.pos[nxt], c.vel[nxt], c.q[nxt], c.type[nxt], c.id[nxt], c.idToIdx, c.posLast, c.xq, c.box,
                              c.refreshed && c.pppm.on && c.hasCharges ? c.meshRec : nullptr, cen, c.buildPos);
    const bool excl = c.exclOff && c.maxExcl > 0;
    if (excl) {
        MDX_REQUIRE(!c.capture || c.exclIdxN == (size_t)c.n * c.maxExcl, 3,
                    "exclusion index rows must exist before a step is captured");
        if (c.exclIdxN != (size_t)c.n * c.maxExcl) {
            if (c.exclIdx) dfree(c, c.exclIdx);
            c.exclIdxN = (size_t)c.n * c.maxExcl;
            c.exclIdx = dalloc<uint32_t>(c, c.exclIdxN);
        }
    }
    if (c.capture) {
        // a captured step keeps fixed buffers: the sorted records come back,
        // with the exclusion rows and the cleared E3B oxygen map
        launch<copy_back_body, 256>(c, cdiv(c.n, threads), 0, rebuild_gate(c), c.n, c.pos[nxt], c.vel[nxt], c.q[nxt], c.type[nxt], c.id[nxt], c.pos[cur], c.vel[cur], c.q[cur], c.type[cur], c.id[cur], c.exclOff, c.exclIds, c.idToIdx, excl ? c.maxExcl : 0, c.exclIdx, c.e3b.on ? c.e3b.atomMolO : nullptr, c.rigid.on ? c.rigid.member : nullptr);
        // The P3M chain needs the sorted sites, their mesh records (the
        // permutation wrote them, the first half kick having refreshed them)
        // and the cells, not the list: its stream forks here, so the list
        // build below runs beside it.
        if (c.refreshed && c.side && c.meshStream && !c.rec && c.pppm.on) {
            MDX_HIP(hipEventRecord(c.evFork, c.stream));
            c.forked = true;
        }
    } else {
        c.cur = nxt;
        if (excl)
            launch<excl_idx_body, 256>(c, cdiv(c.n, threads), 0, nullptr, c.n, c.id[c.cur], c.exclOff, c.exclIds, c.idToIdx, c.maxExcl, c.exclIdx);
    }
    rigid_map(c);
    flex_map(c);
    if (c.e3b.on) {
        // indices were permuted: drop the old map (captured steps did so in k_copy_back)
        if (!c.capture) fill_int(c, c.e3b.atomMolO, c.n, -1, nullptr);
        const int nmolR = c.capture && c.rigid.on ? c.rigid.nmol : 0;  // see rigid_map
        mm.blocks = cdiv(std::max(c.e3b.nmol, nmolR), threads);
        mm.nmol = c.e3b.nmol;
        mm.molIds = c.e3b.molIds;
        mm.idToIdx = c.idToIdx;
        mm.molAtoms = c.e3b.molAtoms;
        mm.atomMolO = c.e3b.atomMolO;
        mm.nmolR = nmolR;
        mm.rigidIds = c.rigid.ids;
        mm.rigidAtoms = c.rigid.atoms;
        mm.member = c.rigid.member;
        // the wave build takes the map as extra workgroups; the team builds launch it first
        if (c.fuseStep && c.buildTeam != 8 && c.buildTeam != 16 && c.buildTeam != 32) {
            mapFused = true;
        } else {
            launch<mol_atoms_body, 256>(c, mm.blocks, 0, rebuild_gate(c), mm.nmol, mm.molIds, mm.idToIdx, mm.molAtoms, mm.atomMolO, mm.nmolR, mm.rigidIds, mm.rigidAtoms, mm.member);
        }
        MDX_REQUIRE(c.rlist >= (double)c.e3b.r2b + c.skin - 1e-9, 1,
                    "E3B needs rCut + padding >= 6.2 A + padding (the molecule list is taken from the "
                    "pair list)");
    }
    for (;;) {
        // a captured step keeps the running maximum of the chunk in dFlags[1..2]
        if (!c.capture) fill_int(c, c.dFlags + 1, 2, 0);
        {
            ProfScope ps(c, PK_BUILD);
            switch (c.buildTeam) {
                case 8: launch_build<8>(c, c.cur); break;
                case 16: launch_build<16>(c, c.cur); break;
                case 32: launch_build<32>(c, c.cur); break;
                default: launch_build_wave(c, c.cur, mapFused ? mm : MolMapArgs{}); break;
            }
        }
        c.waveMaxB = 0;
        if (c.e3b.on) {
            switch (c.e3b.listTeam) {
                case 8: launch_e3b_list<8>(c, c.cur); break;
                case 16: launch_e3b_list<16>(c, c.cur); break;
                default: launch_e3b_list<32>(c, c.cur); break;
            }
        }
        if (c.capture && c.fuseStep) {
            // the row maxima ride on the position refresh that follows; the
            // host checks the rows after the chunk
            c.rowMaxPending = true;
            return;
        }
        launch<row_max_body, 1024>(c, 1, 0, rebuild_gate(c), c.waveMaxA, c.waveMax, c.waveMaxB, c.waveMax + wave_slots_a(c), c.dFlags);
        if (c.capture) return;  // rows checked by the host after the chunk
        // largest row counts -> pinned host memory; checked now (sync) or at
        // the next synchronisation point (capacity_ok)
        MDX_HIP(hipMemcpyAsync(c.pinned, c.dFlags + 1, sizeof(int) * 2, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(c.pinned + 6, c.dFlags + 7, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        if (!sync) {
            MDX_HIP(hipEventRecord(c.capEvent, c.stream));
            c.capPending = true;
            break;
        }
        MDX_HIP(hipStreamSynchronize(c.stream));
        if (grow_if_needed(c)) continue;
        break;
    }
    c.nRebuilds++;
    c.listValid = true;
}

void flag7_check(int v) {
    MDX_REQUIRE(!(v & 1), 1, "an atom's stencil holds over 2^20 candidates: lower the list cutoff or the cell reach");
    MDX_REQUIRE(!(v & 2), 4,
                "a charged site's P3M stencil left the cells next to its own (a site moved beyond skin/2 "
                "between list builds)");
}

// Grows the pair / molecule rows after an overflow seen in pinned[0..1];
// returns true when something grew (the caller rebuilds).  The reference
// reallocates its list at 1.5x (GridGPU.cu:1026-1030).
bool grow_if_needed(Ctx &c) {
    flag7_check(c.pinned[6]);
    bool grew = false;
    if (c.pinned[0] > c.cap) {
        dfree(c, c.nl);
        c.cap = ((int)(c.pinned[0] * 1.25) + 15) & ~15;
        c.nl = dalloc<uint32_t>(c, (size_t)c.n * c.cap);
        grew = true;
    }
    if (c.e3b.on && c.pinned[1] > c.e3b.cap) {
        dfree(c, c.e3b.nl);
        c.e3b.cap = ((int)(c.pinned[1] * 1.25) + 15) & ~15;
        c.e3b.nl = dalloc<int>(c, (size_t)c.e3b.nmol * c.e3b.cap);
        grew = true;
    }
    return grew;
}

// True when every build since the last call fitted its rows.  Waits for the
// last asynchronous build only.
bool capacity_ok(Ctx &c) {
    if (!c.capPending) return true;
    MDX_HIP(hipEventSynchronize(c.capEvent));
    c.capPending = false;
    return !grow_if_needed(c);
}

void nlist_rebuild(Ctx &c, bool forced) {
    if (!c.nlistConfigured) nlist_setup(c);
    if (!forced) {
        int flag = 0;
        MDX_HIP(hipMemcpyAsync(&flag, c.dFlags, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        if (!flag && c.listValid) return;
    }
    run_build(c, true);
    fill_int(c, c.dFlags, 1, 0);
}

void nlist_rebuild_async(Ctx &c) {
    if (!c.nlistConfigured) nlist_setup(c);
    run_build(c, !c.listValid);
}

void fill_int(Ctx &c, int *p, int count, int v, const int *gate) {
    launch<fill_int_body, 256>(c, cdiv(count, 256), 0, gate, count, p, v);
}

namespace {
// Lag-1 rebuild decision for the step loop: per site d = |x - x_last| and the
// prediction p = d + 2 dt |v| (an M site, moved by its molecule, never moves
// further than its fastest heavy site); max d^2 and max p^2 over all sites go
// into one slot of dMax as float bits (non-negative floats order like their
// bit patterns); the other slot is cleared for the next step.  The host reads
// the slot one step later.
// Grid-stride over the sites with a workgroup reduction: one atomic pair per
// workgroup (same-address atomics serialise in L2).
// Per-workgroup maxima into part[2 b], part[2 b + 1] (float bits); a
// one-wave kernel reduces them (no same-address atomics).
__device__ __forceinline__ void check_part_body(int blk, int nblk, int n, const double4 *__restrict__ pos,
                                                    const double4 *__restrict__ last,
                                                    const double4 *__restrict__ vel, const BoxDev *__restrict__ bx,
                                                    double dt2, float *__restrict__ part) {
    const Geometry g = bx->geo;
    __shared__ float red[2][4];
    const int t0 = blk * blockDim.x + threadIdx.x;
    float d2 = 0.f, p2 = 0.f;
    for (int i = t0; i < n; i += nblk * blockDim.x) {
        const double dd = dist2(pos[i], last[i], g);
        const double4 v = vel[i];
        const double p = sqrt(dd) + dt2 * sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        d2 = fmaxf(d2, (float)dd);
        p2 = fmaxf(p2, (float)(p * p));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        d2 = fmaxf(d2, __shfl_xor(d2, o, 64));
        p2 = fmaxf(p2, __shfl_xor(p2, o, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = d2;
        red[1][w] = p2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            d2 = fmaxf(d2, red[0][k]);
            p2 = fmaxf(p2, red[1][k]);
        }
        part[2 * blk] = d2;
        part[2 * blk + 1] = p2;
    }
}

// slot = max(slot, this check) as float bits; the other slot cleared for the next step
__device__ __forceinline__ void check_reduce_body(int blk, int nblk, int nb, const float *__restrict__ part, unsigned *__restrict__ slot,
                               unsigned *__restrict__ other) {
    float d2, p2;
    part_max(nb, part, d2, p2);
    if (threadIdx.x == 0) {
        slot[0] = max(slot[0], __float_as_uint(d2));
        slot[1] = max(slot[1], __float_as_uint(p2));
        other[0] = 0u;
        other[1] = 0u;
    }
}
}  // namespace

static int check_blocks(const Ctx &c) { return std::max(1, std::min(cdiv(c.n, 256), kCheckBlocks)); }

static void check_part(Ctx &c) {
    launch<check_part_body, 256>(c, check_blocks(c), 0, c.n, c.pos[c.cur], c.posLast, c.vel[c.cur], c.box, 2.0 * c.dt, c.checkPart);
}

DecideArgs decide_args(const Ctx &c) {
    DecideArgs d{};
    d.part = c.checkPart;
    d.dmax = c.dMax;
    d.flags = c.dFlags;
    const float half = (float)(0.5 * c.skin);
    d.half2 = half * half;
    d.every = c.thermoOn ? c.thermoEvery : 0;
    d.baroEvery = c.baro.on ? c.baro.every : 0;
    d.thermo = c.fuseThermo ? c.thermo : nullptr;
    return d;
}

void skin_check(Ctx &c) { check_part(c); }

void skin_decide(Ctx &c) {
    const DecideArgs d = decide_args(c);
    const float *part = c.ringPart ? c.ringPart : d.part;
    const int nb = c.ringPart ? c.ringPartNb : check_blocks(c);
    launch<decide_body, 64>(c, 1, 0, nb, part, d.dmax, d.flags, d.half2, d.every, d.baroEvery, d.thermo);
}

void decide_step(Ctx &c) {
    // all-rigid captured steps: the rigid first half wrote the partial maxima
    // and took the decisions in its last workgroup
    if (c.decided) {
        c.decided = false;
        return;
    }
    const int fused = c.fuseCheck ? fused_check_blocks(c) : 0;
    if (!fused) check_part(c);
    const DecideArgs d = decide_args(c);
    launch<decide_body, 64>(c, 1, 0, fused ? fused : check_blocks(c), d.part, d.dmax, d.flags, d.half2, d.every, d.baroEvery, d.thermo);
}

void check_max(Ctx &c, int parity) {
    check_part(c);
    launch<check_reduce_body, 64>(c, 1, 0, check_blocks(c), c.checkPart, c.dMax + 2 * parity, c.dMax + 2 * (1 - parity));
}

}  // namespace mdx
