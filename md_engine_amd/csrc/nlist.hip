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

struct CellGrid {
    int nc[3];
    double cellInv[3];  // nc / L
    int reach;
};

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
__global__ void k_check(int n, const double4 *__restrict__ pos, const double4 *__restrict__ last,
                        Geometry g, double thr2, int *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool move = false;
    if (i < n) move = dist2(pos[i], last[i], g) > thr2;
    if (__any(move) && (threadIdx.x & 63) == 0) flags[0] = 1;
}

// Wrap into the box (periodicWrap) and bin into cells (countNumInGridCells).
__global__ void k_bin(int n, double4 *__restrict__ pos, Geometry g, CellGrid cg,
                      const int *__restrict__ colRank, int *__restrict__ cellOf,
                      int *__restrict__ cellCount) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
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
    int cx = cell_coord(p.x, g.lo[0], cg.cellInv[0], cg.nc[0]);
    int cy = cell_coord(p.y, g.lo[1], cg.cellInv[1], cg.nc[1]);
    int cz = cell_coord(p.z, g.lo[2], cg.cellInv[2], cg.nc[2]);
    int rank = colRank[cx * cg.nc[1] + cy] * cg.nc[2] + cz;
    cellOf[i] = rank;
    atomicAdd(&cellCount[rank], 1);
}

// Exclusive scan of the cell counts by one workgroup (replaces the host
// cumulativeSum round trip, helpers.h:15-24).  Also seeds the fill cursors.
__global__ void __launch_bounds__(1024) k_scan(int ncell, const int *__restrict__ count,
                                               int *__restrict__ start, int *__restrict__ fill) {
    __shared__ int wsum[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int base = 0; base < ncell; base += 1024) {
        int idx = base + threadIdx.x;
        int v = idx < ncell ? count[idx] : 0;
        int x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        if (wid == 0) {
            int s = lane < 16 ? wsum[lane] : 0;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                int y = __shfl_up(s, o, 64);
                if (lane >= o) s += y;
            }
            if (lane < 16) wsum[lane] = s;
        }
        __syncthreads();
        int excl = carry + (wid > 0 ? wsum[wid - 1] : 0) + x - v;
        if (idx < ncell) {
            start[idx] = excl;
            fill[idx] = excl;
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry = excl + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[ncell] = carry;
}

__global__ void k_place(int n, const int *__restrict__ cellOf, int *__restrict__ fill,
                        int *__restrict__ perm) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int slot = atomicAdd(&fill[cellOf[i]], 1);
    perm[slot] = i;
}

// Make the order inside each cell independent of atomic slot order: one
// workgroup per cell ranks its entries by atom id in LDS (ids are unique, so
// rank = number of smaller ids) and writes them in rank order.
__global__ void __launch_bounds__(64) k_cellsort(int ncell, const int *__restrict__ start,
                                                 const int *__restrict__ permIn,
                                                 int *__restrict__ permOut,
                                                 const int *__restrict__ id) {
    constexpr int CAP = 1024;
    __shared__ int sid[CAP];
    __shared__ int sperm[CAP];
    const int c = blockIdx.x;
    const int a = start[c], b = start[c + 1], m = b - a;
    if (m <= CAP) {
        for (int t = threadIdx.x; t < m; t += 64) {
            const int p = permIn[a + t];
            sperm[t] = p;
            sid[t] = id[p];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m; t += 64) {
            const int key = sid[t];
            int r = 0;
            for (int j = 0; j < m; ++j) r += sid[j] < key;
            permOut[a + r] = sperm[t];
        }
    } else if (threadIdx.x == 0) {
        // very crowded cell: serial insertion sort
        for (int s = a; s < b; ++s) permOut[s] = permIn[s];
        for (int s = a + 1; s < b; ++s) {
            int v = permOut[s], key = id[v], t = s - 1;
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
__global__ void k_permute(int n, const int *__restrict__ perm, const double4 *__restrict__ pos0,
                          const double4 *__restrict__ vel0, const float *__restrict__ q0,
                          const int *__restrict__ t0, const int *__restrict__ id0,
                          double4 *__restrict__ pos1, double4 *__restrict__ vel1,
                          float *__restrict__ q1, int *__restrict__ t1, int *__restrict__ id1,
                          int *__restrict__ idToIdx, double4 *__restrict__ last,
                          float4 *__restrict__ xq, Geometry g) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
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
    xq[k] = make_float4((float)(p.x - g.lo[0]), (float)(p.y - g.lo[1]), (float)(p.z - g.lo[2]), qq);
}

__global__ void k_xq(int n, const double4 *__restrict__ pos, const float *__restrict__ q,
                     float4 *__restrict__ xq, Geometry g) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 p = pos[k];
    xq[k] = make_float4((float)(p.x - g.lo[0]), (float)(p.y - g.lo[1]), (float)(p.z - g.lo[2]), q[k]);
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

// Up to kExclReg exclusions of an atom live in registers (fully unrolled, so no
// scratch); atoms with more fall back to reading them from memory.
constexpr int kExclReg = 16;

struct ExclCache {
    uint32_t e[kExclReg];
    int n;
};

__device__ __forceinline__ void excl_load(ExclCache &x, const uint32_t *__restrict__ exclIds, int e0, int e1) {
    x.n = e1 - e0;
#pragma unroll
    for (int k = 0; k < kExclReg; ++k) x.e[k] = (k < x.n) ? exclIds[e0 + k] : 0xFFFFFFFFu;
}

__device__ __forceinline__ uint32_t excl_tag(const ExclCache &x, uint32_t jid, const uint32_t *__restrict__ exclIds,
                                             int e0) {
    uint32_t tag = 0;
#pragma unroll
    for (int k = 0; k < kExclReg; ++k)
        if ((x.e[k] & 0x3FFFFFFFu) == jid && k < x.n) tag = x.e[k] >> 30;
    for (int k = kExclReg; k < x.n; ++k) {
        const uint32_t ex = exclIds[e0 + k];
        if ((ex & 0x3FFFFFFFu) == jid) tag = ex >> 30;
    }
    return tag;
}

struct ZRuns {
    int n, lo0, hi0, lo1, hi1;
};

__device__ __forceinline__ ZRuns z_runs(int cz, int ncz, int reach, int periodic) {
    ZRuns z;
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
            z.n = 2;
        }
        lo = 0;
    } else if (hi >= ncz) {
        if (periodic) {
            z.lo1 = 0;
            z.hi1 = hi - ncz;
            z.n = 2;
        }
        hi = ncz - 1;
    }
    z.lo0 = lo;
    z.hi0 = hi;
    return z;
}

// One team of T lanes per atom.  Candidates come from contiguous z-runs of
// each stencil column; hits are compacted with a ballot restricted to the
// team, so row order is (stencil column, z-run, index) - deterministic.
// FP32 pre-test on xq (x - lo in FP32): candidates with r^2 clearly inside or
// outside r_list^2 (by more than `band`, a bound on the FP32 rounding of the
// positions and of the distance) are decided in FP32; only the few inside the
// band take the canonical FP64 test, so membership is unchanged.
struct Prefilter {
    float L[3], invL[3];
    float lo2, hi2;   // rlist^2 -+ band
};

__device__ __forceinline__ float dist2f(const float4 &a, const float4 &b, const Prefilter &p) {
    float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    dx -= p.L[0] * rintf(dx * p.invL[0]);
    dy -= p.L[1] * rintf(dy * p.invL[1]);
    dz -= p.L[2] * rintf(dz * p.invL[2]);
    return dx * dx + dy * dy + dz * dz;
}

template <int T>
__global__ void __launch_bounds__(256) k_build(
    int n, const double4 *__restrict__ pos, const float4 *__restrict__ xq, const int *__restrict__ id,
    Geometry g, CellGrid cg, Prefilter pf, const int *__restrict__ colRank,
    const int *__restrict__ cellStart, double rlist2, const int *__restrict__ exclOff,
    const uint32_t *__restrict__ exclIds, int exclSpan, int cap, uint32_t *__restrict__ nl,
    int *__restrict__ nlCount, int *__restrict__ flags) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gtid / T;
    const int tl = threadIdx.x & (T - 1);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;  // whole team leaves together (T divides 64)
    const unsigned long long teamMask =
        (T == 64) ? ~0ull : (((1ull << T) - 1ull) << (lane & ~(T - 1)));
    const unsigned long long below = (1ull << lane) - 1ull;

    const double4 pi = pos[i];
    const float4 xi = xq[i];
    const int myid = id[i];
    const int e0 = exclOff ? exclOff[myid] : 0;
    const int e1 = exclOff ? exclOff[myid + 1] : 0;
    ExclCache xc;
    excl_load(xc, exclIds, e0, e1);
    const int cx = cell_coord(pi.x, g.lo[0], cg.cellInv[0], cg.nc[0]);
    const int cy = cell_coord(pi.y, g.lo[1], cg.cellInv[1], cg.nc[1]);
    const int cz = cell_coord(pi.z, g.lo[2], cg.cellInv[2], cg.nc[2]);
    const int ncz = cg.nc[2];
    const int sx = axis_slots(cg.nc[0], cg.reach), sy = axis_slots(cg.nc[1], cg.reach);
    const ZRuns zr = z_runs(cz, ncz, cg.reach, g.periodic[2]);

    uint32_t *row = nl + (size_t)i * cap;
    int cnt = 0;
    for (int a = 0; a < sx; ++a) {
        const int xx = axis_cell(a, cx, cg.nc[0], cg.reach, g.periodic[0]);
        if (xx < 0) continue;
        for (int b = 0; b < sy; ++b) {
            const int yy = axis_cell(b, cy, cg.nc[1], cg.reach, g.periodic[1]);
            if (yy < 0) continue;
            const int colBase = colRank[xx * cg.nc[1] + yy] * ncz;
            for (int r = 0; r < zr.n; ++r) {
                const int s0 = cellStart[colBase + (r ? zr.lo1 : zr.lo0)];
                const int s1 = cellStart[colBase + (r ? zr.hi1 : zr.hi0) + 1];
                for (int base = s0; base < s1; base += T) {
                    const int j = base + tl;
                    bool hit = false;
                    uint32_t entry = 0;
                    if (j < s1 && j != i) {
                        const float r2f = dist2f(xi, xq[j], pf);
                        hit = r2f < pf.lo2;
                        if (!hit && r2f <= pf.hi2) hit = dist2(pi, pos[j], g) < rlist2;
                        if (hit) {
                            uint32_t tag = 0;
                            if (e1 > e0) {
                                const int jid = id[j];
                                if (abs(jid - myid) <= exclSpan) tag = excl_tag(xc, (uint32_t)jid, exclIds, e0);
                            }
                            entry = (uint32_t)j | (tag << 30);
                        }
                    }
                    const unsigned long long m = __ballot(hit) & teamMask;
                    if (hit) {
                        const int k = cnt + __popcll(m & below);
                        if (k < cap) row[k] = entry;
                    }
                    cnt += __popcll(m);
                }
            }
        }
    }
    if (tl == 0) {
        nlCount[i] = cnt;
        atomicMax(&flags[1], cnt);
    }
}

// LDS-staged build: one workgroup per cell.  Every atom of a cell has the same
// stencil, so each stencil column's z-run of candidates is staged in LDS once
// (FP64 position + id) and each wave tests it for the cell's atoms it owns
// (atoms a0 + w, a0 + w + 4, ...; all 64 lanes on one atom).  Row order is
// (stencil column, z-run, index), the same as k_build, and membership uses the
// same canonical FP64 distance, so lists are identical.
constexpr int kStage = 512;
constexpr int kExclLds = 8;   // exclusions per atom cached in LDS (more are read from memory)

__global__ void __launch_bounds__(256) k_build_cell(
    const double4 *__restrict__ pos, const float4 *__restrict__ xq, const int *__restrict__ id,
    Geometry g, CellGrid cg, Prefilter pf, const int2 *__restrict__ colXY,
    const int *__restrict__ colRank, const int *__restrict__ cellStart, double rlist2,
    const int *__restrict__ exclOff, const uint32_t *__restrict__ exclIds, int exclSpan, int cap,
    uint32_t *__restrict__ nl, int *__restrict__ nlCount, int *__restrict__ flags) {
    __shared__ float4 sXq[kStage];
    __shared__ int sId[kStage];
    __shared__ float4 sXi[256];
    __shared__ int sMyId[256], sE0[256], sE1[256], sCnt[256];
    __shared__ uint32_t sEx[256 * kExclLds];
    const int c = blockIdx.x;
    const int ncz = cg.nc[2];
    const int cellA0 = cellStart[c], cellA1 = cellStart[c + 1];
    if (cellA1 == cellA0) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int2 cxy = colXY[c / ncz];
    const int cz = c % ncz;
    const int sx = axis_slots(cg.nc[0], cg.reach), sy = axis_slots(cg.nc[1], cg.reach);
    const ZRuns zr = z_runs(cz, ncz, cg.reach, g.periodic[2]);
    // crowded cells are handled in batches of 256 atoms, each against the whole stencil
    for (int a0 = cellA0; a0 < cellA1; a0 += 256) {
        const int nA = min(cellA1 - a0, 256);
        for (int k = threadIdx.x; k < nA; k += blockDim.x) {
            sXi[k] = xq[a0 + k];
            const int myid = id[a0 + k];
            const int e0 = exclOff ? exclOff[myid] : 0;
            const int e1 = exclOff ? exclOff[myid + 1] : 0;
            sMyId[k] = myid;
            sE0[k] = e0;
            sE1[k] = e1;
            sCnt[k] = 0;
            for (int q = 0; q < kExclLds; ++q) sEx[k * kExclLds + q] = q < e1 - e0 ? exclIds[e0 + q] : 0xFFFFFFFFu;
        }
        __syncthreads();
        for (int a = 0; a < sx; ++a) {
            const int xx = axis_cell(a, cxy.x, cg.nc[0], cg.reach, g.periodic[0]);
            if (xx < 0) continue;
            for (int b = 0; b < sy; ++b) {
                const int yy = axis_cell(b, cxy.y, cg.nc[1], cg.reach, g.periodic[1]);
                if (yy < 0) continue;
                const int colBase = colRank[xx * cg.nc[1] + yy] * ncz;
                for (int r = 0; r < zr.n; ++r) {
                    const int s0 = cellStart[colBase + (r ? zr.lo1 : zr.lo0)];
                    const int s1 = cellStart[colBase + (r ? zr.hi1 : zr.hi0) + 1];
                    for (int base = s0; base < s1; base += kStage) {
                        const int len = min(kStage, s1 - base);
                        for (int t = threadIdx.x; t < len; t += blockDim.x) {
                            sXq[t] = xq[base + t];
                            sId[t] = id[base + t];
                        }
                        __syncthreads();
                        for (int k = w; k < nA; k += 4) {
                            const int i = a0 + k;
                            const float4 xi = sXi[k];
                            const int myid = sMyId[k];
                            const int e0 = sE0[k], e1 = sE1[k];
                            uint32_t *row = nl + (size_t)i * cap;
                            int cnt = sCnt[k];
                            for (int t0 = 0; t0 < len; t0 += 64) {
                                const int t = t0 + lane;
                                bool hit = false;
                                uint32_t entry = 0;
                                if (t < len && base + t != i) {
                                    const float r2f = dist2f(xi, sXq[t], pf);
                                    hit = r2f < pf.lo2;
                                    if (!hit && r2f <= pf.hi2)
                                        hit = dist2(pos[i], pos[base + t], g) < rlist2;
                                    if (hit) {
                                        uint32_t tag = 0;
                                        const int jid = sId[t];
                                        if (e1 > e0 && abs(jid - myid) <= exclSpan) {
                                            const int ne = e1 - e0;
                                            for (int q = 0; q < min(ne, kExclLds); ++q) {
                                                const uint32_t ex = sEx[k * kExclLds + q];
                                                if ((ex & 0x3FFFFFFFu) == (uint32_t)jid) tag = ex >> 30;
                                            }
                                            for (int q = kExclLds; q < ne; ++q) {
                                                const uint32_t ex = exclIds[e0 + q];
                                                if ((ex & 0x3FFFFFFFu) == (uint32_t)jid) tag = ex >> 30;
                                            }
                                        }
                                        entry = (uint32_t)(base + t) | (tag << 30);
                                    }
                                }
                                const unsigned long long m = __ballot(hit);
                                if (hit) {
                                    const int s = cnt + __popcll(m & below);
                                    if (s < cap) row[s] = entry;
                                }
                                cnt += __popcll(m);
                            }
                            if (lane == 0) sCnt[k] = cnt;
                        }
                        __syncthreads();
                    }
                }
            }
        }
        for (int k = threadIdx.x; k < nA; k += blockDim.x) {
            nlCount[a0 + k] = sCnt[k];
            atomicMax(&flags[1], sCnt[k]);
        }
        __syncthreads();
    }
}

// E3B molecule list from the atom list: for each water, the oxygens among its
// oxygen's neighbors within (r2b + skin) at build time.  Replaces the private
// molecule GridGPU of FixE3B (FixE3B.cu:656-684) that the reference rebuilds
// every step; with the skin criterion the per-step r2b set is a subset.
template <int T>
__global__ void __launch_bounds__(256) k_e3b_list(
    int nmol, const int4 *__restrict__ molAtoms, const int *__restrict__ atomMolO,
    const double4 *__restrict__ pos, Geometry g, double rcut2, int cap, const uint32_t *__restrict__ nl,
    const int *__restrict__ nlCount, int mcap, int *__restrict__ mnl, int *__restrict__ mcount,
    int *__restrict__ flags) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gtid / T;
    const int tl = threadIdx.x & (T - 1);
    const int lane = threadIdx.x & 63;
    if (m >= nmol) return;
    const unsigned long long teamMask =
        (T == 64) ? ~0ull : (((1ull << T) - 1ull) << (lane & ~(T - 1)));
    const unsigned long long below = (1ull << lane) - 1ull;
    const int oi = molAtoms[m].x;
    const double4 pi = pos[oi];
    const int cnt_i = min(nlCount[oi], cap);
    const uint32_t *row = nl + (size_t)oi * cap;
    int cnt = 0;
    for (int base = 0; base < cnt_i; base += T) {
        const int k = base + tl;
        bool hit = false;
        int mj = -1;
        if (k < cnt_i) {
            const int j = (int)(row[k] & 0x3FFFFFFFu);
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
    if (tl == 0) {
        mcount[m] = cnt;
        atomicMax(&flags[2], cnt);
    }
}

__global__ void k_mol_atoms(int nmol, const int4 *__restrict__ molIds, const int *__restrict__ idToIdx,
                            int4 *__restrict__ molAtoms, int *__restrict__ atomMolO) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    int4 ids = molIds[m];
    int4 a = make_int4(idToIdx[ids.x], idToIdx[ids.y], idToIdx[ids.z], idToIdx[ids.w]);
    molAtoms[m] = a;
    atomMolO[a.x] = m;
}

__global__ void k_fill_int(int n, int *__restrict__ p, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

uint32_t morton2(uint32_t x, uint32_t y) {
    uint64_t r = 0;
    for (int b = 0; b < 16; ++b) {
        r |= (uint64_t)((x >> b) & 1u) << (2 * b);
        r |= (uint64_t)((y >> b) & 1u) << (2 * b + 1);
    }
    return (uint32_t)r;
}

CellGrid make_grid(const Ctx &c) {
    CellGrid cg;
    cg.reach = c.reach;
    for (int a = 0; a < 3; ++a) {
        cg.nc[a] = c.nc[a];
        cg.cellInv[a] = (double)c.nc[a] / c.geo.L[a];
    }
    return cg;
}

// Band that covers the FP32 error of r^2 for wrapped positions (x - lo in
// [0, L]) and r <= r_list.  Per component the FP32 difference is off by at
// most ~5 L 2^-24 (two rounded inputs, the subtraction, the rounded L and the
// image shift); r^2 is then off by at most 2 sqrt(3) r_list of that plus a few
// ulp of r^2 for the squares and the sum.  The band takes 16 L 2^-24 and
// 32 r^2 2^-24, about three times both.
Prefilter make_prefilter(const Ctx &c) {
    Prefilter p;
    double lmax = 0.0;
    for (int a = 0; a < 3; ++a) {
        p.L[a] = (float)c.geo.L[a];
        p.invL[a] = c.geo.periodic[a] ? (float)(1.0 / c.geo.L[a]) : 0.0f;
        lmax = std::max(lmax, c.geo.L[a]);
    }
    const double r2 = c.rlist * c.rlist;
    const double eps = std::ldexp(1.0, -23);
    const double band = 8.0 * 2.0 * c.rlist * 1.01 * lmax * eps * std::sqrt(3.0) + 16.0 * r2 * eps;
    p.lo2 = (float)(r2 - band);
    p.hi2 = (float)(r2 + band);
    // an open axis leaves positions unwrapped (no bound on |x - lo|): every
    // candidate then takes the FP64 test
    if (!c.geo.periodic[0] || !c.geo.periodic[1] || !c.geo.periodic[2]) {
        p.lo2 = -1.0f;
        p.hi2 = INFINITY;
    }
    return p;
}

template <int T>
void launch_build(Ctx &c, const CellGrid &cg, int cur) {
    int threads = 256;
    int blocks = cdiv((long long)c.n * T, threads);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_build<T>), dim3(blocks), dim3(threads), 0, c.stream, c.n,
                       c.pos[cur], c.xq, c.id[cur], c.geo, cg, make_prefilter(c), c.colRank,
                       c.cellStart, c.rlist * c.rlist, c.exclOff, c.exclIds, c.exclSpan, c.cap, c.nl,
                       c.nlCount, c.dFlags);
}

template <int T>
void launch_e3b_list(Ctx &c, int cur) {
    int threads = 256;
    int blocks = cdiv((long long)c.e3b.nmol * T, threads);
    double rc = (double)c.e3b.r2b + c.skin;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_e3b_list<T>), dim3(blocks), dim3(threads), 0, c.stream,
                       c.e3b.nmol, c.e3b.molAtoms, c.e3b.atomMolO, c.pos[cur], c.geo, rc * rc, c.cap,
                       c.nl, c.nlCount, c.e3b.cap, c.e3b.nl, c.e3b.nlCount, c.dFlags);
}

}  // namespace

void nlist_setup(Ctx &c) {
    MDX_REQUIRE(c.n > 0, 3, "system not set");
    MDX_REQUIRE(c.rlist > 0, 1, "neighbor list cutoff must be positive");
    for (int a = 0; a < 3; ++a) {
        if (c.geo.periodic[a])
            MDX_REQUIRE(c.rlist < 0.5 * c.geo.L[a], 1,
                        "neighbor cutoff (rCut + padding) must be below half the box length");
        double edge = c.rlist / c.reach * (1.0 + 1e-9);
        int nca = (int)std::floor(c.geo.L[a] / edge);
        c.nc[a] = std::max(1, nca);
    }
    int ncell = c.nc[0] * c.nc[1] * c.nc[2];
    if (ncell != c.ncell) {
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
        c.cellCount = dalloc<int>(c, ncell);
        c.cellStart = dalloc<int>(c, ncell + 1);
        c.cellFill = dalloc<int>(c, ncell);
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
    // team width: next power of two of the expected z-run population
    double dens = c.n / (c.geo.L[0] * c.geo.L[1] * c.geo.L[2]);
    double run = dens * (c.geo.L[0] / c.nc[0]) * (c.geo.L[1] / c.nc[1]) * (c.geo.L[2] / c.nc[2]) *
                 std::min(c.nc[2], 2 * c.reach + 1);
    int T = 8;
    while (T < 64 && T < run) T *= 2;
    if (!c.buildTeamUser) c.buildTeam = T;
    if (c.cap == 0) {
        double expect = dens * 4.18879020479 * c.rlist * c.rlist * c.rlist;
        c.cap = ((int)(expect * 1.3) + 32 + 15) & ~15;
        c.nl = dalloc<uint32_t>(c, (size_t)c.n * c.cap);
        c.nlCount = dalloc<int>(c, c.n);
    }
    c.nlistConfigured = true;
}

void refresh_xq(Ctx &c) {
    int threads = 256;
    hipLaunchKernelGGL(k_xq, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.pos[c.cur],
                       c.q[c.cur], c.xq, c.geo);
}

void nlist_check(Ctx &c) {
    int threads = 256;
    double thr = 0.5 * c.skin;
    hipLaunchKernelGGL(k_check, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n,
                       c.pos[c.cur], c.posLast, c.geo, thr * thr, c.dFlags);
}

static void run_build(Ctx &c, bool sync) {
    CellGrid cg = make_grid(c);
    int threads = 256;
    int cur = c.cur, nxt = 1 - c.cur;
    hipLaunchKernelGGL(k_fill_int, dim3(cdiv(c.ncell, threads)), dim3(threads), 0, c.stream, c.ncell,
                       c.cellCount, 0);
    hipLaunchKernelGGL(k_bin, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.pos[cur],
                       c.geo, cg, c.colRank, c.cellOf, c.cellCount);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, c.stream, c.ncell, c.cellCount, c.cellStart,
                       c.cellFill);
    hipLaunchKernelGGL(k_place, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.cellOf,
                       c.cellFill, c.perm);
    hipLaunchKernelGGL(k_cellsort, dim3(c.ncell), dim3(64), 0, c.stream, c.ncell, c.cellStart, c.perm,
                       c.perm2, c.id[cur]);
    hipLaunchKernelGGL(k_permute, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.perm2,
                       c.pos[cur], c.vel[cur], c.q[cur], c.type[cur], c.id[cur], c.pos[nxt],
                       c.vel[nxt], c.q[nxt], c.type[nxt], c.id[nxt], c.idToIdx, c.posLast, c.xq,
                       c.geo);
    c.cur = nxt;
    rigid_map(c);
    if (c.e3b.on) {
        fill_int(c, c.e3b.atomMolO, c.n, -1);  // indices were permuted: drop the old map
        hipLaunchKernelGGL(k_mol_atoms, dim3(cdiv(c.e3b.nmol, threads)), dim3(threads), 0, c.stream,
                           c.e3b.nmol, c.e3b.molIds, c.idToIdx, c.e3b.molAtoms, c.e3b.atomMolO);
        MDX_REQUIRE(c.rlist >= (double)c.e3b.r2b + c.skin - 1e-9, 1,
                    "E3B needs rCut + padding >= 6.2 A + padding (the molecule list is taken from the "
                    "pair list)");
    }
    for (;;) {
        fill_int(c, c.dFlags + 1, 2, 0);
        {
            ProfScope ps(c, PK_BUILD);
            const double perCell = (double)c.n / c.ncell;
            if (c.buildMode == 1 || (c.buildMode == 0 && perCell >= 8.0)) {
                hipLaunchKernelGGL(k_build_cell, dim3(c.ncell), dim3(256), 0, c.stream, c.pos[c.cur],
                                   c.xq, c.id[c.cur], c.geo, cg, make_prefilter(c), c.colXY, c.colRank,
                                   c.cellStart, c.rlist * c.rlist, c.exclOff, c.exclIds, c.exclSpan,
                                   c.cap, c.nl, c.nlCount, c.dFlags);
            } else {
                switch (c.buildTeam) {
                    case 8: launch_build<8>(c, cg, c.cur); break;
                    case 16: launch_build<16>(c, cg, c.cur); break;
                    case 32: launch_build<32>(c, cg, c.cur); break;
                    default: launch_build<64>(c, cg, c.cur); break;
                }
            }
        }
        if (c.e3b.on) {
            switch (c.e3b.team) {
                case 8: launch_e3b_list<8>(c, c.cur); break;
                case 16: launch_e3b_list<16>(c, c.cur); break;
                default: launch_e3b_list<32>(c, c.cur); break;
            }
        }
        // largest row counts -> pinned host memory; checked now (sync) or at
        // the next synchronisation point (capacity_ok)
        MDX_HIP(hipMemcpyAsync(c.pinned, c.dFlags + 1, sizeof(int) * 2, hipMemcpyDeviceToHost, c.stream));
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

// Grows the pair / molecule rows after an overflow seen in pinned[0..1];
// returns true when something grew (the caller rebuilds).  The reference
// reallocates its list at 1.5x (GridGPU.cu:1026-1030).
bool grow_if_needed(Ctx &c) {
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

void fill_int(Ctx &c, int *p, int count, int v) {
    hipLaunchKernelGGL(k_fill_int, dim3(cdiv(count, 256)), dim3(256), 0, c.stream, count, p, v);
}

namespace {
// Lag-1 rebuild decision for the step loop: per site d = |x - x_last| and the
// prediction p = d + 2 dt |v| (an M site, moved by its molecule, never moves
// further than its fastest heavy site); max d^2 and max p^2 over all sites go
// into one slot of dMax as float bits (non-negative floats order like their
// bit patterns); the other slot is cleared for the next step.  The host reads
// the slot one step later.
__global__ void k_check_max(int n, const double4 *__restrict__ pos, const double4 *__restrict__ last,
                            const double4 *__restrict__ vel, Geometry g, double dt2,
                            unsigned *__restrict__ slot, unsigned *__restrict__ other) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        other[0] = 0u;
        other[1] = 0u;
    }
    float d2 = 0.f, p2 = 0.f;
    if (i < n) {
        const double dd = dist2(pos[i], last[i], g);
        const double4 v = vel[i];
        const double p = sqrt(dd) + dt2 * sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        d2 = (float)dd;
        p2 = (float)(p * p);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        d2 = fmaxf(d2, __shfl_xor(d2, o, 64));
        p2 = fmaxf(p2, __shfl_xor(p2, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&slot[0], __float_as_uint(d2));
        atomicMax(&slot[1], __float_as_uint(p2));
    }
}
}  // namespace

void check_max(Ctx &c, int parity) {
    int threads = 256;
    hipLaunchKernelGGL(k_check_max, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n,
                       c.pos[c.cur], c.posLast, c.vel[c.cur], c.geo, 2.0 * c.dt, c.dMax + 2 * parity,
                       c.dMax + 2 * (1 - parity));
}

}  // namespace mdx
