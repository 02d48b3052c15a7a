// Deserno-Holm P3M long-range electrostatics.  Replaces the mesh part of
// FixChargeEwald (src/Fixes/FixChargeEwald.cu): charge spread
// map_charge_to_grid_order_{1,3,5,7}_cu (80-280), Green_function_cu (314-366),
// E_field_cu (387-434), the gather Ewald_long_range_forces_order_*_cu
// (437-697), Energy_cu / singlePointEng (705-728, 1481-1611), virials_cu
// (731-802) and the alpha secant solve (853-967).
//
// Design (gfx950):
//   * spread is deterministic: each weight is added to a 64-bit fixed-point
//     mesh (value * 2^40) with integer atomics, then one pass converts the mesh
//     to FP32 and clears it for the next step;
//   * real-to-complex FFTs (hipFFT) on the half spectrum replace the C2C
//     transforms; the field is i k G rho_hat with the component's k set to 0 on
//     its own Nyquist plane, which is exactly the part Re(C2C^-1) keeps in the
//     reference, so C2R reproduces its field;
//   * the three field components go through one batched C2R plan;
//   * zero-charge sites (water oxygens) are skipped in spread and gather: their
//     contributions are exactly zero.
#include <cmath>

#include "mdx_internal.h"

namespace mdx {

namespace {

// t / n for 0 <= t < 2^16 and 0 < n <= 2^10 from a float reciprocal: the
// quotient of (t + 0.5) / n lies at least 0.5 / n from an integer, far above
// the reciprocal's rounding, so truncation gives the exact integer quotient.
__device__ __forceinline__ int fft_div(int t, float rn) { return (int)(((float)t + 0.5f) * rn); }

constexpr double kPi = 3.14159265358979323846;

// W_p for orders 1/3/5/7 (FixChargeEwald.cu:111-115, 163-172, 220-234) as
// polynomials in the offset x, coefficients in ascending powers, evaluated by
// Horner's rule with explicit fused multiply-adds: the stencil index may be a
// compile-time constant (weights) or a run-time one (weight_at, coefficients
// copied to LDS), and both give the same bits.
template <int ORDER>
struct WCoef;
template <>
struct WCoef<1> {
    static constexpr float c[1][1] = {{1.0f}};
};
template <>
struct WCoef<3> {
    static constexpr float c[3][3] = {{0.125f, -0.5f, 0.5f}, {0.75f, 0.0f, -1.0f}, {0.125f, 0.5f, 0.5f}};
};
#define MDX_W(d, a, b, c_, e, f) {(float)(a / d), (float)(b / d), (float)(c_ / d), (float)(e / d), (float)(f / d)}
template <>
struct WCoef<5> {
    static constexpr float c[5][5] = {MDX_W(384.0, 1.0, -8.0, 24.0, -32.0, 16.0),
                                      MDX_W(96.0, 19.0, -44.0, 24.0, 16.0, -16.0),
                                      MDX_W(192.0, 115.0, 0.0, -120.0, 0.0, 48.0),
                                      MDX_W(96.0, 19.0, 44.0, 24.0, -16.0, -16.0),
                                      MDX_W(384.0, 1.0, 8.0, 24.0, 32.0, 16.0)};
};
#undef MDX_W
#define MDX_W(d, a, b, c_, e, f, g, h) \
    {(float)(a / d), (float)(b / d), (float)(c_ / d), (float)(e / d), (float)(f / d), (float)(g / d), (float)(h / d)}
template <>
struct WCoef<7> {
    static constexpr float c[7][7] = {MDX_W(46080.0, 1.0, -12.0, 60.0, -160.0, 240.0, -192.0, 64.0),
                                      MDX_W(23040.0, 361.0, -1416.0, 2220.0, -1600.0, 240.0, 384.0, -192.0),
                                      MDX_W(46080.0, 10543.0, -17340.0, 4740.0, 6880.0, -4080.0, -960.0, 960.0),
                                      MDX_W(11520.0, 5887.0, 0.0, -4620.0, 0.0, 1680.0, 0.0, -320.0),
                                      MDX_W(46080.0, 10543.0, 17340.0, 4740.0, -6880.0, -4080.0, 960.0, 960.0),
                                      MDX_W(23040.0, 361.0, 1416.0, 2220.0, 1600.0, 240.0, -384.0, -192.0),
                                      MDX_W(46080.0, 1.0, 12.0, 60.0, 160.0, 240.0, 192.0, 64.0)};
};
#undef MDX_W

template <int ORDER>
__device__ __forceinline__ float horner(const float *c, float x) {
    float w = c[ORDER - 1];
#pragma unroll
    for (int m = ORDER - 2; m >= 0; --m) w = __builtin_fmaf(w, x, c[m]);
    return w;
}

template <int ORDER>
__device__ __forceinline__ void weights(float x, float *w) {
#pragma unroll
    for (int j = 0; j < ORDER; ++j) w[j] = horner<ORDER>(WCoef<ORDER>::c[j], x);
}

// weight of run-time stencil index j from the coefficient table in LDS
template <int ORDER>
__device__ __forceinline__ float weight_at(const float *lcoef, int j, float x) {
    return horner<ORDER>(lcoef + j * ORDER, x);
}

// weight of stencil index k (0 <= k < ORDER) without a dynamically indexed
// local array (which would live in scratch memory)
template <int ORDER>
__device__ __forceinline__ float pick(const float (&w)[ORDER], int k) {
    float r = w[0];
#pragma unroll
    for (int t = 1; t < ORDER; ++t) r = k == t ? w[t] : r;
    return r;
}

// v in (-2n, 2n) to [0, n) without an integer division
__device__ __forceinline__ int wrap_near(int v, int n) {
    v += v < 0 ? n : 0;
    v += v < 0 ? n : 0;
    return v >= n ? v - n : v;
}

__device__ __forceinline__ int wrapi(int p, int n) {
    p %= n;
    return p < 0 ? p + n : p;
}

template <int ORDER>
__device__ __forceinline__ void spread_body(int blk, int nblk, int n, const double4 *__restrict__ pos,
                                                const float *__restrict__ q, const BoxDev *__restrict__ bx,
                                                float qunit, unsigned long long *__restrict__ mesh) {
    const MeshGeo m = bx->mesh;
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float qi = q[i];
    if (qi == 0.0f) return;
    const double4 p = pos[i];
    int g[3];
    float d[3];
    mesh_point(p.x, 0, m, g[0], d[0]);
    mesh_point(p.y, 1, m, g[1], d[1]);
    mesh_point(p.z, 2, m, g[2], d[2]);
    float wx[ORDER], wy[ORDER], wz[ORDER];
    weights<ORDER>(d[0], wx);
    weights<ORDER>(d[1], wy);
    weights<ORDER>(d[2], wz);
    constexpr int H = (ORDER - 1) / 2;
    const float qq = qunit * qi;
#pragma unroll
    for (int ix = 0; ix < ORDER; ++ix) {
        const int px = wrapi(g[0] + ix - H, m.sz[0]);
        const float a = qq * wx[ix];
#pragma unroll
        for (int iy = 0; iy < ORDER; ++iy) {
            const int py = wrapi(g[1] + iy - H, m.sz[1]);
            const float b = a * wy[iy];
            const size_t rowb = ((size_t)px * m.sz[1] + py) * m.sz[2];
#pragma unroll
            for (int iz = 0; iz < ORDER; ++iz) {
                const int pz = wrapi(g[2] + iz - H, m.sz[2]);
                const long long v = __double2ll_rn((double)(b * wz[iz]) * kMeshScale);
                atomicAdd(&mesh[rowb + pz], (unsigned long long)v);
            }
        }
    }
}

// LDS-tiled spread.  One workgroup per (cell column, chunk of zs cells in z):
// its atoms are contiguous in the cell-ordered arrays and, with positions
// within skin/2 of their cell since the last build, every stencil point lies in
// a box of mesh points known from the cell geometry.  Weights are added to that
// box in LDS (64-bit integer atomics, order independent), and the box is then
// added to the global mesh with one atomic per point whose addresses are
// consecutive across a wave.  Rows that leave the box (never, unless an atom
// drifted further than the margin) go straight to global memory.
struct TileGeo {  // cell widths and margin come from the device box
    int ncz, zs, nzc;
    int ext[3];
    // cb x cb cell columns a tile (1 or 2): the columns are in 2-D Morton
    // order, so with even column counts the four columns of an aligned 2 x 2
    // block hold ranks 4k..4k+3, rank 4k its lowest corner
    int cb;
};

__device__ __forceinline__ int tile_origin(double lo_pos, double h) {
    // same rounding as mesh_point on the lowest reachable coordinate
    return (int)floor((lo_pos + 0.5 * h) / h);
}

template <int ORDER>
__device__ __forceinline__ void spread_tiled_body(int blk, int nblk, const double4 *__restrict__ pos,
                                                      const float *__restrict__ q, const BoxDev *__restrict__ bx,
                                                      float qunit, const int *__restrict__ cellStart,
                                                      const int2 *__restrict__ colXY, TileGeo tg, int nlog,
                                                      unsigned long long *__restrict__ mesh,
                                                      const float4 *__restrict__ rec) {
    extern __shared__ unsigned long long tile[];
    const int lb = xcd_block(blk, nlog);
    if (lb >= nlog) return;  // uniform over the workgroup
    const MeshGeo m = bx->mesh;
    constexpr int H = (ORDER - 1) / 2;
    const int col = lb / tg.nzc;
    const int zc = lb % tg.nzc;
    const int z0 = zc * tg.zs;
    const int z1 = min(z0 + tg.zs, tg.ncz);
    const int a0 = cellStart[col * tg.ncz + z0];
    const int a1 = cellStart[col * tg.ncz + z1];
    const int2 cxy = colXY[col];
    const double margin = bx->tileMargin;
    const int ox = tile_origin(cxy.x * bx->tileCellW[0] - margin, m.h[0]) - H;
    const int oy = tile_origin(cxy.y * bx->tileCellW[1] - margin, m.h[1]) - H;
    const int oz = tile_origin(z0 * bx->tileCellW[2] - margin, m.h[2]) - H;
    const int ex = tg.ext[0], ey = tg.ext[1], ez = tg.ext[2];
    const int vol = ex * ey * ez;
    for (int k = threadIdx.x; k < vol; k += blockDim.x) tile[k] = 0ull;
    __syncthreads();
    const int items = (a1 - a0) * ORDER;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
        const int i = a0 + it / ORDER;
        const int ix = it % ORDER;
        const float qi = q[i];
        int g[3];
        float d[3];
        if (rec) {  // the position refresh's mesh point and offsets, loaded with the charge
            const float4 rc = rec[i];
            if (qi == 0.0f) continue;
            g[0] = rec_point(rc, 0);
            g[1] = rec_point(rc, 1);
            g[2] = rec_point(rc, 2);
            d[0] = rc.x;
            d[1] = rc.y;
            d[2] = rc.z;
        } else {
            if (qi == 0.0f) continue;
            const double4 p = pos[i];
            mesh_point(p.x, 0, m, g[0], d[0]);
            mesh_point(p.y, 1, m, g[1], d[1]);
            mesh_point(p.z, 2, m, g[2], d[2]);
        }
        float wx[ORDER], wy[ORDER], wz[ORDER];
        weights<ORDER>(d[0], wx);
        weights<ORDER>(d[1], wy);
        weights<ORDER>(d[2], wz);
        const float a = qunit * qi * wx[ix];
        const int rx = wrapi(g[0] - H + ix - ox, m.sz[0]);
        const int ry = wrapi(g[1] - H - oy, m.sz[1]);
        const int rz = wrapi(g[2] - H - oz, m.sz[2]);
        if (rx < ex && ry + ORDER <= ey && rz + ORDER <= ez) {
#pragma unroll
            for (int iy = 0; iy < ORDER; ++iy) {
                const float b = a * wy[iy];
                const int rb = (rx * ey + ry + iy) * ez + rz;
#pragma unroll
                for (int iz = 0; iz < ORDER; ++iz) {
                    const long long v = __double2ll_rn((double)(b * wz[iz]) * kMeshScale);
                    atomicAdd(&tile[rb + iz], (unsigned long long)v);
                }
            }
        } else {
            const int px = wrapi(g[0] + ix - H, m.sz[0]);
#pragma unroll
            for (int iy = 0; iy < ORDER; ++iy) {
                const int py = wrapi(g[1] + iy - H, m.sz[1]);
                const float b = a * wy[iy];
                const size_t rowb = ((size_t)px * m.sz[1] + py) * m.sz[2];
#pragma unroll
                for (int iz = 0; iz < ORDER; ++iz) {
                    const int pz = wrapi(g[2] + iz - H, m.sz[2]);
                    const long long v = __double2ll_rn((double)(b * wz[iz]) * kMeshScale);
                    atomicAdd(&mesh[rowb + pz], (unsigned long long)v);
                }
            }
        }
    }
    __syncthreads();
    // flush: a wave takes whole z rows of the box (as many as fit in its 64
    // lanes), so the row's coordinates are computed once per row and the
    // atomics of a row go to consecutive addresses
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int per = 64 / ez;  // rows per wave pass (ez <= 64)
    const int sub = lane / ez, kz = lane - sub * ez;
    const int gz = wrap_near(oz + kz, m.sz[2]);
    for (int row = (threadIdx.x >> 6) * per + sub; sub < per && row < ex * ey; row += nw * per) {
        const int kx = row / ey, ky = row - kx * ey;
        const unsigned long long v = tile[row * ez + kz];
        if (v == 0ull) continue;
        const int gx = wrap_near(ox + kx, m.sz[0]), gy = wrap_near(oy + ky, m.sz[1]);
        atomicAdd(&mesh[((size_t)gx * m.sz[1] + gy) * m.sz[2] + gz], v);
    }
}

// First mesh point of cell k along an axis of sz points and nc cells,
// ceil(k sz / nc), for k >= -nc: point p belongs to cell floor(p nc / sz), so
// the cells partition the mesh by integer arithmetic alone, whatever the box.
__host__ __device__ __forceinline__ int own_start(int k, int sz, int nc) {
    return k >= 0 ? (k * sz + nc - 1) / nc : -((-k * sz) / nc);
}

// Owner-computes spread.  A workgroup owns the mesh points of zs cells of one
// cell column and sums into LDS every contribution to them: from the charged
// sites of its own cells and of the cells within reach (one column either
// side in x and y, rz cells in z), each site's mesh point and offsets taken
// from the record the position refresh wrote.  A site stays within skin/2 of
// the cell it was sorted into (the list's own condition for a rebuild), so
// its stencil lies in the cells within reach: the host checks that from the
// geometry and every workgroup checks it for its own sites (bit 1 of
// flags[7] otherwise).  The workgroup then writes its points to the real mesh
// in FP32: no global atomics and no pass converting and clearing a
// fixed-point mesh.  The sums are the same 64-bit integers as the other
// spreads', so the real mesh is the same bit for bit.
//
// Pass 1 (each wave takes whole candidate runs, a lane a site): the sites
// whose stencil meets the owned block are listed with one item per x plane
// that meets it; pass 2 (a thread an item): the plane's 25 (order 5) points,
// those outside the block masked.  Candidates are taken kOwnCap at a time in
// their run order, so crowded cells only add rounds.
struct OwnGeo {
    int nc[3];      // cells per axis
    int sz[3];      // mesh points per axis
    int zs, nzc;    // z cells per workgroup, workgroups per column
    int rz;         // z reach in cells (x and y reach one column)
    int whole;      // zs + 2 rz >= nc[2]: the z range is the whole column
};

constexpr int kOwnRanges = 18;  // 9 columns x at most 2 z runs
constexpr int kOwnCap = 1024;   // candidates per round
constexpr int kOwnZMax = 64;    // z cells per workgroup at most

template <int ORDER, int NT>
__device__ __forceinline__ void spread_own_body(int blk, int nblk, const float4 *__restrict__ rec,
                                                    const float *__restrict__ q, const int *__restrict__ cellStart,
                                                    const int *__restrict__ colRank, const int2 *__restrict__ colXY,
                                                    OwnGeo og, float qunit, int nlog, float *__restrict__ real,
                                                    int *__restrict__ flags) {
    extern __shared__ unsigned long long own[];
    __shared__ int rS[kOwnRanges], rP[kOwnRanges + 1];  // run start, candidate prefix
    __shared__ float4 eRec[kOwnCap];
    __shared__ float eQ[kOwnCap];
    __shared__ unsigned short items[kOwnCap * ORDER];  // listed site << 3 | x plane
    __shared__ float lcoef[ORDER * ORDER];
    __shared__ int nE, nI;
    const int lb = xcd_block(blk, nlog);
    if (lb >= nlog) return;  // uniform over the workgroup
    constexpr int H = (ORDER - 1) / 2;
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col = lb / og.nzc, zc = lb % og.nzc;
    const int ncz = og.nc[2];
    const int z0 = zc * og.zs, z1 = min(z0 + og.zs, ncz);
    const int2 cxy = colXY[col];
    const int X0 = own_start(cxy.x, og.sz[0], og.nc[0]), nx = own_start(cxy.x + 1, og.sz[0], og.nc[0]) - X0;
    const int Y0 = own_start(cxy.y, og.sz[1], og.nc[1]), ny = own_start(cxy.y + 1, og.sz[1], og.nc[1]) - Y0;
    const int Z0 = own_start(z0, og.sz[2], ncz), nz = own_start(z1, og.sz[2], ncz) - Z0;
    const int vol = nx * ny * nz;
    for (int k = threadIdx.x; k < vol; k += NT) own[k] = 0ull;
    if (threadIdx.x < ORDER * ORDER) lcoef[threadIdx.x] = WCoef<ORDER>::c[threadIdx.x / ORDER][threadIdx.x % ORDER];
    __shared__ int rN[kOwnRanges], ownZ[kOwnZMax + 1];
    if (threadIdx.x < 9) {
        // candidate runs: 9 columns, each a z range that wraps at one end at most
        const int t = threadIdx.x;
        int cx = cxy.x + t / 3 - 1, cy = cxy.y + t % 3 - 1;
        cx += cx < 0 ? og.nc[0] : (cx >= og.nc[0] ? -og.nc[0] : 0);
        cy += cy < 0 ? og.nc[1] : (cy >= og.nc[1] ? -og.nc[1] : 0);
        const int base = colRank[cx * og.nc[1] + cy] * ncz;
        int s0, e0, s1 = 0, e1 = 0;
        if (og.whole) {
            s0 = 0;
            e0 = ncz;
        } else {
            const int zlo = z0 - og.rz, zhi = z1 + og.rz;
            if (zlo < 0) {
                s0 = zlo + ncz; e0 = ncz; s1 = 0; e1 = zhi;
            } else if (zhi > ncz) {
                s0 = zlo; e0 = ncz; s1 = 0; e1 = zhi - ncz;
            } else {
                s0 = zlo; e0 = zhi;
            }
        }
        const int a0 = cellStart[base + s0], b0 = cellStart[base + e0];
        const int a1 = cellStart[base + s1], b1 = cellStart[base + e1];
        rS[2 * t] = a0;
        rN[2 * t] = b0 - a0;
        rS[2 * t + 1] = a1;
        rN[2 * t + 1] = b1 - a1;
    } else if (threadIdx.x - 9 <= z1 - z0) {
        // the own column's cell starts, for the own sites' cells
        ownZ[threadIdx.x - 9] = cellStart[col * ncz + z0 + threadIdx.x - 9];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int pre = 0;
        for (int r = 0; r < kOwnRanges; ++r) {
            rP[r] = pre;
            pre += rN[r];
        }
        rP[kOwnRanges] = pre;
    }
    __syncthreads();
    const int total = rP[kOwnRanges];
    // this workgroup's own sites and the x / y span their stencils must stay in
    const int ownA = ownZ[0], ownB = ownZ[z1 - z0];
    const int sxl = own_start(cxy.x - 1, og.sz[0], og.nc[0]), sxn = own_start(cxy.x + 2, og.sz[0], og.nc[0]) - sxl;
    const int syl = own_start(cxy.y - 1, og.sz[1], og.nc[1]), syn = own_start(cxy.y + 2, og.sz[1], og.nc[1]) - syl;
    for (int w0 = 0; w0 < total; w0 += kOwnCap) {
        if (threadIdx.x == 0) {
            nE = 0;
            nI = 0;
        }
        __syncthreads();
        // pass 1: candidates f in [w0, w0 + kOwnCap) of the flattened runs
        for (int r = wv; r < kOwnRanges; r += NW) {
            const int f0 = max(rP[r], w0), f1 = min(rP[r + 1], w0 + kOwnCap);
            for (int f = f0 + lane; f < f1; f += 64) {
                const int i = rS[r] + f - rP[r];
                const float4 rc = rec[i];
                const float qi = q[i];
                const int gx = rec_point(rc, 0), gy = rec_point(rc, 1), gz = rec_point(rc, 2);
                unsigned xm = 0;
                bool ym = false, zm = false;
#pragma unroll
                for (int k = 0; k < ORDER; ++k) {
                    xm |= (wrap_near(gx - H + k - X0, og.sz[0]) < nx ? 1u : 0u) << k;
                    ym = ym || wrap_near(gy - H + k - Y0, og.sz[1]) < ny;
                    zm = zm || wrap_near(gz - H + k - Z0, og.sz[2]) < nz;
                }
                if (xm && ym && zm && qi != 0.0f) {
                    const int e = atomicAdd(&nE, 1);
                    eRec[e] = rc;
                    eQ[e] = qi;
                    int it = atomicAdd(&nI, __popc(xm));
#pragma unroll
                    for (int k = 0; k < ORDER; ++k)
                        if (xm >> k & 1u) items[it++] = (unsigned short)(e << 3 | k);
                }
                if (i >= ownA && i < ownB) {
                    // the site's own cell in z, then its stencil against the cells within reach
                    int zo = z0;
                    for (int z = 1; z < z1 - z0; ++z) zo += ownZ[z] <= i ? 1 : 0;
                    const int szl = own_start(zo - og.rz, og.sz[2], ncz);
                    const int szn = own_start(zo + 1 + og.rz, og.sz[2], ncz) - szl;
                    const bool inx = wrap_near(gx - H - sxl, og.sz[0]) + ORDER <= sxn;
                    const bool iny = wrap_near(gy - H - syl, og.sz[1]) + ORDER <= syn;
                    const bool inz = og.whole || wrap_near(gz - H - szl, og.sz[2]) + ORDER <= szn;
                    if (!(inx && iny && inz)) atomicOr(&flags[7], 2);
                }
            }
        }
        __syncthreads();
        // pass 2: one listed x plane per thread
        const int ni = nI;
        for (int t = threadIdx.x; t < ni; t += NT) {
            const int e = items[t] >> 3, ix = items[t] & 7;
            const float4 rc = eRec[e];
            const float a = qunit * eQ[e] * weight_at<ORDER>(lcoef, ix, rc.x);
            const int lx = wrap_near(rec_point(rc, 0) - H + ix - X0, og.sz[0]);
            const int yb = rec_point(rc, 1) - H - Y0, zb = rec_point(rc, 2) - H - Z0;
            float wy[ORDER], wz[ORDER];
            weights<ORDER>(rc.y, wy);
            weights<ORDER>(rc.z, wz);
#pragma unroll
            for (int iy = 0; iy < ORDER; ++iy) {
                const int ly = wrap_near(yb + iy, og.sz[1]);
                const float b = a * wy[iy];
                const int rb = (lx * ny + ly) * nz;
#pragma unroll
                for (int iz = 0; iz < ORDER; ++iz) {
                    const int lz = wrap_near(zb + iz, og.sz[2]);
                    if (ly < ny && lz < nz) {
                        const long long v = __double2ll_rn((double)(b * wz[iz]) * kMeshScale);
                        atomicAdd(&own[rb + lz], (unsigned long long)v);
                    }
                }
            }
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < vol; k += NT) {
        const int kz = k % nz, ky = (k / nz) % ny, kx = k / (nz * ny);
        real[((size_t)(X0 + kx) * og.sz[1] + (Y0 + ky)) * og.sz[2] + (Z0 + kz)] =
            (float)((double)(long long)own[k] * kMeshInv);
    }
}

// LDS-tiled spread from the mesh records.  The tile's sites are taken SL at
// a time: a thread per site loads its record and charge and puts its x-plane
// amplitudes qunit q w_x, its y and z weights and its base in the box into
// LDS; then every thread takes stencil rows (site, x plane, y row) and adds
// the row's points into the LDS box.  A site whose stencil leaves the box
// (it drifted further than the margin) adds its points to the global mesh
// directly.  The values are the other spreads' 64-bit integers.
template <int ORDER, int NT>
__device__ __forceinline__ void spread_rows_body(int blk, int nblk, const float *__restrict__ q,
                                                     const BoxDev *__restrict__ bx, float qunit,
                                                     const int *__restrict__ cellStart,
                                                     const int2 *__restrict__ colXY, TileGeo tg, int nlog,
                                                     unsigned long long *__restrict__ mesh,
                                                     const float4 *__restrict__ rec) {
    constexpr int SL = 128;  // sites per pass
    constexpr int H = (ORDER - 1) / 2;
    constexpr int RPS = ORDER * ORDER;  // stencil rows per site
    extern __shared__ unsigned long long tile[];
    __shared__ float sA[SL][ORDER], sY[SL][ORDER], sZ[SL][ORDER];
    __shared__ int sB[SL];
    const int lb = xcd_block(blk, nlog);
    if (lb >= nlog) return;  // uniform over the workgroup
    const MeshGeo m = bx->mesh;
    const int ncol = tg.cb * tg.cb;
    const int col0 = (lb / tg.nzc) * ncol;
    const int zc = lb % tg.nzc;
    const int z0 = zc * tg.zs;
    const int z1 = min(z0 + tg.zs, tg.ncz);
    const int2 cxy = colXY[col0];
    const double margin = bx->tileMargin;
    const int ox = tile_origin(cxy.x * bx->tileCellW[0] - margin, m.h[0]) - H;
    const int oy = tile_origin(cxy.y * bx->tileCellW[1] - margin, m.h[1]) - H;
    const int oz = tile_origin(z0 * bx->tileCellW[2] - margin, m.h[2]) - H;
    const int ex = tg.ext[0], ey = tg.ext[1], ez = tg.ext[2];
    const int vol = ex * ey * ez;
    for (int k = threadIdx.x; k < vol; k += NT) tile[k] = 0ull;
    for (int cu = 0; cu < ncol; ++cu) {
    const int a0 = cellStart[(col0 + cu) * tg.ncz + z0];
    const int a1 = cellStart[(col0 + cu) * tg.ncz + z1];
    for (int base = a0; base < a1; base += SL) {
        const int ns = min(SL, a1 - base);
        if (threadIdx.x < ns) {
            const int i = base + threadIdx.x;
            const float4 rc = rec[i];
            const float qi = q[i];
            int b = -1;
            if (qi != 0.0f) {
                const int gx = rec_point(rc, 0), gy = rec_point(rc, 1), gz = rec_point(rc, 2);
                float wx[ORDER], wy[ORDER], wz[ORDER];
                weights<ORDER>(rc.x, wx);
                weights<ORDER>(rc.y, wy);
                weights<ORDER>(rc.z, wz);
                const float qq = qunit * qi;
                const int rx = wrap_near(gx - H - ox, m.sz[0]), ry = wrap_near(gy - H - oy, m.sz[1]),
                          rz = wrap_near(gz - H - oz, m.sz[2]);
                if (rx + ORDER <= ex && ry + ORDER <= ey && rz + ORDER <= ez) {
                    b = (rx * ey + ry) * ez + rz;
#pragma unroll
                    for (int k = 0; k < ORDER; ++k) {
                        sA[threadIdx.x][k] = qq * wx[k];
                        sY[threadIdx.x][k] = wy[k];
                        sZ[threadIdx.x][k] = wz[k];
                    }
                } else {
#pragma unroll
                    for (int ix = 0; ix < ORDER; ++ix) {
                        const int px = wrap_near(gx + ix - H, m.sz[0]);
                        const float a = qq * wx[ix];
#pragma unroll
                        for (int iy = 0; iy < ORDER; ++iy) {
                            const int py = wrap_near(gy + iy - H, m.sz[1]);
                            const float bb = a * wy[iy];
                            const size_t rowb = ((size_t)px * m.sz[1] + py) * m.sz[2];
#pragma unroll
                            for (int iz = 0; iz < ORDER; ++iz) {
                                const long long v = __double2ll_rn((double)(bb * wz[iz]) * kMeshScale);
                                atomicAdd(&mesh[rowb + wrap_near(gz + iz - H, m.sz[2])], (unsigned long long)v);
                            }
                        }
                    }
                }
            }
            sB[threadIdx.x] = b;
        }
        __syncthreads();
        for (int r = threadIdx.x; r < ns * RPS; r += NT) {
            const int st = r / RPS, rw = r - st * RPS;
            const int ix = rw / ORDER, iy = rw - ix * ORDER;
            const int bb = sB[st];
            if (bb < 0) continue;
            const float b = sA[st][ix] * sY[st][iy];
            const int off = bb + (ix * ey + iy) * ez;
#pragma unroll
            for (int iz = 0; iz < ORDER; ++iz) {
                const long long v = __double2ll_rn((double)(b * sZ[st][iz]) * kMeshScale);
                atomicAdd(&tile[off + iz], (unsigned long long)v);
            }
        }
        __syncthreads();
    }
    }
    // flush by z rows per wave, as the tiled spread
    const int lane = threadIdx.x & 63, nw = NT >> 6;
    const int per = 64 / ez;
    const int sub = fft_div(lane, 1.0f / (float)ez), kz = lane - sub * ez;
    const int gz = wrap_near(oz + kz, m.sz[2]);
    const float rey = 1.0f / (float)ey;
    for (int row = (threadIdx.x >> 6) * per + sub; sub < per && row < ex * ey; row += nw * per) {
        const int kx = fft_div(row, rey), ky = row - kx * ey;
        const unsigned long long v = tile[row * ez + kz];
        if (v == 0ull) continue;
        const int gx = wrap_near(ox + kx, m.sz[0]), gy = wrap_near(oy + ky, m.sz[1]);
        atomicAdd(&mesh[((size_t)gx * m.sz[1] + gy) * m.sz[2] + gz], v);
    }
}

__device__ __forceinline__ void mesh_to_real_body(int blk, int nblk, int ng, long long *__restrict__ mesh, float *__restrict__ out) {
    int g = blk * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    out[g] = (float)((double)mesh[g] * kMeshInv);
    mesh[g] = 0;
}

__device__ __forceinline__ double kcomp(int id, int n, double L) {
    int m = id > n / 2 ? id - n : id;
    return 2.0 * kPi * m / L;
}

// sinc with the reference's small-argument series (FixChargeEwald.cu:302-312)
__device__ __forceinline__ double sinc_d(double x) {
    if (x < 0.1 && x > -0.1) {
        double x2 = x * x;
        return 1.0 - x2 * 0.16666666667 + x2 * x2 * 0.008333333333333333 -
               x2 * x2 * x2 * 0.00019841269841269841;
    }
    return sin(x) / x;
}

// optimal influence function (calc_Green_function, FixChargeEwald.cu:314-434)
__device__ double green_point_any(const double *k, const MeshGeo &m, double f4a2, int order, int lim) {
    double s1 = 0.0, s2 = 0.0;
    for (int mx = -lim; mx <= lim; ++mx)
        for (int my = -lim; my <= lim; ++my)
            for (int mz = -lim; mz <= lim; ++mz) {
                double kp0 = k[0] + 2.0 * kPi * mx / m.h[0];
                double kp1 = k[1] + 2.0 * kPi * my / m.h[1];
                double kp2 = k[2] + 2.0 * kPi * mz / m.h[2];
                double kl = kp0 * kp0 + kp1 * kp1 + kp2 * kp2;
                double W = sinc_d(kp0 * m.h[0] * 0.5) * sinc_d(kp1 * m.h[1] * 0.5) * sinc_d(kp2 * m.h[2] * 0.5);
                double W2 = 1.0;
                for (int o = 0; o < 2 * order; ++o) W2 *= W;
                s1 += 4.0 * kPi * exp(-kl * f4a2) * (k[0] * kp0 + k[1] * kp1 + k[2] * kp2) / kl * W2;
                s2 += W2;
            }
    return s1 / (s2 * s2);
}

// Optimal influence function on the half spectrum, FP64 arithmetic, one
// thread per point with the alias sums written out (any image range).
__device__ __forceinline__ void green_direct_body(int blk, int nblk, const int *__restrict__ gate, const BoxDev *__restrict__ bx, int szh, int order,
                               float *__restrict__ G) {
    if (gate && *gate == 0) return;
    const MeshGeo m = bx->mesh;
    const double alpha = bx->alpha;
    const int lim = bx->greenLim;
    int idx = blk * blockDim.x + threadIdx.x;
    int total = m.sz[0] * m.sz[1] * szh;
    if (idx >= total) return;
    int iz = idx % szh, iy = (idx / szh) % m.sz[1], ix = idx / (szh * m.sz[1]);
    double k[3] = {kcomp(ix, m.sz[0], m.L[0]), kcomp(iy, m.sz[1], m.L[1]), kcomp(iz, m.sz[2], m.L[2])};
    double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    if (k2 == 0.0) {
        G[idx] = 0.0f;
        return;
    }
    G[idx] = (float)(green_point_any(k, m, 0.25 / alpha / alpha, order, lim) / k2);
}

// One point of the tabulated Green function: the 3 (2 LIM + 1) table entries
// of its indices are loaded first, then the (2 LIM + 1)^3 images are summed
// in registers in the order of k_green_direct.  FP32, like the reference's
// Green_function_cu (the result is stored in FP32).
struct GreenWin {
    int s[3], n[3];  // per axis: first mesh index of the window and its length
    int off1, off2;  // table offsets of the y and z windows
};

template <int LIM>
__device__ __forceinline__ float green_tab_point(const float *tkp, const float *tex, const float *tw,
                                                 const GreenWin &w, int ix, int iy, int iz) {
    constexpr int NI = 2 * LIM + 1;
    float kx[NI], ey[NI], wy[NI], ky[NI], ex[NI], wx[NI], kz[NI], ez[NI], wz[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        const int a = j * w.n[0] + ix - w.s[0], b = w.off1 + j * w.n[1] + iy - w.s[1],
                  c = w.off2 + j * w.n[2] + iz - w.s[2];
        kx[j] = tkp[a];
        ex[j] = tex[a];
        wx[j] = tw[a];
        ky[j] = tkp[b];
        ey[j] = tex[b];
        wy[j] = tw[b];
        kz[j] = tkp[c];
        ez[j] = tex[c];
        wz[j] = tw[c];
    }
    // image 0 of each axis is k itself
    const float k0 = kx[LIM], k1 = ky[LIM], k2v = kz[LIM];
    const float k2 = k0 * k0 + k1 * k1 + k2v * k2v;
    if (k2 == 0.0f) return 0.0f;
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int a = 0; a < NI; ++a)
#pragma unroll
        for (int b = 0; b < NI; ++b) {
            const float wab = wx[a] * wy[b], eab = ex[a] * ey[b];
            const float kab = kx[a] * kx[a] + ky[b] * ky[b], dab = k0 * kx[a] + k1 * ky[b];
#pragma unroll
            for (int cc = 0; cc < NI; ++cc) {
                const float W2 = wab * wz[cc];
                const float kl = kab + kz[cc] * kz[cc];
                s1 += eab * ez[cc] * (dab + k2v * kz[cc]) / kl * W2;
                s2 += W2;
            }
        }
    return 4.0f * (float)kPi * s1 / (s2 * s2) / k2;
}

// sinc with the reference's small-argument series (FixChargeEwald.cu:302-312), FP32
__device__ __forceinline__ float sinc_f(float x) {
    if (x < 0.1f && x > -0.1f) {
        const float x2 = x * x;
        return 1.0f - x2 * 0.16666666667f + x2 * x2 * 0.008333333333333333f - x2 * x2 * x2 * 0.00019841269841269841f;
    }
    return sinf(x) / x;
}

// Same function with the per-axis factors tabulated: the Gaussian
// exp(-|k'|^2 / 4 alpha^2) and the assignment factor W(k')^2p factor over the
// axes, so for every axis d, mesh index i and image j the workgroup first
// computes k'_d = k_d + 2 pi j / h_d, exp(-k'_d^2 / 4 alpha^2) and
// sinc(k'_d h_d / 2)^2p once into LDS for the index windows its points use
// (one or two x planes, the y rows they touch, the z half-spectrum columns:
// at most 3 (2 lim + 1) (sz_x + sz_y + sz_z) entries); each point then sums
// its (2 lim + 1)^3 images from products of table entries.  lim >
// kGreenTabLim takes the direct FP64 sums.
constexpr int kGreenTabLim = 2;
// Workgroups from gblocks on dilate the system instead (dilate_thread): the
// barostat's two box-change passes in one launch; both only read the box.
__device__ __forceinline__ void green_body(int blk, int nblk, const int *__restrict__ gate, const BoxDev *__restrict__ bx,
                                               int szh, int order, float *__restrict__ G, int gblocks,
                                               DilateArgs dl) {
    extern __shared__ float gtab[];
    if ((int)blk >= nblk - dl.vblocks) {  // the thermostat's rescale, under its own gate
        if (!dl.tgate || *dl.tgate != 0)
            vscale_thread(((int)blk - (nblk - dl.vblocks)) * blockDim.x + threadIdx.x, dl.n, dl.vel, dl.thermo, dl.tconv,
                          dl.T0, dl.tmode, dl.coupling);
        return;
    }
    if (gate && *gate == 0) return;
    if ((int)blk >= gblocks) {  // uniform over the workgroup
        dilate_thread(((int)blk - gblocks) * blockDim.x + threadIdx.x, dl, bx);
        return;
    }
    const MeshGeo m = bx->mesh;
    const double alpha = bx->alpha;
    const int lim = bx->greenLim;
    const int idx = blk * blockDim.x + threadIdx.x;
    const int total = m.sz[0] * m.sz[1] * szh;
    const double f4a2 = 0.25 / alpha / alpha;
    if (lim > kGreenTabLim) {  // uniform over the grid
        if (idx >= total) return;
        const int iz = idx % szh, iy = (idx / szh) % m.sz[1], ix = idx / (szh * m.sz[1]);
        const double k[3] = {kcomp(ix, m.sz[0], m.L[0]), kcomp(iy, m.sz[1], m.L[1]), kcomp(iz, m.sz[2], m.L[2])};
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        G[idx] = k2 == 0.0 ? 0.0f : (float)(green_point_any(k, m, f4a2, order, lim) / k2);
        return;
    }
    const int NI = 2 * lim + 1;
    // index windows of this workgroup's points (iz fastest, then iy, then ix)
    GreenWin w;
    {
        const int first = blk * blockDim.x, last = min(first + (int)blockDim.x, total) - 1;
        const int fz = first % szh, fy = (first / szh) % m.sz[1], fx = first / (szh * m.sz[1]);
        const int lz = last % szh, ly = (last / szh) % m.sz[1], lx = last / (szh * m.sz[1]);
        w.s[0] = fx;
        w.n[0] = lx - fx + 1;
        const bool onePlane = lx == fx;
        w.s[1] = onePlane ? fy : 0;
        w.n[1] = onePlane ? ly - fy + 1 : m.sz[1];
        const bool oneRow = onePlane && ly == fy;
        w.s[2] = oneRow ? fz : 0;
        w.n[2] = oneRow ? lz - fz + 1 : szh;
        w.off1 = NI * w.n[0];
        w.off2 = NI * (w.n[0] + w.n[1]);
    }
    const int ne = NI * (m.sz[0] + m.sz[1] + m.sz[2]);  // the table's allocation (its upper bound)
    const int nu = NI * (w.n[0] + w.n[1] + w.n[2]);      // entries this workgroup fills
    float *tkp = gtab, *tex = gtab + ne, *tw = gtab + 2 * ne;
    const float f4a2f = (float)f4a2;
    for (int e = threadIdx.x; e < nu; e += blockDim.x) {
        // per-axis values by selection, not by a dynamic index (which would
        // put the local structs in scratch memory)
        const bool z = e >= w.off2, y = !z && e >= w.off1;
        const int n = z ? m.sz[2] : (y ? m.sz[1] : m.sz[0]);
        const int nw = z ? w.n[2] : (y ? w.n[1] : w.n[0]);
        const int ws = z ? w.s[2] : (y ? w.s[1] : w.s[0]);
        const double Ld = z ? m.L[2] : (y ? m.L[1] : m.L[0]);
        const double hd = z ? m.h[2] : (y ? m.h[1] : m.h[0]);
        const int r = e - (z ? w.off2 : (y ? w.off1 : 0)), j = r / nw, i = ws + (r - j * nw);
        const float q = (float)(kcomp(i, n, Ld) + 2.0 * kPi * (j - lim) / hd);
        const float W = sinc_f(q * (float)hd * 0.5f);
        float W2 = 1.0f;
        for (int o = 0; o < 2 * order; ++o) W2 *= W;
        tkp[e] = q;
        tex[e] = expf(-q * q * f4a2f);
        tw[e] = W2;
    }
    __syncthreads();
    if (idx >= total) return;
    const int iz = idx % szh, iy = (idx / szh) % m.sz[1], ix = idx / (szh * m.sz[1]);
    float g;
    switch (lim) {
        case 0: g = green_tab_point<0>(tkp, tex, tw, w, ix, iy, iz); break;
        case 1: g = green_tab_point<1>(tkp, tex, tw, w, ix, iy, iz); break;
        default: g = green_tab_point<2>(tkp, tex, tw, w, ix, iy, iz); break;
    }
    G[idx] = g;
}

// Energy 0.5 sum |rho_hat|^2 G / V over the full spectrum and the k-space
// virial (virials_cu): half-spectrum entries with 0 < iz < sz/2 count twice.
__device__ __forceinline__ void kspace_ev_point(const MeshGeo &m, double alpha, double invV, int ix, int iy, int iz,
                                                float g, float2 r, double vals[7]) {
    double w = (iz == 0 || 2 * iz == m.sz[2]) ? 1.0 : 2.0;
    double k[3] = {kcomp(ix, m.sz[0], m.L[0]), kcomp(iy, m.sz[1], m.L[1]), kcomp(iz, m.sz[2], m.L[2])};
    double kl = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    double E = w * ((double)r.x * r.x + (double)r.y * r.y) * g;
    vals[0] = 0.5 * E * invV;
    if (kl != 0.0) {
        double diff = -2.0 * (1.0 / kl + 0.25 / (alpha * alpha));
        vals[1] = 0.5 * invV * (1.0 + diff * k[0] * k[0]) * E;
        vals[2] = 0.5 * invV * (1.0 + diff * k[1] * k[1]) * E;
        vals[3] = 0.5 * invV * (1.0 + diff * k[2] * k[2]) * E;
        vals[4] = 0.5 * invV * diff * k[0] * k[1] * E;
        vals[5] = 0.5 * invV * diff * k[0] * k[2] * E;
        vals[6] = 0.5 * invV * diff * k[1] * k[2] * E;
    }
}

// workgroup sums of the seven values, one fixed-point add per value
__device__ __forceinline__ void kspace_ev_reduce(const double vals[7], unsigned long long *__restrict__ acc) {
    __shared__ double wred[16][7];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        double x = team_sum_d<64>(vals[s]);
        if (lane == 0) wred[wv][s] = x;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double x = 0.0;
        for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) x += wred[ww][threadIdx.x];
        fx_add(&acc_rep(acc)[threadIdx.x == 0 ? ACC_KSPACE : ACC_VIR + threadIdx.x - 1], x);
    }
}

// E_hat_d = i k_d G rho_hat (E_field_cu), k_d := 0 on the d-th Nyquist plane.
// With ev set (and the gate, if any, open) the same pass also sums the
// k-space energy and virial (kspace_ev_point, one read of G and rho_hat for
// both); ev and the gate are uniform over the grid.
__device__ __forceinline__ void efield_body(int blk, int nblk, const BoxDev *__restrict__ bx, int szh, const float *__restrict__ G,
                         const float2 *__restrict__ rho, float2 *__restrict__ E, int ev, const int *__restrict__ gate,
                         unsigned long long *__restrict__ acc) {
    const MeshGeo m = bx->mesh;
    int idx = blk * blockDim.x + threadIdx.x;
    int total = m.sz[0] * m.sz[1] * szh;
    const bool sums = ev && (!gate || *gate != 0);
    double vals[7] = {0, 0, 0, 0, 0, 0, 0};
    if (idx < total) {
        int iz = idx % szh, iy = (idx / szh) % m.sz[1], ix = idx / (szh * m.sz[1]);
        float kx = (2 * ix == m.sz[0]) ? 0.0f : (float)kcomp(ix, m.sz[0], m.L[0]);
        float ky = (2 * iy == m.sz[1]) ? 0.0f : (float)kcomp(iy, m.sz[1], m.L[1]);
        float kz = (2 * iz == m.sz[2]) ? 0.0f : (float)kcomp(iz, m.sz[2], m.L[2]);
        float g = G[idx];
        float2 r = rho[idx];
        float gr = g * r.x, gi = g * r.y;
        E[idx] = make_float2(-kx * gi, kx * gr);
        E[total + idx] = make_float2(-ky * gi, ky * gr);
        E[2 * total + idx] = make_float2(-kz * gi, kz * gr);
        if (sums) kspace_ev_point(m, bx->alpha, 1.0 / bx->vol, ix, iy, iz, g, r, vals);
    }
    if (sums) kspace_ev_reduce(vals, acc);
}

// Field interpolation (Ewald_long_range_forces_order_*_cu): a team of T
// lanes per charged atom splits the ORDER^2 stencil rows (the field's three
// components are consecutive real meshes, as the batched inverse FFT writes
// them), and the team sums by butterfly shuffles before one fixed-point add
// per component.
// the three real field components interleaved per mesh point, so the gather
// takes one 16-byte load per stencil point
__device__ __forceinline__ void field_pack_body(int blk, int nblk, int ng, const float *__restrict__ E, float4 *__restrict__ E4) {
    const int g = blk * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    E4[g] = make_float4(E[g], E[ng + g], E[2 * (size_t)ng + g], 0.0f);
}

// ---------------------------------------------------------------------------
// The mesh solve in three passes, for power-of-two meshes whose planes fit
// one workgroup's LDS: the transforms of E_field_cu's pipeline
// (FixChargeEwald.cu: forward R2C of the charge mesh, E_hat_d =
// i k_d G rho_hat, inverse C2R of each component), here as
//   pass 1, one workgroup per x plane: the plane's charges (fixed point, read
//     and cleared for the next spread; or the owner spread's real mesh), real
//     FFTs along z two rows at a time (one complex FFT of row a + i row b,
//     split by Hermitian symmetry), complex FFTs along y -> S1[x][ky][kz];
//   pass 2, one wave per (ky, group of kz columns, component d): complex FFT
//     along x gives rho_hat(kx, ky, kz), and i k_d G rho_hat goes back along
//     x -> S2[d][x][ky][kz]; with energy or virial requested a fourth wave per
//     column group sums them (kspace_ev_point) from its own rho_hat;
//   pass 3, one workgroup per (x plane, component): inverse complex FFTs along
//     y, then the Hermitian inverse along z two rows at a time (row a + i row
//     b, the imaginary parts of the kz = 0 and nz/2 terms dropped as a real
//     transform does) -> the planar real field E[d][x][y][z].
// Unnormalised in both directions, forward exp(-i), inverse exp(+i), as the
// library transforms of the other path; S1 and S2 live in that path's
// half-spectrum arrays (rho_hat, E_hat).  Every transform belongs to one
// wave (its stages need no workgroup barrier): a Stockham autosort FFT
// (natural order in and out) of radix-8 stages and one radix-4 or radix-2
// stage, each lane doing whole radix-R butterflies in registers between two
// LDS buffers (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC 2008).
constexpr int kFftMax = 128;      // longest axis; twiddles exp(-2 pi i k / 128), k < 128
constexpr int kFftThreads = 256;  // pass 2: 4 waves, one (ky, columns, component) task each
constexpr int kFftPlaneThreads = 512;  // passes 1 and 3: 8 waves, one group of lines each
constexpr int kFftWavePts = 512;  // points a wave transforms at a time (lines x length)

__device__ __forceinline__ int fft_log2(int n) { return __ffs(n) - 1; }


// Workgroup barrier ordering LDS only: the passes read global memory that
// earlier launches wrote and write what later launches read, so their
// barriers need not wait for outstanding global loads and stores.
__device__ __forceinline__ void fft_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Between the stages of one wave's transforms: the wave's LDS accesses are
// processed in order, so this only keeps the compiler from moving them.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ void fft_twiddles(float2 *tw) {
    for (int k = threadIdx.x; k < kFftMax; k += blockDim.x) {
        float sn, cs;
        sincospif((float)k / (kFftMax / 2), &sn, &cs);
        tw[k] = make_float2(cs, -sn);
    }
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// multiplication by -i (forward) or +i (inverse)
__device__ __forceinline__ float2 rot_i(float2 a, bool inv) {
    return inv ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R>
__device__ __forceinline__ void small_dft(float2 (&v)[R], bool inv);

template <>
__device__ __forceinline__ void small_dft<2>(float2 (&v)[2], bool) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
}

template <>
__device__ __forceinline__ void small_dft<4>(float2 (&v)[4], bool inv) {
    const float2 t0 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 t1 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 t2 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 t3 = rot_i(make_float2(v[1].x - v[3].x, v[1].y - v[3].y), inv);
    v[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
    v[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
    v[1] = make_float2(t1.x + t3.x, t1.y + t3.y);
    v[3] = make_float2(t1.x - t3.x, t1.y - t3.y);
}

template <>
__device__ __forceinline__ void small_dft<8>(float2 (&v)[8], bool inv) {
    float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    small_dft<4>(e, inv);
    small_dft<4>(o, inv);
    const float h = 0.70710678118654752f;
    // o_k times exp(-+ 2 pi i k / 8)
    o[1] = inv ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
               : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    o[2] = rot_i(o[2], inv);
    o[3] = inv ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
               : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = make_float2(e[k].x + o[k].x, e[k].y + o[k].y);
        v[k + 4] = make_float2(e[k].x - o[k].x, e[k].y - o[k].y);
    }
}

// One Stockham stage of radix R (sub-transforms of length ns done so far) of
// the calling wave's nl transforms of n points, element e of line l at
// l * ls + e * es.  Neighbouring lanes take neighbouring lines.
template <int R>
__device__ __forceinline__ void wave_stage(const float2 *src, float2 *dst, int n, int ns, int nl, int ls, int es,
                                           const float2 *tw, bool inv) {
    const int nb = n / R;
    const int step = kFftMax / (ns * R);
    const float rnl = 1.0f / (float)nl;
    for (int t = threadIdx.x & 63; t < nl * nb; t += 64) {
        const int j = fft_div(t, rnl), l = t - j * nl;
        const int jm = j & (ns - 1);
        const float2 *sp = src + l * ls;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = sp[(j + r * nb) * es];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            float2 w = tw[(r * jm * step) & (kFftMax - 1)];
            if (inv) w.y = -w.y;
            v[r] = cmul(v[r], w);
        }
        small_dft<R>(v, inv);
        float2 *dp = dst + l * ls;
        const int base = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) dp[(base + r * ns) * es] = v[r];
    }
}

// The calling wave's nl transforms of n points (a power of two, 2..128) from
// a, with b the other buffer (the same layout); returns the one holding the
// result.  The wave's earlier LDS writes to a are visible to it.
__device__ __forceinline__ float2 *wave_fft(float2 *a, float2 *b, int n, int nl, int ls, int es, const float2 *tw,
                                            bool inv) {
    int ns = 1;
    wave_sync();
    while (ns < n) {
        const int rem = n / ns;
        if (rem >= 8) {
            wave_stage<8>(a, b, n, ns, nl, ls, es, tw, inv);
            ns *= 8;
        } else if (rem == 4) {
            wave_stage<4>(a, b, n, ns, nl, ls, es, tw, inv);
            ns *= 4;
        } else {
            wave_stage<2>(a, b, n, ns, nl, ls, es, tw, inv);
            ns *= 2;
        }
        float2 *x = a;
        a = b;
        b = x;
        wave_sync();
    }
    return a;
}

// true when an n-point wave_fft ends in its first buffer (even stage count)
__device__ __forceinline__ bool fft_ends_in_a(int n) {
    int stages = 0;
    for (int ns = 1; ns < n;) {
        const int rem = n / ns;
        ns *= rem >= 8 ? 8 : rem;
        ++stages;
    }
    return (stages & 1) == 0;
}

__device__ __forceinline__ void fft_pass1_body(int blk, int nblk, const BoxDev *__restrict__ bx, long long *__restrict__ mesh,
                                               const float *__restrict__ real, float2 *__restrict__ S1) {
    extern __shared__ float2 fsh[];
    const MeshGeo m = bx->mesh;
    const int ny = m.sz[1], nz = m.sz[2], szh = nz / 2 + 1;
    const int lz = fft_log2(nz);
    const int rs = nz + 1;  // padded row pitch of the z transforms
    const int cap = max((ny / 2) * rs, ny * szh);
    float2 *tw = fsh;
    float2 *A = tw + kFftMax, *B = A + cap, *Pl = B + cap;
    fft_twiddles(tw);
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // the plane as ny / 2 row pairs (pitch rs): row 2r in .x, row 2r + 1 in .y
    const size_t base = (size_t)blk * ny * nz;
    const int total = ny * nz;
    constexpr int U = 8;
    for (int t0 = threadIdx.x; t0 < total; t0 += U * kFftPlaneThreads) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * kFftPlaneThreads;
            if (t < total) v[u] = mesh ? (float)((double)mesh[base + t] * kMeshInv) : real[base + t];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * kFftPlaneThreads;
            if (t < total) {
                if (mesh) mesh[base + t] = 0;
                const int y = t >> lz, z = t & (nz - 1);
                ((float *)(A + (y >> 1) * rs + z))[y & 1] = v[u];
            }
        }
    }
    fft_sync();
    // z: groups of row pairs per wave; each wave splits its own rows into Pl
    const int lzw = kFftWavePts / nz;
    const float2 *Z = A;
    for (int r0 = wave * lzw; r0 < ny / 2; r0 += nw * lzw) {
        const int nl = min(lzw, ny / 2 - r0);
        Z = wave_fft(A + r0 * rs, B + r0 * rs, nz, nl, rs, 1, tw, false) - r0 * rs;
        const float rszh = 1.0f / (float)szh;
        for (int t = threadIdx.x & 63; t < nl * szh; t += 64) {
            const int rr = fft_div(t, rszh), k = t - rr * szh, r = r0 + rr;
            const float2 P = Z[r * rs + k], Q = Z[r * rs + ((nz - k) & (nz - 1))];
            Pl[2 * r * szh + k] = make_float2(0.5f * (P.x + Q.x), 0.5f * (P.y - Q.y));
            Pl[(2 * r + 1) * szh + k] = make_float2(0.5f * (P.y + Q.y), -0.5f * (P.x - Q.x));
        }
    }
    fft_sync();
    // y: groups of kz columns per wave, written out column group by group
    const int lyw = kFftWavePts / ny;
    float2 *out = S1 + (size_t)blk * ny * szh;
    for (int c0 = wave * lyw; c0 < szh; c0 += nw * lyw) {
        const int nl = min(lyw, szh - c0);
        const float2 *Y = wave_fft(Pl + c0, A + c0, ny, nl, 1, szh, tw, false) - c0;
        const float rnl = 1.0f / (float)nl;
        for (int t = threadIdx.x & 63; t < ny * nl; t += 64) {
            const int y = fft_div(t, rnl), k = c0 + t - y * nl;
            out[y * szh + k] = Y[y * szh + k];
        }
    }
}

__device__ __forceinline__ void fft_pass2_body(int blk, int nblk, const BoxDev *__restrict__ bx, const float *__restrict__ G,
                                               const float2 *__restrict__ S1, float2 *__restrict__ S2, int ev,
                                               const int *__restrict__ gate, unsigned long long *__restrict__ acc) {
    extern __shared__ float2 fsh[];
    float2 *tw = fsh;
    fft_twiddles(tw);
    fft_sync();
    const MeshGeo m = bx->mesh;
    const int nx = m.sz[0], ny = m.sz[1], szh = m.sz[2] / 2 + 1;
    const int nh = nx * ny * szh;
    const int lw = kFftWavePts / nx;  // kz columns of a wave (a power of two)
    const int llw = fft_log2(lw);
    const int ngr = (szh + lw - 1) / lw;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int task = blk * (blockDim.x >> 6) + wave;
    // task kinds: components 0..2, and with ev a fourth wave per column group
    // that only sums the k-space energy and virial (off the field waves' path)
    const int kinds = ev ? 4 : 3;
    if (task >= ny * ngr * kinds) return;
    const int d = task % kinds, ky = (task / kinds) / ngr, kz0 = ((task / kinds) % ngr) * lw;
    const bool sums = d == 3 && (!gate || *gate != 0);
    if (d == 3 && !sums) return;
    const int nl = min(lw, szh - kz0);
    float2 *A = tw + kFftMax + wave * 2 * kFftWavePts, *Bf = A + kFftWavePts;
    // element (x, column l) at A[l + x * lw]; columns beyond szh are zero
    constexpr int U = kFftWavePts / 64;
    float2 v[U];
    float g[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int t = lane + 64 * u, l = t & (lw - 1), x = t >> llw;
        v[u] = make_float2(0.0f, 0.0f);
        g[u] = 0.0f;
        if (l < nl) {
            const int i = (x * ny + ky) * szh + kz0 + l;
            v[u] = S1[i];
            g[u] = G[i];
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) A[lane + 64 * u] = v[u];
    float2 *R = wave_fft(A, Bf, nx, lw, 1, lw, tw, false);
    float2 *F = R == A ? Bf : A;
    if (sums) {
        // kspace_ev_point's sums with the per-axis factors taken out of the
        // loop (k_d = m_d (2 pi / L_d), 1 / (4 alpha^2)): one division a point
        double vals[7] = {0, 0, 0, 0, 0, 0, 0};
        const double invV = 1.0 / bx->vol, hV = 0.5 * invV;
        const double c4 = 0.25 / (bx->alpha * bx->alpha);
        const double ku0 = 2.0 * kPi / m.L[0], ku2 = 2.0 * kPi / m.L[2];
        const double kyd = (double)(2 * ky > ny ? ky - ny : ky) * (2.0 * kPi / m.L[1]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = lane + 64 * u, l = t & (lw - 1), kx = t >> llw, kz = kz0 + l;
            if (l < nl) {
                const double k0 = (double)(2 * kx > nx ? kx - nx : kx) * ku0;
                const double k2 = (double)(2 * kz > m.sz[2] ? kz - m.sz[2] : kz) * ku2;
                const double kl = k0 * k0 + kyd * kyd + k2 * k2;
                const float2 r = R[t];
                const double wgt = (kz == 0 || 2 * kz == m.sz[2]) ? 1.0 : 2.0;
                const double E = wgt * ((double)r.x * r.x + (double)r.y * r.y) * g[u];
                vals[0] += hV * E;
                if (kl != 0.0) {
                    const double diff = -2.0 * (1.0 / kl + c4);
                    const double hE = hV * E, dE = diff * hE;
                    vals[1] += hE + dE * k0 * k0;
                    vals[2] += hE + dE * kyd * kyd;
                    vals[3] += hE + dE * k2 * k2;
                    vals[4] += dE * k0 * kyd;
                    vals[5] += dE * k0 * k2;
                    vals[6] += dE * kyd * k2;
                }
            }
        }
#pragma unroll
        for (int s2 = 0; s2 < 7; ++s2) {
            const double x = team_sum_d<64>(vals[s2]);
            if (lane == 0) fx_add(&acc_rep(acc)[s2 == 0 ? ACC_KSPACE : ACC_VIR + s2 - 1], x);
        }
        return;
    }
    // k_d of this wave's component: 2 pi m / L_d in FP32 (0 on the Nyquist plane)
    const int nd = m.sz[d];
    const float kunit = (float)(2.0 * kPi / m.L[d]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int t = lane + 64 * u, l = t & (lw - 1), kx = t >> llw, kz = kz0 + l;
        const float2 r = R[t];
        const int id = d == 0 ? kx : (d == 1 ? ky : kz);
        const int mi = 2 * id > nd ? id - nd : id;
        const float kd = 2 * id == nd ? 0.0f : kunit * (float)mi;
        const float gr = g[u] * r.x, gi = g[u] * r.y;
        F[t] = make_float2(-kd * gi, kd * gr);
    }
    const float2 *X = wave_fft(F, R, nx, lw, 1, lw, tw, true);
    float2 *out = S2 + d * nh;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int t = lane + 64 * u, l = t & (lw - 1), x = t >> llw;
        if (l < nl) out[(x * ny + ky) * szh + kz0 + l] = X[t];
    }
}

__device__ __forceinline__ void fft_pass3_body(int blk, int nblk, const BoxDev *__restrict__ bx,
                                               const float2 *__restrict__ S2, float *__restrict__ E) {
    extern __shared__ float2 fsh[];
    const MeshGeo m = bx->mesh;
    const int nx = m.sz[0], ny = m.sz[1], nz = m.sz[2], szh = nz / 2 + 1;
    const size_t nh = (size_t)nx * ny * szh, ng = (size_t)nx * ny * nz;
    const int d = blk / nx, x = blk - d * nx;
    const int lz = fft_log2(nz);
    const int rs = nz + 1;  // padded row pitch of the z transforms
    const int cap = max((ny / 2) * rs, ny * szh);
    float2 *tw = fsh;
    float2 *A = tw + kFftMax, *B = A + cap, *Q = B + cap;
    fft_twiddles(tw);
    fft_sync();
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = threadIdx.x & 63;
    // y: each wave loads and transforms its groups of kz columns
    const float2 *in = S2 + d * nh + (size_t)x * ny * szh;
    const int lyw = kFftWavePts / ny;
    const float2 *P = fft_ends_in_a(ny) ? A : B;  // the same for every wave's columns
    for (int c0 = wave * lyw; c0 < szh; c0 += nw * lyw) {
        const int nl = min(lyw, szh - c0);
        constexpr int U = 8;
        const float rnl = 1.0f / (float)nl;
        for (int t0 = lane; t0 < ny * nl; t0 += 64 * U) {
            float2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + 64 * u;
                if (t < ny * nl) {
                    const int y = fft_div(t, rnl), k = c0 + t - y * nl;
                    v[u] = in[y * szh + k];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + 64 * u;
                if (t < ny * nl) {
                    const int y = fft_div(t, rnl), k = c0 + t - y * nl;
                    A[y * szh + k] = v[u];
                }
            }
        }
        wave_fft(A + c0, B + c0, ny, nl, 1, szh, tw, true);
    }
    fft_sync();
    // z: each wave builds and transforms its groups of row pairs
    float2 *Pw = P == A ? B : A;  // free now (every wave's y transforms are done)
    const int lzw = kFftWavePts / nz;
    float *out = E + d * ng + (size_t)x * ny * nz;
    for (int r0 = wave * lzw; r0 < ny / 2; r0 += nw * lzw) {
        const int nl = min(lzw, ny / 2 - r0);
        for (int t = lane; t < nl * nz; t += 64) {
            const int rr = t >> lz, k = t & (nz - 1), r = r0 + rr;
            float2 a, b;
            if (2 * k <= nz) {
                a = P[2 * r * szh + k];
                b = P[(2 * r + 1) * szh + k];
                if (k == 0 || 2 * k == nz) {
                    a.y = 0.0f;
                    b.y = 0.0f;
                }
            } else {
                a = P[2 * r * szh + nz - k];
                b = P[(2 * r + 1) * szh + nz - k];
                a.y = -a.y;
                b.y = -b.y;
            }
            Q[r * rs + k] = make_float2(a.x - b.y, a.y + b.x);
        }
        const float2 *Z = wave_fft(Q + r0 * rs, Pw + r0 * rs, nz, nl, rs, 1, tw, true) - r0 * rs;
        for (int t = lane; t < 2 * nl * nz; t += 64) {
            const int y = 2 * r0 + t / nz, z = t & (nz - 1);
            out[y * nz + z] = ((const float *)(Z + (y >> 1) * rs + z))[y & 1];
        }
    }
}

template <int ORDER, int T>
__device__ __forceinline__ void gather_body(int blk, int nblk, int n, const double4 *__restrict__ pos,
                                                const float *__restrict__ q, const BoxDev *__restrict__ bx,
                                                const float4 *__restrict__ E4,
                                                unsigned long long *__restrict__ force,
                                                const int *__restrict__ ids, long long *__restrict__ stored,
                                                const float4 *__restrict__ rec) {
    const int nlog = (int)(((long long)n * T + blockDim.x - 1) / blockDim.x);
    const int lb = xcd_block(blk, nlog);
    if (lb >= nlog) return;  // uniform over the workgroup
    const MeshGeo m = bx->mesh;
    const float scale = bx->gatherScale;
    const int gt = lb * blockDim.x + threadIdx.x;
    const int i = gt / T;
    const int tl = threadIdx.x & (T - 1);
    if (i >= n) return;  // whole teams leave together
    const float qi = q[i];
    int g[3];
    float d[3];
    if (rec) {  // the position refresh's mesh point and offsets, loaded with the charge
        const float4 rc = rec[i];
        if (qi == 0.0f) return;
        g[0] = rec_point(rc, 0);
        g[1] = rec_point(rc, 1);
        g[2] = rec_point(rc, 2);
        d[0] = rc.x;
        d[1] = rc.y;
        d[2] = rc.z;
    } else {
        if (qi == 0.0f) return;
        const double4 p = pos[i];
        mesh_point(p.x, 0, m, g[0], d[0]);
        mesh_point(p.y, 1, m, g[1], d[1]);
        mesh_point(p.z, 2, m, g[2], d[2]);
    }
    float wx[ORDER], wy[ORDER], wz[ORDER];
    weights<ORDER>(d[0], wx);
    weights<ORDER>(d[1], wy);
    weights<ORDER>(d[2], wz);
    constexpr int H = (ORDER - 1) / 2;
    float ex = 0.f, ey = 0.f, ez = 0.f;
    // a lane's stencil rows r = tl, tl + T, ...: every row's field points are
    // loaded before any is used (the rows one after another left each lane
    // waiting on one row's loads at a time); the sums run in row order
    constexpr int NRW = (ORDER * ORDER + T - 1) / T;
    int zi[ORDER];
#pragma unroll
    for (int iz = 0; iz < ORDER; ++iz) zi[iz] = wrapi(g[2] + iz - H, m.sz[2]);
    float4 e[NRW][ORDER];
#pragma unroll
    for (int rr = 0; rr < NRW; ++rr) {
        const int r = min(tl + rr * T, ORDER * ORDER - 1);
        const int ix = r / ORDER, iy = r - ix * ORDER;
        const int px = wrapi(g[0] + ix - H, m.sz[0]);
        const int py = wrapi(g[1] + iy - H, m.sz[1]);
        const size_t rowb = ((size_t)px * m.sz[1] + py) * m.sz[2];
#pragma unroll
        for (int iz = 0; iz < ORDER; ++iz) e[rr][iz] = E4[rowb + zi[iz]];
    }
#pragma unroll
    for (int rr = 0; rr < NRW; ++rr) {
        const int r = tl + rr * T;
        if (r >= ORDER * ORDER) break;
        const int ix = r / ORDER, iy = r - ix * ORDER;
        const float wxy = pick<ORDER>(wx, ix) * pick<ORDER>(wy, iy);
#pragma unroll
        for (int iz = 0; iz < ORDER; ++iz) {
            const float w = wxy * wz[iz];
            ex += w * e[rr][iz].x;
            ey += w * e[rr][iz].y;
            ez += w * e[rr][iz].z;
        }
    }
    ex = team_sum<T>(ex);
    ey = team_sum<T>(ey);
    ez = team_sum<T>(ez);
    if (tl == 0) {
        // F = Qunit q sum W (-E / V)  (Ewald_long_range_forces_order_5_cu)
        const float s = scale * qi;
        const long long f0 = to_fx(s * ex), f1 = to_fx(s * ey), f2 = to_fx(s * ez);
        atomicAdd(&force[3 * i + 0], (unsigned long long)f0);
        atomicAdd(&force[3 * i + 1], (unsigned long long)f1);
        atomicAdd(&force[3 * i + 2], (unsigned long long)f2);
        if (stored) {  // the same fixed-point increments, by atom id, for the steps between
            const int id = ids[i];
            stored[3 * id + 0] = f0;
            stored[3 * id + 1] = f1;
            stored[3 * id + 2] = f2;
        }
    }
}

// LDS-tiled field interpolation.  A workgroup takes the charged sites of one
// tile of the LDS spread (a cell column's chunk of zs cells): it loads the
// box of mesh points their stencils can reach - the three components from
// the planar output of the inverse transform (or the field interleaved per
// mesh point, one 16-byte load a point), every lane of the workgroup busy,
// into three LDS planes - once, then a team of T lanes per site sums the
// stencil rows from LDS
// (rows r = lane, lane + T, ..., as the row-team gather).  A site whose
// stencil leaves the box (it drifted further than the margin) reads the rows
// from global memory instead.  The site's charge and record are loaded before
// the box, so both round trips overlap.
template <int ORDER, int T, int NT>
__device__ __forceinline__ void gather_tile_body(int blk, int nblk, const float4 *__restrict__ rec,
                                                     const float *__restrict__ q, const BoxDev *__restrict__ bx,
                                                     const int *__restrict__ cellStart, const int2 *__restrict__ colXY,
                                                     TileGeo tg, int nlog, const float *__restrict__ E,
                                                     const float4 *__restrict__ E4,
                                                     unsigned long long *__restrict__ force,
                                                     const int *__restrict__ ids, long long *__restrict__ stored) {
    extern __shared__ float fb[];  // 3 planes of the box
    __shared__ float lcoef[ORDER * ORDER];
    const int lb = xcd_block(blk, nlog);
    if (lb >= nlog) return;  // uniform over the workgroup
    constexpr int H = (ORDER - 1) / 2;
    constexpr int SPR = NT / T;  // sites per round
    const MeshGeo m = bx->mesh;
    const int sx = m.sz[0], sy = m.sz[1], sz = m.sz[2];
    const int ncol = tg.cb * tg.cb;
    const int col0 = (lb / tg.nzc) * ncol, zc = lb % tg.nzc;
    const int z0 = zc * tg.zs, z1 = min(z0 + tg.zs, tg.ncz);
    int a0 = cellStart[col0 * tg.ncz + z0], a1 = cellStart[col0 * tg.ncz + z1];
    const int tl = threadIdx.x & (T - 1);
    // first round's site: charge and record in flight with the box loads
    int s = threadIdx.x / T;
    float qi = 0.0f;
    float4 rc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a0 + s < a1) {
        qi = q[a0 + s];
        rc = rec[a0 + s];
    }
    const int2 cxy = colXY[col0];
    const double margin = bx->tileMargin;
    const int ox = tile_origin(cxy.x * bx->tileCellW[0] - margin, m.h[0]) - H;
    const int oy = tile_origin(cxy.y * bx->tileCellW[1] - margin, m.h[1]) - H;
    const int oz = tile_origin(z0 * bx->tileCellW[2] - margin, m.h[2]) - H;
    const int ex = tg.ext[0], ey = tg.ext[1], ez = tg.ext[2];
    const int vol = ex * ey * ez;
    const size_t ng = (size_t)sx * sy * sz;
    if (threadIdx.x < ORDER * ORDER) lcoef[threadIdx.x] = WCoef<ORDER>::c[threadIdx.x / ORDER][threadIdx.x % ORDER];
    // UL points per thread and pass, all their loads issued before any LDS store
    constexpr int UL = 4;
    const float rez = 1.0f / (float)ez, rey = 1.0f / (float)ey;
    for (int k0 = threadIdx.x; k0 < vol; k0 += UL * NT) {
        size_t g[UL];
        bool in[UL];
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            const int k = k0 + u * NT;
            in[u] = k < vol;
            const int kk = in[u] ? k : 0;
            const int row = fft_div(kk, rez), kz = kk - row * ez;
            const int kx = fft_div(row, rey), ky = row - kx * ey;
            g[u] = ((size_t)wrap_near(ox + kx, sx) * sy + wrap_near(oy + ky, sy)) * sz + wrap_near(oz + kz, sz);
        }
        float v0[UL], v1[UL], v2[UL];
        if (E4) {
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                const float4 v = E4[g[u]];
                v0[u] = v.x;
                v1[u] = v.y;
                v2[u] = v.z;
            }
        } else {
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                v0[u] = E[g[u]];
                v1[u] = E[ng + g[u]];
                v2[u] = E[2 * ng + g[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            if (!in[u]) continue;
            const int k = k0 + u * NT;
            fb[k] = v0[u];
            fb[vol + k] = v1[u];
            fb[2 * vol + k] = v2[u];
        }
    }
    __syncthreads();
    const float scale = bx->gatherScale;
    constexpr int NRW = (ORDER * ORDER + T - 1) / T;
    for (int cu = 0;;) {
        const int i = a0 + s;
        if (i < a1 && qi != 0.0f) {
            const int gx = rec_point(rc, 0), gy = rec_point(rc, 1), gz = rec_point(rc, 2);
            const int rx = wrap_near(gx - H - ox, sx), ry = wrap_near(gy - H - oy, sy), rz = wrap_near(gz - H - oz, sz);
            const bool inside = rx + ORDER <= ex && ry + ORDER <= ey && rz + ORDER <= ez;
            float wz[ORDER];
            weights<ORDER>(rc.z, wz);
            float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
            for (int rr = 0; rr < NRW; ++rr) {
                const int r = tl + rr * T;
                if (r < ORDER * ORDER) {
                    const int ix = r / ORDER, iy = r - ix * ORDER;
                    const float wxy = weight_at<ORDER>(lcoef, ix, rc.x) * weight_at<ORDER>(lcoef, iy, rc.y);
                    float e0[ORDER], e1[ORDER], e2[ORDER];
                    if (inside) {
                        const int b = ((rx + ix) * ey + (ry + iy)) * ez + rz;
#pragma unroll
                        for (int iz = 0; iz < ORDER; ++iz) {
                            e0[iz] = fb[b + iz];
                            e1[iz] = fb[vol + b + iz];
                            e2[iz] = fb[2 * vol + b + iz];
                        }
                    } else {
                        const size_t b = ((size_t)wrap_near(gx - H + ix, sx) * sy + wrap_near(gy - H + iy, sy)) * sz;
#pragma unroll
                        for (int iz = 0; iz < ORDER; ++iz) {
                            const size_t g = b + wrap_near(gz - H + iz, sz);
                            e0[iz] = E[g];
                            e1[iz] = E[ng + g];
                            e2[iz] = E[2 * ng + g];
                        }
                    }
#pragma unroll
                    for (int iz = 0; iz < ORDER; ++iz) {
                        const float w = wxy * wz[iz];
                        fx += w * e0[iz];
                        fy += w * e1[iz];
                        fz += w * e2[iz];
                    }
                }
            }
            fx = team_sum<T>(fx);
            fy = team_sum<T>(fy);
            fz = team_sum<T>(fz);
            if (tl == 0) {
                const float sc = scale * qi;
                const long long f0 = to_fx(sc * fx), f1 = to_fx(sc * fy), f2 = to_fx(sc * fz);
                atomicAdd(&force[3 * i + 0], (unsigned long long)f0);
                atomicAdd(&force[3 * i + 1], (unsigned long long)f1);
                atomicAdd(&force[3 * i + 2], (unsigned long long)f2);
                if (stored) {
                    const int id = ids[i];
                    stored[3 * id + 0] = f0;
                    stored[3 * id + 1] = f1;
                    stored[3 * id + 2] = f2;
                }
            }
        }
        s += SPR;
        if (a0 + s - (int)(threadIdx.x / T) >= a1) {  // uniform: the round's first site is past the end
            if (++cu >= ncol) break;  // the tile's next cell column
            a0 = cellStart[(col0 + cu) * tg.ncz + z0];
            a1 = cellStart[(col0 + cu) * tg.ncz + z1];
            s = threadIdx.x / T;
        }
        qi = 0.0f;
        if (a0 + s < a1) {
            qi = q[a0 + s];
            rc = rec[a0 + s];
        }
    }
}

// the stored mesh forces by id added to the current order's force rows
// (applyStoredForces, FixChargeEwald.cu:804-814)
__device__ __forceinline__ void apply_stored_body(int blk, int nblk, int n, const int *__restrict__ ids, const long long *__restrict__ stored,
                               unsigned long long *__restrict__ force) {
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int id = ids[i];
    const long long f0 = stored[3 * id], f1 = stored[3 * id + 1], f2 = stored[3 * id + 2];
    if ((f0 | f1 | f2) == 0) return;
    atomicAdd(&force[3 * i + 0], (unsigned long long)f0);
    atomicAdd(&force[3 * i + 1], (unsigned long long)f1);
    atomicAdd(&force[3 * i + 2], (unsigned long long)f2);
}


// Tile shape for the LDS spread: z-chunks of zs cells so that the tile stays
// within 48 KB of LDS.  Returns false (plain spread) when no cell list is
// valid, the box is not periodic, or even one cell does not fit.
static bool tiled_geometry_cols(const Ctx &c, TileGeo &tg, int budget, int cb) {
    const int order = c.pppm.order;
    tg.cb = cb;
    tg.ncz = c.nc[2];
    // the kernel takes cell widths and margin from the device box (box_derive);
    // the extents keep two mesh points of slack for a box that changes by a
    // barostat while the grid lives
    const double margin = 0.5 * c.skin + 1e-6 * c.geo.L[0];
    auto extent = [&](int a, int ncells) {
        double h = c.geo.L[a] / c.pppm.sz[a];
        return (int)std::floor((ncells * (c.geo.L[a] / c.nc[a]) + 2.0 * margin) / h) + order + 2;
    };
    tg.ext[0] = extent(0, cb);
    tg.ext[1] = extent(1, cb);
    if (tg.ext[0] >= c.pppm.sz[0] || tg.ext[1] >= c.pppm.sz[1]) return false;
    int zs = 0;
    for (int z = 1; z <= tg.ncz; ++z) {
        int ez = extent(2, z);
        if (ez >= c.pppm.sz[2] || tg.ext[0] * tg.ext[1] * ez > budget) break;
        zs = z;
    }
    if (zs == 0) return false;
    tg.zs = zs;
    // the spread flushes the box by z rows per wave: at most 64 points a row
    while (tg.zs > 1 && extent(2, tg.zs) > 64) --tg.zs;
    if (extent(2, tg.zs) > 64) return false;
    zs = tg.zs;
    tg.nzc = (tg.ncz + zs - 1) / zs;
    tg.ext[2] = extent(2, zs);
    return true;
}

// Tiles of 2 x 2 cell columns (option p3m_tile_cols 2, even column counts)
// where they fit the budget, else of single columns.
bool tiled_geometry(const Ctx &c, TileGeo &tg, int budget = 48 * 1024 / 8, int maxCols = 2) {
    if (!c.listValid || c.ncell == 0 || !c.colXY || c.mdxDisableTiledSpread) return false;
    for (int a = 0; a < 3; ++a)
        if (!c.geo.periodic[a]) return false;
    if (c.pppm.tileCols == 2 && maxCols >= 2 && c.nc[0] % 2 == 0 && c.nc[1] % 2 == 0 &&
        tiled_geometry_cols(c, tg, budget, 2))
        return true;
    return tiled_geometry_cols(c, tg, budget, 1);
}

// Geometry of the owner-computes spread, or false (another spread) when no
// cell list is valid, an axis is open, a mesh axis is beyond the records'
// 10 bits, x or y has fewer than 3 cells, or a site within the margin of its
// cell could reach past the cells next to it in x or y or past 8 cells in z.
// The margin is skin/2 with a quarter's slack (a barostat changes h and the
// margin's share of it while the grid lives).
bool own_geometry(const Ctx &c, OwnGeo &og, size_t &shm) {
    if (!c.listValid || c.ncell == 0 || !c.colXY || !c.meshRec || !c.hasCharges || c.mdxDisableTiledSpread)
        return false;
    const int order = c.pppm.order, H = (order - 1) / 2;
    for (int a = 0; a < 3; ++a) {
        if (!c.geo.periodic[a] || c.pppm.sz[a] > 1024 || c.pppm.sz[a] < order + 1) return false;
        og.nc[a] = c.nc[a];
        og.sz[a] = c.pppm.sz[a];
    }
    if (og.nc[0] < 3 || og.nc[1] < 3) return false;
    // reach: every cell's sites keep their stencils within [cell - R, cell + R]
    auto reaches = [&](int a, int R) {
        const int nc = og.nc[a], sz = og.sz[a];
        const double h = c.geo.L[a] / sz, r = (double)sz / nc;
        const double mu = (0.5 * c.skin + 1e-6 * c.geo.L[0]) / h * 1.25 + 0.01;
        for (int k = 0; k < nc; ++k) {
            const int glo = (int)std::floor(k * r - mu + 0.5), ghi = (int)std::floor((k + 1) * r + mu + 0.5);
            if (glo - H < own_start(k - R, sz, nc) || ghi + H >= own_start(k + 1 + R, sz, nc)) return false;
        }
        return true;
    };
    if (!reaches(0, 1) || !reaches(1, 1)) return false;
    og.rz = 0;
    for (int R = 1; R <= 8 && R <= og.nc[2] && !og.rz; ++R)
        if (reaches(2, R)) og.rz = R;
    if (!og.rz) return false;
    // z cells per workgroup: about 512 workgroups, the owned points within 48 KB of LDS
    auto most = [&](int a, int span) {
        int m = 0;
        for (int k = 0; k < og.nc[a]; ++k) {
            const int e = std::min(k + span, og.nc[a]);
            m = std::max(m, own_start(e, og.sz[a], og.nc[a]) - own_start(k, og.sz[a], og.nc[a]));
        }
        return m;
    };
    const int mx = most(0, 1), my = most(1, 1);
    const int ncol = og.nc[0] * og.nc[1];
    int zs = std::max(1, std::min({og.nc[2], ncol * og.nc[2] / 512, kOwnZMax}));
    while (zs > 1 && (size_t)mx * my * most(2, zs) * 8 > 48 * 1024) --zs;
    if ((size_t)mx * my * most(2, zs) * 8 > 48 * 1024) return false;
    og.zs = zs;
    og.nzc = (og.nc[2] + zs - 1) / zs;
    og.whole = zs + 2 * og.rz >= og.nc[2] ? 1 : 0;
    shm = (size_t)mx * my * most(2, zs) * 8;
    return true;
}

// clears n 64-bit words (the charge mesh of the scratch)
__device__ __forceinline__ void scr_zero_body(int blk, int nblk, long long n, unsigned long long *__restrict__ p) {
    const long long i = (long long)blk * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of mdx_forces_compute_noalloc: the P3M work arrays for meshes of up
// to kScratchMeshMax points (charge mesh int64, real mesh, half spectrum, the
// three field spectra, the three real field components, the packed field),
// about 14 MiB per GPU, held until the process ends.  Host code never uses
// the symbol itself (a host-side stand-in): it takes the device address from
// hipGetSymbolAddress, once per device, with that device current.  Calls that
// overlap on one GPU share this memory and race.
namespace {
constexpr size_t kScrAlign = 256;
constexpr size_t kScrNh = kScratchMeshMax / 2 + kScratchMeshMax / 32;  // sx sy (sz / 2 + 1) for sz >= 32
constexpr size_t scr_round(size_t b) { return (b + kScrAlign - 1) / kScrAlign * kScrAlign; }
constexpr size_t kScrMesh = 0;
constexpr size_t kScrReal = kScrMesh + scr_round(sizeof(long long) * kScratchMeshMax);
constexpr size_t kScrRho = kScrReal + scr_round(sizeof(float) * kScratchMeshMax);
constexpr size_t kScrEHat = kScrRho + scr_round(sizeof(float2) * kScrNh);
constexpr size_t kScrEReal = kScrEHat + scr_round(sizeof(float2) * 3 * kScrNh);
constexpr size_t kScrField = kScrEReal + scr_round(sizeof(float) * 3 * kScratchMeshMax);
constexpr size_t kScrBytes = kScrField + scr_round(sizeof(float4) * kScratchMeshMax);
constexpr int kScrMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScrBytes];

static unsigned char *g_scratch[kScrMaxDevices] = {};

void pppm_scratch_init(int device) {
    MDX_REQUIRE(device >= 0 && device < kScrMaxDevices, 1, "device index beyond the scratch table");
    if (g_scratch[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    MDX_REQUIRE(e == hipSuccess && p, 2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (unsigned char *)p;
}

PppmScratchScope::PppmScratchScope(Ctx &c_) : c(c_) {
    PppmState &P = c.pppm;
    mesh = P.mesh;
    real = P.real;
    rhoHat = P.rhoHat;
    eHat = P.eHat;
    eReal = P.eReal;
    eField4 = P.eField4;
    const long long ng = (long long)P.sz[0] * P.sz[1] * P.sz[2];
    const long long nh = (long long)P.sz[0] * P.sz[1] * (P.sz[2] / 2 + 1);
    MDX_REQUIRE(ng <= kScratchMeshMax && nh <= (long long)kScrNh, 1,
                "mesh larger than the scratch of mdx_forces_compute_noalloc (64^3 points)");
    unsigned char *base = g_scratch[c.device];
    MDX_REQUIRE(base, 3, "scratch address not looked up on this device");
    P.mesh = (long long *)(base + kScrMesh);
    P.real = (float *)(base + kScrReal);
    P.rhoHat = (float2 *)(base + kScrRho);
    P.eHat = (float2 *)(base + kScrEHat);
    P.eReal = (float *)(base + kScrEReal);
    P.eField4 = (float4 *)(base + kScrField);
    // the spread adds into the charge mesh, which the last call on this GPU
    // left in any state: clear it with a kernel (no memset node)
    launch<scr_zero_body, 256>(c, cdiv(ng, 256), 0, ng, (unsigned long long *)P.mesh);
}

PppmScratchScope::~PppmScratchScope() {
    PppmState &P = c.pppm;
    P.mesh = mesh;
    P.real = real;
    P.rhoHat = rhoHat;
    P.eHat = eHat;
    P.eReal = eReal;
    P.eField4 = eField4;
}

void pppm_release(Ctx &c) {
    if (c.pppm.plansMade) {
        hipfftDestroy(c.pppm.planR2C);
        hipfftDestroy(c.pppm.planC2R);
        c.pppm.plansMade = false;
    }
}

// alpha + influence function for the current box (handleBoundsChangeInternal,
// FixChargeEwald.cu:1206-1229): recomputed whenever the box changed.
void pppm_prepare(Ctx &c) {
    PppmState &P = c.pppm;
    if (P.errorTarget > 0) {
        bool same = true;
        for (int a = 0; a < 3; ++a) same = same && (P.boxAtGreen[a] == c.geo.L[a]);
        if (!same || !P.plansMade) pppm_mesh_for_error(c);
    }
    const int *sz = P.sz;
    const int szh = sz[2] / 2 + 1;
    const size_t ng = (size_t)sz[0] * sz[1] * sz[2];
    const size_t nh = (size_t)sz[0] * sz[1] * szh;
    if (!P.plansMade) {
        P.mesh = dalloc<long long>(c, ng);
        P.real = dalloc<float>(c, ng);
        P.rhoHat = dalloc<float2>(c, nh);
        P.eHat = dalloc<float2>(c, 3 * nh);
        P.eReal = dalloc<float>(c, 3 * ng);
        P.eField4 = dalloc<float4>(c, ng);
        P.green = dalloc<float>(c, nh);
        MDX_HIP(hipMemsetAsync(P.mesh, 0, sizeof(long long) * ng, c.stream));
        MDX_REQUIRE(hipfftPlan3d(&P.planR2C, sz[0], sz[1], sz[2], HIPFFT_R2C) == HIPFFT_SUCCESS, 2,
                    "hipfftPlan3d R2C failed");
        int dims[3] = {sz[0], sz[1], sz[2]};
        MDX_REQUIRE(hipfftPlanMany(&P.planC2R, 3, dims, nullptr, 1, (int)nh, nullptr, 1, (int)ng,
                                   HIPFFT_C2R, 3) == HIPFFT_SUCCESS,
                    2, "hipfftPlanMany C2R failed");
        hipfftSetStream(P.planR2C, c.stream);
        hipfftSetStream(P.planC2R, c.stream);
        P.planStream = c.stream;
        P.plansMade = true;
    }
    bool same = true;
    for (int a = 0; a < 3; ++a) same = same && (P.boxAtGreen[a] == c.geo.L[a]);
    if (same && P.alpha > 0) return;
    if (P.alphaUser > 0) {
        P.alpha = P.alphaUser;
        P.err = 0.0;
    } else {
        P.alpha = pppm_solve_alpha(P.alphaAtoms > 0 ? P.alphaAtoms : c.n, P.totalQ2, c.geo.L, sz, P.order, P.rcut,
                                   &P.err);
    }
    c.alpha = (float)P.alpha;
    box_upload(c);  // the Green function reads alpha and the mesh from the device box
    pppm_green_device(c, nullptr);
    for (int a = 0; a < 3; ++a) P.boxAtGreen[a] = c.geo.L[a];
}

bool pppm_green_device(Ctx &c, const int *gate, const DilateArgs *dl) {
    const PppmState &P = c.pppm;
    const int szh = P.sz[2] / 2 + 1;
    const size_t nh = (size_t)P.sz[0] * P.sz[1] * szh;
    // table of 3 (2 kGreenTabLim + 1) (sz_x + sz_y + sz_z) floats; meshes too
    // large for it take the direct kernel
    const size_t tab = sizeof(float) * 3 * (2 * kGreenTabLim + 1) * (P.sz[0] + P.sz[1] + P.sz[2]);
    const int gb = cdiv(nh, 256);
    if (tab <= 64 * 1024) {
        DilateArgs none{};
        const int db = dl ? cdiv(std::max(dl->nmol, dl->n), 256) : 0;
        const int vb = dl ? dl->vblocks : 0;
        launch<green_body, 256>(c, gb + db + vb, tab, gate, c.box, szh, P.order, P.green, gb, dl ? *dl : none);
        return dl != nullptr;
    }
    launch<green_direct_body, 256>(c, gb, 0, gate, c.box, szh, P.order, P.green);
    return false;
}

// The transforms run on the stream their context's work is on (a ring
// polymer's beads share one, a batch records its replicas onto one).
static void pppm_plans_on(PppmState &P, hipStream_t s) {
    if (P.planStream == s) return;
    MDX_REQUIRE(hipfftSetStream(P.planR2C, s) == HIPFFT_SUCCESS && hipfftSetStream(P.planC2R, s) == HIPFFT_SUCCESS, 2,
                "hipfftSetStream failed");
    P.planStream = s;
}

// LDS bytes of passes 1 and 3 of the mesh solve, 0 where the mesh does not
// allow it (every axis a power of two from 4 to 128, the planes within
// 60 KB).  Pass 2 takes kFftMax + 2 * kFftWavePts points per wave.
static size_t lds_fft_shm(const PppmState &P) {
    for (int a = 0; a < 3; ++a) {
        const int n = P.sz[a];
        if (n < 4 || n > kFftMax || (n & (n - 1))) return 0;
    }
    const size_t szh = P.sz[2] / 2 + 1;
    const size_t cap = std::max((size_t)(P.sz[1] / 2) * (P.sz[2] + 1), (size_t)P.sz[1] * szh);
    const size_t b = sizeof(float2) * (kFftMax + 3 * cap);
    return b <= 60 * 1024 ? b : 0;
}

bool pppm_lds_fft(const PppmState &P) { return P.fftMode == 1 && lds_fft_shm(P) > 0; }

// Charge mesh -> the planar real field P.eReal, with the k-space energy and
// virial when ev is set (and the gate, if any, open).  fromReal: the charges
// are in P.real (owner spread) instead of the fixed-point mesh.
static void mesh_solve(Ctx &c, bool fromReal, int ev, const int *gate) {
    PppmState &P = c.pppm;
    const int *sz = P.sz;
    const int szh = sz[2] / 2 + 1;
    const int ng = sz[0] * sz[1] * sz[2];
    const int nh = sz[0] * sz[1] * szh;
    const int threads = 256;
    const size_t shm = pppm_lds_fft(P) ? lds_fft_shm(P) : 0;
    if (shm) {
        P.lastFft = 1;
        ProfScope ps(c, PK_FFT);
        const int waves = kFftThreads / 64;
        const int lw = kFftWavePts / sz[0], tasks = sz[1] * ((szh + lw - 1) / lw) * (ev ? 4 : 3);
        const size_t shm2 = sizeof(float2) * (kFftMax + waves * 2 * kFftWavePts);
        launch<fft_pass1_body, kFftPlaneThreads>(c, sz[0], shm, c.box, fromReal ? nullptr : P.mesh,
                                                 fromReal ? P.real : nullptr, P.rhoHat);
        launch<fft_pass2_body, kFftThreads>(c, cdiv(tasks, waves), shm2, c.box, P.green, P.rhoHat, P.eHat, ev, gate,
                                            c.acc);
        launch<fft_pass3_body, kFftPlaneThreads>(c, 3 * sz[0], shm, c.box, P.eHat, P.eReal);
        return;
    }
    P.lastFft = 0;
    if (!fromReal) {
        ProfScope ps(c, PK_SPREAD);
        launch<mesh_to_real_body, 256>(c, cdiv(ng, threads), 0, ng, P.mesh, P.real);
    }
    {
        ProfScope ps(c, PK_FFT);
        stream_op(c, [&P](hipStream_t s) {
            pppm_plans_on(P, s);
            MDX_REQUIRE(hipfftExecR2C(P.planR2C, P.real, (hipfftComplex *)P.rhoHat) == HIPFFT_SUCCESS, 2,
                        "hipfftExecR2C failed");
        }, 1);
    }
    {
        ProfScope ps(c, PK_EFIELD);
        launch<efield_body, 256>(c, cdiv(nh, threads), 0, c.box, szh, P.green, P.rhoHat, P.eHat, ev, gate, c.acc);
    }
    {
        ProfScope ps(c, PK_FFT);
        stream_op(c, [&P](hipStream_t s) {
            pppm_plans_on(P, s);
            MDX_REQUIRE(hipfftExecC2R(P.planC2R, (hipfftComplex *)P.eHat, P.eReal) == HIPFFT_SUCCESS, 2,
                        "hipfftExecC2R failed");
        }, 2);
    }
}

void pppm_forces(Ctx &c, bool energy, bool virial) {
    PppmState &P = c.pppm;
    pppm_prepare(c);
    const int *sz = P.sz;
    const int szh = sz[2] / 2 + 1;
    const int ng = sz[0] * sz[1] * sz[2];
    const int nh = sz[0] * sz[1] * szh;
    const float qunit = (float)std::sqrt(P.qqr);
    const int threads = 256;
    const int cur = c.cur;
    {
    ProfScope ps(c, PK_SPREAD);
    TileGeo tg;
    OwnGeo og;
    size_t oshm = 0;
    if (P.spreadMode == 2 && own_geometry(c, og, oshm)) {
        P.lastSpread = 2;
        const int nlog = c.nc[0] * c.nc[1] * og.nzc;
        switch (P.order) {
            case 1: launch<spread_own_body<1, 256>, 256>(c, xcd_grid(nlog), oshm, c.meshRec, c.q[cur], c.cellStart, c.colRank, c.colXY, og, qunit, nlog, P.real, c.dFlags); break;
            case 3: launch<spread_own_body<3, 256>, 256>(c, xcd_grid(nlog), oshm, c.meshRec, c.q[cur], c.cellStart, c.colRank, c.colXY, og, qunit, nlog, P.real, c.dFlags); break;
            case 5: launch<spread_own_body<5, 256>, 256>(c, xcd_grid(nlog), oshm, c.meshRec, c.q[cur], c.cellStart, c.colRank, c.colXY, og, qunit, nlog, P.real, c.dFlags); break;
            default: launch<spread_own_body<7, 256>, 256>(c, xcd_grid(nlog), oshm, c.meshRec, c.q[cur], c.cellStart, c.colRank, c.colXY, og, qunit, nlog, P.real, c.dFlags); break;
        }
    } else {
    const bool tiled = P.spreadMode >= 1 && tiled_geometry(c, tg, P.spreadKB * 1024 / 8);
    P.lastSpread = tiled ? 1 : 0;
    const float4 *rec = c.meshRec && c.hasCharges ? c.meshRec : nullptr;
    if (tiled && rec) {
        const int nblk = c.nc[0] * c.nc[1] / (tg.cb * tg.cb) * tg.nzc;
        const size_t shm = sizeof(unsigned long long) * tg.ext[0] * tg.ext[1] * tg.ext[2];
        switch (P.order) {
            case 1: launch<spread_rows_body<1, 256>, 256>(c, xcd_grid(nblk), shm, c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
            case 3: launch<spread_rows_body<3, 256>, 256>(c, xcd_grid(nblk), shm, c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
            case 5: launch<spread_rows_body<5, 256>, 256>(c, xcd_grid(nblk), shm, c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
            default: launch<spread_rows_body<7, 256>, 256>(c, xcd_grid(nblk), shm, c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
        }
    } else if (tiled && tiled_geometry(c, tg, 48 * 1024 / 8, 1)) {  // single-column tiles (spread_tiled_body)
        const int nblk = c.nc[0] * c.nc[1] / (tg.cb * tg.cb) * tg.nzc;
        const size_t shm = sizeof(unsigned long long) * tg.ext[0] * tg.ext[1] * tg.ext[2];
        switch (P.order) {
            case 1: launch<spread_tiled_body<1>, 256>(c, xcd_grid(nblk), shm, c.pos[cur], c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
            case 3: launch<spread_tiled_body<3>, 256>(c, xcd_grid(nblk), shm, c.pos[cur], c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
            case 5: launch<spread_tiled_body<5>, 256>(c, xcd_grid(nblk), shm, c.pos[cur], c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
            default: launch<spread_tiled_body<7>, 256>(c, xcd_grid(nblk), shm, c.pos[cur], c.q[cur], c.box, qunit, c.cellStart, c.colXY, tg, nblk, (unsigned long long *)P.mesh, rec); break;
        }
    } else
    switch (P.order) {
        case 1: launch<spread_body<1>, 256>(c, cdiv(c.n, threads), 0, c.n, c.pos[cur], c.q[cur], c.box, qunit, (unsigned long long *)P.mesh); break;
        case 3: launch<spread_body<3>, 256>(c, cdiv(c.n, threads), 0, c.n, c.pos[cur], c.q[cur], c.box, qunit, (unsigned long long *)P.mesh); break;
        case 5: launch<spread_body<5>, 256>(c, cdiv(c.n, threads), 0, c.n, c.pos[cur], c.q[cur], c.box, qunit, (unsigned long long *)P.mesh); break;
        default: launch<spread_body<7>, 256>(c, cdiv(c.n, threads), 0, c.n, c.pos[cur], c.q[cur], c.box, qunit, (unsigned long long *)P.mesh); break;
    }
    }
    }
    // energy and virial in the field pass; a trace-only request (a
    // barostat's) is skipped on steps it does not serve
    mesh_solve(c, P.lastSpread == 2, (energy || virial) ? 1 : 0, c.vtrace ? baro_gate(c) : nullptr);
    // a built list puts the charged atoms first: only they take the gather
    const int na = c.listValid ? c.nCharged : c.n;
    ProfScope ps(c, PK_GATHER);
    if (na == 0) return;
    long long *stored = P.interval > 1 ? P.stored : nullptr;
    TileGeo gg;
    if (P.gatherMode == 2 && c.meshRec && c.hasCharges && tiled_geometry(c, gg, P.tileKB * 1024 / 12)) {
        P.lastGather = 2;
        const float4 *e4 = nullptr;
        if (!P.tilePlanar) {
            launch<field_pack_body, 256>(c, cdiv(ng, threads), 0, ng, P.eReal, P.eField4);
            e4 = P.eField4;
        }
        const int nlog = c.nc[0] * c.nc[1] / (gg.cb * gg.cb) * gg.nzc;
        const size_t shm = sizeof(float) * 3 * gg.ext[0] * gg.ext[1] * gg.ext[2];
#define MDX_GTILE(O, T) \
    launch<gather_tile_body<O, T, 256>, 256>(c, xcd_grid(nlog), shm, c.meshRec, c.q[cur], c.box, c.cellStart, c.colXY, gg, nlog, P.eReal, e4, c.force, c.id[cur], stored)
#define MDX_GTILE_T(O) \
    if (P.tileTeam == 4) MDX_GTILE(O, 4); else if (P.tileTeam == 16) MDX_GTILE(O, 16); else MDX_GTILE(O, 8)
        switch (P.order) {
            case 1: MDX_GTILE_T(1); break;
            case 3: MDX_GTILE_T(3); break;
            case 5: MDX_GTILE_T(5); break;
            default: MDX_GTILE_T(7); break;
        }
#undef MDX_GTILE_T
#undef MDX_GTILE
        return;
    }
    launch<field_pack_body, 256>(c, cdiv(ng, threads), 0, ng, P.eReal, P.eField4);
    // 1 (and 2 where the tiles do not fit): the mesh points and offsets from
    // the position refresh's records; 0: computed here from the positions
    const float4 *grec = P.gatherMode >= 1 && c.meshRec && c.hasCharges ? c.meshRec : nullptr;
    P.lastGather = grec ? 1 : 0;
    const int gridx = xcd_grid(cdiv((long long)na * P.gatherTeam, threads));
#define MDX_GATHER(O, T) \
    launch<gather_body<O, T>, 256>(c, gridx, 0, na, c.pos[cur], c.q[cur], c.box, P.eField4, c.force, c.id[cur], stored, grec)
    const bool wide = P.gatherTeam == 16;
    switch (P.order) {
        case 1: if (wide) MDX_GATHER(1, 16); else MDX_GATHER(1, 8); break;
        case 3: if (wide) MDX_GATHER(3, 16); else MDX_GATHER(3, 8); break;
        case 5: if (wide) MDX_GATHER(5, 16); else MDX_GATHER(5, 8); break;
        default: if (wide) MDX_GATHER(7, 16); else MDX_GATHER(7, 8); break;
    }
#undef MDX_GATHER
}

// Mesh part of a ring polymer on the first bead's mesh: charges spread from
// the centroids (by id), field gathered at the centroids into r.cf, then the
// centroid force onto every bead (FixChargeEwald.cu:1255-1266, 1395-1405).
// Same kernels as a single system's chain; alpha and G are the first bead's.
void ring_pppm(Ctx &c0, const RingSpec &r) {
    PppmState &P = c0.pppm;
    MDX_REQUIRE(P.on && P.plansMade, 3, "ring contraction needs P3M set up on the first bead");
    const int *sz = P.sz;
    const int szh = sz[2] / 2 + 1;
    const int ng = sz[0] * sz[1] * sz[2];
    const int nh = sz[0] * sz[1] * szh;
    const float qunit = (float)std::sqrt(P.qqr);
    const int threads = 256;
    ring_centroids(r, c0);
    {
        ProfScope ps(c0, PK_SPREAD);
        switch (P.order) {
            case 1: launch<spread_body<1>, 256>(c0, cdiv(r.natoms, threads), 0, r.natoms, r.cen, r.q, c0.box, qunit, (unsigned long long *)P.mesh); break;
            case 3: launch<spread_body<3>, 256>(c0, cdiv(r.natoms, threads), 0, r.natoms, r.cen, r.q, c0.box, qunit, (unsigned long long *)P.mesh); break;
            case 5: launch<spread_body<5>, 256>(c0, cdiv(r.natoms, threads), 0, r.natoms, r.cen, r.q, c0.box, qunit, (unsigned long long *)P.mesh); break;
            default: launch<spread_body<7>, 256>(c0, cdiv(r.natoms, threads), 0, r.natoms, r.cen, r.q, c0.box, qunit, (unsigned long long *)P.mesh); break;
        }
    }
    mesh_solve(c0, false, 0, nullptr);
    ProfScope ps(c0, PK_GATHER);
    launch<field_pack_body, 256>(c0, cdiv(ng, threads), 0, ng, P.eReal, P.eField4);
    const dim3 grid(xcd_grid(cdiv((long long)r.natoms * 8, threads)));
    unsigned long long *cf = (unsigned long long *)r.cf;
    switch (P.order) {
        case 1: launch<gather_body<1, 8>, 256>(c0, grid.x, 0, r.natoms, r.cen, r.q, c0.box, P.eField4, cf, nullptr, nullptr, nullptr); break;
        case 3: launch<gather_body<3, 8>, 256>(c0, grid.x, 0, r.natoms, r.cen, r.q, c0.box, P.eField4, cf, nullptr, nullptr, nullptr); break;
        case 5: launch<gather_body<5, 8>, 256>(c0, grid.x, 0, r.natoms, r.cen, r.q, c0.box, P.eField4, cf, nullptr, nullptr, nullptr); break;
        default: launch<gather_body<7, 8>, 256>(c0, grid.x, 0, r.natoms, r.cen, r.q, c0.box, P.eField4, cf, nullptr, nullptr, nullptr); break;
    }
    ring_apply(r, c0);
}

void pppm_step(Ctx &c, bool energy, bool virial) {
    PppmState &P = c.pppm;
    if (P.interval > 1 && c.inRun) {
        if (!P.stored) {
            P.stored = dalloc<long long>(c, 3 * (size_t)c.n);
            MDX_HIP(hipMemsetAsync(P.stored, 0, sizeof(long long) * 3 * c.n, c.stream));
            P.storedValid = false;
        }
        const bool due = (c.turn - P.turnInit) % P.interval == 0;
        if (!due && P.storedValid && !energy && !virial && !c.vtrace) {
            launch<apply_stored_body, 256>(c, cdiv(c.n, 256), 0, c.n, c.id[c.cur], P.stored, c.force);
            return;
        }
        pppm_forces(c, energy, virial);
        P.storedValid = c.listValid;  // a gather without a valid list visits every atom; stored is complete too
        return;
    }
    pppm_forces(c, energy, virial);
}

void pppm_mesh_for_error(Ctx &c) {
    PppmState &P = c.pppm;
    int sz[3] = {P.sz[0], P.sz[1], P.sz[2]};
    double err = 0.0;
    pppm_solve_alpha(c.n, P.totalQ2, c.geo.L, sz, P.order, P.rcut, &err);
    int tries = 0;
    while (tries < 100 && (err > P.errorTarget || err != err || err < 0)) {
        // double the axis with the fewest mesh points per length
        int k = 0;
        for (int a = 1; a < 3; ++a)
            if (sz[a] / c.geo.L[a] < sz[k] / c.geo.L[k]) k = a;
        MDX_REQUIRE(sz[k] < 1024, 1, "P3M error target needs a mesh beyond 1024 points per axis");
        sz[k] *= 2;
        pppm_solve_alpha(c.n, P.totalQ2, c.geo.L, sz, P.order, P.rcut, &err);
        ++tries;
    }
    if (sz[0] == P.sz[0] && sz[1] == P.sz[1] && sz[2] == P.sz[2]) return;
    // a new mesh: new work arrays, plans, Green function and device box
    if (P.plansMade) {
        pppm_release(c);
        dfree(c, P.mesh);
        dfree(c, P.real);
        dfree(c, P.rhoHat);
        dfree(c, P.eHat);
        dfree(c, P.eReal);
        dfree(c, P.eField4);
        dfree(c, P.green);
        P.mesh = nullptr;
        P.real = nullptr;
        P.rhoHat = P.eHat = nullptr;
        P.eReal = nullptr;
        P.eField4 = nullptr;
        P.green = nullptr;
    }
    for (int a = 0; a < 3; ++a) {
        P.sz[a] = sz[a];
        P.boxAtGreen[a] = 0.0;
    }
    P.alpha = 0.0;
    P.storedValid = false;
}

}  // namespace mdx
