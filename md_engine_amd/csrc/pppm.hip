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

constexpr double kPi = 3.14159265358979323846;

// W_p for orders 1/3/5/7 (FixChargeEwald.cu:111-115, 163-172, 220-234)
template <int ORDER>
__device__ __forceinline__ void weights(float x, float *w) {
    const float x2 = x * x;
    if (ORDER == 1) {
        w[0] = 1.0f;
    } else if (ORDER == 3) {
        w[0] = 0.125f - 0.5f * x + 0.5f * x2;
        w[1] = 0.75f - x2;
        w[2] = 0.125f + 0.5f * x + 0.5f * x2;
    } else if (ORDER == 5) {
        const float x3 = x2 * x, x4 = x2 * x2;
        w[0] = (1.0f / 384.0f) * (1.0f - 8.0f * x + 24.0f * x2 - 32.0f * x3 + 16.0f * x4);
        w[1] = (1.0f / 96.0f) * (19.0f - 44.0f * x + 24.0f * x2 + 16.0f * x3 - 16.0f * x4);
        w[2] = (1.0f / 192.0f) * (115.0f - 120.0f * x2 + 48.0f * x4);
        w[3] = (1.0f / 96.0f) * (19.0f + 44.0f * x + 24.0f * x2 - 16.0f * x3 - 16.0f * x4);
        w[4] = (1.0f / 384.0f) * (1.0f + 8.0f * x + 24.0f * x2 + 32.0f * x3 + 16.0f * x4);
    } else {
        const float x3 = x2 * x, x4 = x2 * x2, x5 = x3 * x2, x6 = x3 * x3;
        w[0] = (1.0f / 46080.0f) * (1.0f - 12.0f * x + 60.0f * x2 - 160.0f * x3 + 240.0f * x4 - 192.0f * x5 + 64.0f * x6);
        w[1] = (1.0f / 23040.0f) * (361.0f - 1416.0f * x + 2220.0f * x2 - 1600.0f * x3 + 240.0f * x4 + 384.0f * x5 - 192.0f * x6);
        w[2] = (1.0f / 46080.0f) * (10543.0f - 17340.0f * x + 4740.0f * x2 + 6880.0f * x3 - 4080.0f * x4 - 960.0f * x5 + 960.0f * x6);
        w[3] = (1.0f / 11520.0f) * (5887.0f - 4620.0f * x2 + 1680.0f * x4 - 320.0f * x6);
        w[4] = (1.0f / 46080.0f) * (10543.0f + 17340.0f * x + 4740.0f * x2 - 6880.0f * x3 - 4080.0f * x4 + 960.0f * x5 + 960.0f * x6);
        w[5] = (1.0f / 23040.0f) * (361.0f + 1416.0f * x + 2220.0f * x2 + 1600.0f * x3 + 240.0f * x4 - 384.0f * x5 - 192.0f * x6);
        w[6] = (1.0f / 46080.0f) * (1.0f + 12.0f * x + 60.0f * x2 + 160.0f * x3 + 240.0f * x4 + 192.0f * x5 + 64.0f * x6);
    }
}

struct MeshGeo {
    int sz[3];
    double lo[3], L[3], invL[3], h[3];
    int periodic[3];
};

// Nearest mesh point and offset, FP64 like the oracle: the coordinate is
// wrapped into [lo, lo + L) first, then ngp = int((pos + h/2) / h) and
// d = pos / h - ngp (FixChargeEwald.cu:178-190).
__device__ __forceinline__ void mesh_point(double x, int a, const MeshGeo &m, int &ngp, float &d) {
    double img = floor((x - m.lo[a]) * m.invL[a]);
    double xw = x - img * m.L[a];
    double p = xw - m.lo[a];
    ngp = (int)((p + 0.5 * m.h[a]) / m.h[a]);
    d = (float)(p / m.h[a] - ngp);
}

__device__ __forceinline__ int wrapi(int p, int n) {
    p %= n;
    return p < 0 ? p + n : p;
}

template <int ORDER>
__global__ void __launch_bounds__(256) k_spread(int n, const double4 *__restrict__ pos,
                                                const float *__restrict__ q, MeshGeo m, float qunit,
                                                unsigned long long *__restrict__ mesh) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
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
struct TileGeo {
    int ncz, zs, nzc;
    int ext[3];
    double cellw[3];
    double margin;
};

__device__ __forceinline__ int tile_origin(double lo_pos, double h) {
    // same rounding as mesh_point on the lowest reachable coordinate
    return (int)floor((lo_pos + 0.5 * h) / h);
}

template <int ORDER>
__global__ void __launch_bounds__(256) k_spread_tiled(const double4 *__restrict__ pos,
                                                      const float *__restrict__ q, MeshGeo m,
                                                      float qunit, const int *__restrict__ cellStart,
                                                      const int2 *__restrict__ colXY, TileGeo tg,
                                                      unsigned long long *__restrict__ mesh) {
    extern __shared__ unsigned long long tile[];
    constexpr int H = (ORDER - 1) / 2;
    const int col = blockIdx.x / tg.nzc;
    const int zc = blockIdx.x % tg.nzc;
    const int z0 = zc * tg.zs;
    const int z1 = min(z0 + tg.zs, tg.ncz);
    const int a0 = cellStart[col * tg.ncz + z0];
    const int a1 = cellStart[col * tg.ncz + z1];
    const int2 cxy = colXY[col];
    const int ox = tile_origin(cxy.x * tg.cellw[0] - tg.margin, m.h[0]) - H;
    const int oy = tile_origin(cxy.y * tg.cellw[1] - tg.margin, m.h[1]) - H;
    const int oz = tile_origin(z0 * tg.cellw[2] - tg.margin, m.h[2]) - H;
    const int ex = tg.ext[0], ey = tg.ext[1], ez = tg.ext[2];
    const int vol = ex * ey * ez;
    for (int k = threadIdx.x; k < vol; k += blockDim.x) tile[k] = 0ull;
    __syncthreads();
    const int items = (a1 - a0) * ORDER;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
        const int i = a0 + it / ORDER;
        const int ix = it % ORDER;
        const float qi = q[i];
        if (qi == 0.0f) continue;
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
    for (int k = threadIdx.x; k < vol; k += blockDim.x) {
        const unsigned long long v = tile[k];
        if (v == 0ull) continue;
        const int kz = k % ez, ky = (k / ez) % ey, kx = k / (ez * ey);
        const int gx = wrapi(ox + kx, m.sz[0]), gy = wrapi(oy + ky, m.sz[1]), gz = wrapi(oz + kz, m.sz[2]);
        atomicAdd(&mesh[((size_t)gx * m.sz[1] + gy) * m.sz[2] + gz], v);
    }
}

__global__ void k_mesh_to_real(int ng, long long *__restrict__ mesh, float *__restrict__ out) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
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

// Optimal influence function on the half spectrum, FP64 arithmetic.
__global__ void k_green(MeshGeo m, int szh, double alpha, int order, int lim, float *__restrict__ G) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int total = m.sz[0] * m.sz[1] * szh;
    if (idx >= total) return;
    int iz = idx % szh, iy = (idx / szh) % m.sz[1], ix = idx / (szh * m.sz[1]);
    double k[3] = {kcomp(ix, m.sz[0], m.L[0]), kcomp(iy, m.sz[1], m.L[1]), kcomp(iz, m.sz[2], m.L[2])};
    double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    if (k2 == 0.0) {
        G[idx] = 0.0f;
        return;
    }
    double f4a2 = 0.25 / alpha / alpha;
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
    G[idx] = (float)(s1 / (s2 * s2) / k2);
}

// E_hat_d = i k_d G rho_hat (E_field_cu), k_d := 0 on the d-th Nyquist plane.
__global__ void k_efield(MeshGeo m, int szh, const float *__restrict__ G,
                         const float2 *__restrict__ rho, float2 *__restrict__ E) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int total = m.sz[0] * m.sz[1] * szh;
    if (idx >= total) return;
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
}

// Energy 0.5 sum |rho_hat|^2 G / V over the full spectrum and the k-space
// virial (virials_cu): half-spectrum entries with 0 < iz < sz/2 count twice.
__global__ void __launch_bounds__(256) k_kspace_ev(MeshGeo m, int szh, double alpha, double invV,
                                                   const float *__restrict__ G,
                                                   const float2 *__restrict__ rho,
                                                   unsigned long long *__restrict__ acc) {
    __shared__ double wred[4][7];
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int total = m.sz[0] * m.sz[1] * szh;
    double vals[7] = {0, 0, 0, 0, 0, 0, 0};
    if (idx < total) {
        int iz = idx % szh, iy = (idx / szh) % m.sz[1], ix = idx / (szh * m.sz[1]);
        double w = (iz == 0 || 2 * iz == m.sz[2]) ? 1.0 : 2.0;
        double k[3] = {kcomp(ix, m.sz[0], m.L[0]), kcomp(iy, m.sz[1], m.L[1]), kcomp(iz, m.sz[2], m.L[2])};
        double kl = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        float2 r = rho[idx];
        double E = w * ((double)r.x * r.x + (double)r.y * r.y) * G[idx];
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

// weight of stencil index k (0 <= k < ORDER) without a dynamically indexed
// local array (which would live in scratch memory)
template <int ORDER>
__device__ __forceinline__ float pick(const float (&w)[ORDER], int k) {
    float r = w[0];
#pragma unroll
    for (int t = 1; t < ORDER; ++t) r = k == t ? w[t] : r;
    return r;
}

// Field interpolation (Ewald_long_range_forces_order_*_cu): a team of T
// lanes per charged atom splits the ORDER^2 stencil rows (the field's three
// components are consecutive real meshes, as the batched inverse FFT writes
// them), and the team sums by butterfly shuffles before one fixed-point add
// per component.
template <int ORDER, int T>
__global__ void __launch_bounds__(256) k_gather(int n, const double4 *__restrict__ pos,
                                                const float *__restrict__ q, MeshGeo m, float scale,
                                                const float *__restrict__ E,
                                                unsigned long long *__restrict__ force) {
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gt / T;
    const int tl = threadIdx.x & (T - 1);
    if (i >= n) return;  // whole teams leave together
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
    const size_t ng = (size_t)m.sz[0] * m.sz[1] * m.sz[2];
    float ex = 0.f, ey = 0.f, ez = 0.f;
    for (int r = tl; r < ORDER * ORDER; r += T) {
        const int ix = r / ORDER, iy = r - ix * ORDER;
        const int px = wrapi(g[0] + ix - H, m.sz[0]);
        const int py = wrapi(g[1] + iy - H, m.sz[1]);
        const float wxy = pick<ORDER>(wx, ix) * pick<ORDER>(wy, iy);
        const size_t rowb = ((size_t)px * m.sz[1] + py) * m.sz[2];
        float e0[ORDER], e1[ORDER], e2[ORDER];
#pragma unroll
        for (int iz = 0; iz < ORDER; ++iz) {
            const size_t gidx = rowb + wrapi(g[2] + iz - H, m.sz[2]);
            e0[iz] = E[gidx];
            e1[iz] = E[ng + gidx];
            e2[iz] = E[2 * ng + gidx];
        }
#pragma unroll
        for (int iz = 0; iz < ORDER; ++iz) {
            const float w = wxy * wz[iz];
            ex += w * e0[iz];
            ey += w * e1[iz];
            ez += w * e2[iz];
        }
    }
    ex = team_sum<T>(ex);
    ey = team_sum<T>(ey);
    ez = team_sum<T>(ez);
    if (tl == 0) {
        // F = Qunit q sum W (-E / V)  (Ewald_long_range_forces_order_5_cu)
        const float s = scale * qi;
        fx_add(&force[3 * i + 0], s * ex);
        fx_add(&force[3 * i + 1], s * ey);
        fx_add(&force[3 * i + 2], s * ez);
    }
}

MeshGeo make_mesh(const Ctx &c) {
    MeshGeo m;
    for (int a = 0; a < 3; ++a) {
        m.sz[a] = c.pppm.sz[a];
        m.lo[a] = c.geo.lo[a];
        m.L[a] = c.geo.L[a];
        m.invL[a] = c.geo.invL[a];
        m.h[a] = c.geo.L[a] / c.pppm.sz[a];
        m.periodic[a] = c.geo.periodic[a];
    }
    return m;
}

const double kAmpTable[7][7] = {
    {2.0 / 3.0, 0, 0, 0, 0, 0, 0},
    {1.0 / 50.0, 5.0 / 294.0, 0, 0, 0, 0, 0},
    {1.0 / 588.0, 7.0 / 1440.0, 21.0 / 3872.0, 0, 0, 0, 0},
    {1.0 / 4320.0, 3.0 / 1936.0, 7601.0 / 2271360.0, 143.0 / 28800.0, 0, 0, 0},
    {1.0 / 23232.0, 7601.0 / 12628160.0, 143.0 / 69120.0, 517231.0 / 106536960.0,
     106640677.0 / 11737571328.0, 0, 0},
    {691.0 / 68140800.0, 13.0 / 57600.0, 47021.0 / 35512320.0, 9694607.0 / 2095994880.0,
     733191589.0 / 59609088000.0, 326190917.0 / 11700633600.0, 0},
    {1.0 / 345600.0, 3617.0 / 35512320.0, 745739.0 / 838397952.0, 56399353.0 / 12773376000.0,
     25091609.0 / 1560084480.0, 1755948832039.0 / 36229939200000.0, 48887769399.0 / 37838389248.0}};

double deltaf_k(double a, int natoms, double q2, const double *L, const int *sz, int order) {
    double h[3] = {L[0] / sz[0], L[1] / sz[1], L[2] / sz[2]};
    double sum[3] = {0, 0, 0};
    for (int mm = 0; mm < order; ++mm) {
        double amp = kAmpTable[order - 1][mm];
        for (int d = 0; d < 3; ++d) sum[d] += amp * std::pow(h[d] * a, 2 * mm);
    }
    double acc = 0.0;
    for (int d = 0; d < 3; ++d)
        acc += 1.0 / (L[d] * L[d]) * std::pow(a * h[d], order) *
               std::sqrt(a * L[d] / natoms * std::sqrt(2.0 * M_PI) * sum[d]);
    return q2 / 3.0 * acc;
}

double deltaf_real(double a, int natoms, double q2, const double *L, double rcut) {
    return 2 * q2 / std::sqrt(natoms * rcut * L[0] * L[1] * L[2]) * std::exp(-a * a * rcut * rcut);
}

// Tile shape for the LDS spread: z-chunks of zs cells so that the tile stays
// within 48 KB of LDS.  Returns false (plain spread) when no cell list is
// valid, the box is not periodic, or even one cell does not fit.
bool tiled_geometry(const Ctx &c, TileGeo &tg) {
    if (!c.listValid || c.ncell == 0 || !c.colXY || c.mdxDisableTiledSpread) return false;
    for (int a = 0; a < 3; ++a)
        if (!c.geo.periodic[a]) return false;
    const int order = c.pppm.order;
    tg.ncz = c.nc[2];
    tg.margin = 0.5 * c.skin + 1e-6 * c.geo.L[0];
    for (int a = 0; a < 3; ++a) tg.cellw[a] = c.geo.L[a] / c.nc[a];
    auto extent = [&](int a, int ncells) {
        double h = c.geo.L[a] / c.pppm.sz[a];
        return (int)std::floor((ncells * tg.cellw[a] + 2.0 * tg.margin) / h) + order + 2;
    };
    tg.ext[0] = extent(0, 1);
    tg.ext[1] = extent(1, 1);
    if (tg.ext[0] >= c.pppm.sz[0] || tg.ext[1] >= c.pppm.sz[1]) return false;
    const int budget = 48 * 1024 / 8;
    int zs = 0;
    for (int z = 1; z <= tg.ncz; ++z) {
        int ez = extent(2, z);
        if (ez >= c.pppm.sz[2] || tg.ext[0] * tg.ext[1] * ez > budget) break;
        zs = z;
    }
    if (zs == 0) return false;
    tg.zs = zs;
    tg.nzc = (tg.ncz + zs - 1) / zs;
    tg.ext[2] = extent(2, zs);
    return true;
}

}  // namespace

// Secant solve of DeltaF_k(alpha) = DeltaF_real(alpha)
// (find_optimal_parameters, FixChargeEwald.cu:935-967), FP64 on the host.
double pppm_solve_alpha(int natoms, double q2, const double *L, const int *sz, int order, double rcut,
                        double *err) {
    double xa = 0.0, xb = 4.79853 / rcut;
    double ya = deltaf_k(xa, natoms, q2, L, sz, order) - deltaf_real(xa, natoms, q2, L, rcut);
    double yb = deltaf_k(xb, natoms, q2, L, sz, order) - deltaf_real(xb, natoms, q2, L, rcut);
    double tol = 1e-5;
    int it = 0;
    while ((std::fabs(yb) / deltaf_real(xb, natoms, q2, L, rcut) > tol) && it < 100) {
        double kinv = (xb - xa) / (yb - ya);
        ya = yb;
        xa = xb;
        xb = xa - ya * kinv;
        yb = deltaf_k(xb, natoms, q2, L, sz, order) - deltaf_real(xb, natoms, q2, L, rcut);
        ++it;
    }
    if (err) *err = deltaf_k(xb, natoms, q2, L, sz, order) + deltaf_real(xb, natoms, q2, L, rcut);
    return xb;
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
        P.plansMade = true;
    }
    bool same = true;
    for (int a = 0; a < 3; ++a) same = same && (P.boxAtGreen[a] == c.geo.L[a]);
    if (same && P.alpha > 0) return;
    if (P.alphaUser > 0) {
        P.alpha = P.alphaUser;
        P.err = 0.0;
    } else {
        P.alpha = pppm_solve_alpha(c.n, P.totalQ2, c.geo.L, sz, P.order, P.rcut, &P.err);
    }
    c.alpha = (float)P.alpha;
    MeshGeo m = make_mesh(c);
    double hm = std::pow(m.h[0] * m.h[1] * m.h[2], 1.0 / 3.0);
    int lim = (int)(P.alpha * hm / 3.14159 * (std::sqrt(-std::log(10E-7)))) + 1;
    hipLaunchKernelGGL(k_green, dim3(cdiv(nh, 256)), dim3(256), 0, c.stream, m, szh, P.alpha, P.order,
                       lim, P.green);
    for (int a = 0; a < 3; ++a) P.boxAtGreen[a] = c.geo.L[a];
}

void pppm_forces(Ctx &c, bool energy, bool virial) {
    PppmState &P = c.pppm;
    pppm_prepare(c);
    const int *sz = P.sz;
    const int szh = sz[2] / 2 + 1;
    const int ng = sz[0] * sz[1] * sz[2];
    const int nh = sz[0] * sz[1] * szh;
    MeshGeo m = make_mesh(c);
    const float qunit = (float)std::sqrt(P.qqr);
    const int threads = 256;
    const int cur = c.cur;
    {
    ProfScope ps(c, PK_SPREAD);
    TileGeo tg;
    if (tiled_geometry(c, tg)) {
        const int nblk = c.nc[0] * c.nc[1] * tg.nzc;
        const size_t shm = sizeof(unsigned long long) * tg.ext[0] * tg.ext[1] * tg.ext[2];
        switch (P.order) {
            case 1: hipLaunchKernelGGL(k_spread_tiled<1>, dim3(nblk), dim3(threads), shm, c.stream, c.pos[cur], c.q[cur], m, qunit, c.cellStart, c.colXY, tg, (unsigned long long *)P.mesh); break;
            case 3: hipLaunchKernelGGL(k_spread_tiled<3>, dim3(nblk), dim3(threads), shm, c.stream, c.pos[cur], c.q[cur], m, qunit, c.cellStart, c.colXY, tg, (unsigned long long *)P.mesh); break;
            case 5: hipLaunchKernelGGL(k_spread_tiled<5>, dim3(nblk), dim3(threads), shm, c.stream, c.pos[cur], c.q[cur], m, qunit, c.cellStart, c.colXY, tg, (unsigned long long *)P.mesh); break;
            default: hipLaunchKernelGGL(k_spread_tiled<7>, dim3(nblk), dim3(threads), shm, c.stream, c.pos[cur], c.q[cur], m, qunit, c.cellStart, c.colXY, tg, (unsigned long long *)P.mesh); break;
        }
    } else
    switch (P.order) {
        case 1: hipLaunchKernelGGL(k_spread<1>, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.pos[cur], c.q[cur], m, qunit, (unsigned long long *)P.mesh); break;
        case 3: hipLaunchKernelGGL(k_spread<3>, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.pos[cur], c.q[cur], m, qunit, (unsigned long long *)P.mesh); break;
        case 5: hipLaunchKernelGGL(k_spread<5>, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.pos[cur], c.q[cur], m, qunit, (unsigned long long *)P.mesh); break;
        default: hipLaunchKernelGGL(k_spread<7>, dim3(cdiv(c.n, threads)), dim3(threads), 0, c.stream, c.n, c.pos[cur], c.q[cur], m, qunit, (unsigned long long *)P.mesh); break;
    }
    hipLaunchKernelGGL(k_mesh_to_real, dim3(cdiv(ng, threads)), dim3(threads), 0, c.stream, ng, P.mesh, P.real);
    }
    {
    ProfScope ps(c, PK_FFT);
    MDX_REQUIRE(hipfftExecR2C(P.planR2C, P.real, (hipfftComplex *)P.rhoHat) == HIPFFT_SUCCESS, 2,
                "hipfftExecR2C failed");
    }
    if (energy || virial) {
        double V = c.geo.L[0] * c.geo.L[1] * c.geo.L[2];
        hipLaunchKernelGGL(k_kspace_ev, dim3(cdiv(nh, threads)), dim3(threads), 0, c.stream, m, szh,
                           P.alpha, 1.0 / V, P.green, P.rhoHat, c.acc);
    }
    {
        ProfScope ps(c, PK_EFIELD);
        hipLaunchKernelGGL(k_efield, dim3(cdiv(nh, threads)), dim3(threads), 0, c.stream, m, szh, P.green,
                           P.rhoHat, P.eHat);
    }
    {
        ProfScope ps(c, PK_FFT);
        MDX_REQUIRE(hipfftExecC2R(P.planC2R, (hipfftComplex *)P.eHat, P.eReal) == HIPFFT_SUCCESS, 2,
                    "hipfftExecC2R failed");
    }
    const double V = c.geo.L[0] * c.geo.L[1] * c.geo.L[2];
    const float scale = (float)(-std::sqrt(P.qqr) / V);
    // a built list puts the charged atoms first: only they take the gather
    const int na = c.listValid ? c.nCharged : c.n;
    ProfScope ps(c, PK_GATHER);
    if (na == 0) return;
    const dim3 grid(cdiv((long long)na * P.gatherTeam, threads));
#define MDX_GATHER(O, T)                                                                                   \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gather<O, T>), grid, dim3(threads), 0, c.stream, na, c.pos[cur], \
                       c.q[cur], m, scale, P.eReal, c.force)
    const bool wide = P.gatherTeam == 16;
    switch (P.order) {
        case 1: if (wide) MDX_GATHER(1, 16); else MDX_GATHER(1, 8); break;
        case 3: if (wide) MDX_GATHER(3, 16); else MDX_GATHER(3, 8); break;
        case 5: if (wide) MDX_GATHER(5, 16); else MDX_GATHER(5, 8); break;
        default: if (wide) MDX_GATHER(7, 16); else MDX_GATHER(7, 8); break;
    }
#undef MDX_GATHER
}

}  // namespace mdx
