// Ring-polymer (PIMD) drift: the free ring polymer of every atom propagated
// exactly in normal modes over one time step, as nve_xPIMD_cu / preForcePIMD_cu
// do (src/Integrators/IntegratorVerlet.cu:63-304, 342-593): positions and
// velocities of the P beads go to normal-mode coordinates with the orthonormal
// matrix C (k = 0: 1/sqrt P; 1 <= k < P/2: sqrt(2/P) cos(2 pi k n / P);
// k = P/2: (-1)^n / sqrt P; P/2 < k < P: sqrt(2/P) sin(2 pi k n / P), n the
// bead number from 1), mode 0 drifts, mode k rotates at
// w_k = 2 w_P sin(k pi / P), and the beads come back with C^T.  Beads are
// first put in the periodic image of bead 0, as the reference does.
//
// Here each bead is a whole system in its own context (its own neighbor
// list and cell order), so atom `id` of bead n sits at idToIdx_n[id].
// One team of P lanes per atom: lane n loads bead n, computes mode n, and
// after an LDS exchange bead n again.  Massless sites (M) are left alone;
// the force evaluation places them.
//
// Ring-polymer contraction of the mesh part (FixChargeEwald.cu:1255-1266,
// 1395-1405): the charges are spread from the centroid of each atom's ring
// (computeCentroid, FixChargeEwald.cu:22-50: bead 0 plus the mean of the
// other beads' minimum-image offsets from it, wrapped into the box), the
// field is interpolated at the centroid, and the whole force q E goes onto
// every bead of the atom.  k_ring_centroid and k_ring_apply are the two ends
// of that; the mesh work in between is the single-system P3M chain.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kRingThreads = 256;

__device__ __forceinline__ void ring_body(int blk, int nblk, int natoms, int P, void *const *__restrict__ tab,
                                                       const double *__restrict__ C,
                                                       const double *__restrict__ cw,
                                                       const BoxDev *__restrict__ bx) {
    extern __shared__ double sh[];  // [teams][P][6]: x, v of every bead
    const int teams = kRingThreads / P;
    const int team = threadIdx.x / P, n = threadIdx.x % P;
    const int id = blk * teams + team;
    double *me = sh + (size_t)team * P * 6;
    const Geometry g = bx->geo;
    double4 *pos = (double4 *)tab[kRingTab * n];
    double4 *vel = (double4 *)tab[kRingTab * n + 1];
    const int *idToIdx = (const int *)tab[kRingTab * n + 2];
    const bool live = id < natoms;
    int idx = 0;
    double4 x = make_double4(0, 0, 0, 0), v = make_double4(0, 0, 0, 0);
    if (live) {
        idx = idToIdx[id];
        x = pos[idx];
        v = vel[idx];
    }
    const bool moves = live && v.w != 0.0;  // the same for every bead of the atom
    me[n * 6 + 0] = x.x;
    me[n * 6 + 1] = x.y;
    me[n * 6 + 2] = x.z;
    __syncthreads();
    // bead n in bead 0's image
    double xn[3] = {x.x, x.y, x.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double x0 = me[d];
        double dd = xn[d] - x0;
        if (g.periodic[d]) dd -= g.L[d] * rint(dd * g.invL[d]);
        xn[d] = x0 + dd;
    }
    __syncthreads();
    me[n * 6 + 0] = xn[0];
    me[n * 6 + 1] = xn[1];
    me[n * 6 + 2] = xn[2];
    me[n * 6 + 3] = v.x;
    me[n * 6 + 4] = v.y;
    me[n * 6 + 5] = v.z;
    __syncthreads();
    // mode k = n
    const int k = n;
    double xt[3] = {0, 0, 0}, vt[3] = {0, 0, 0};
    for (int m = 0; m < P; ++m) {
        const double c = C[m * P + k];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xt[d] += c * me[m * 6 + d];
            vt[d] += c * me[m * 6 + 3 + d];
        }
    }
    const double cs = cw[3 * k], sw = cw[3 * k + 1], ws = cw[3 * k + 2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double x1 = xt[d] * cs + vt[d] * sw;
        const double v1 = vt[d] * cs - xt[d] * ws;
        xt[d] = x1;
        vt[d] = v1;
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        me[k * 6 + d] = xt[d];
        me[k * 6 + 3 + d] = vt[d];
    }
    __syncthreads();
    double xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0};
    for (int kk = 0; kk < P; ++kk) {
        const double c = C[n * P + kk];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xb[d] += c * me[kk * 6 + d];
            vb[d] += c * me[kk * 6 + 3 + d];
        }
    }
    if (moves) {
        pos[idx] = make_double4(xb[0], xb[1], xb[2], x.w);
        vel[idx] = make_double4(vb[0], vb[1], vb[2], v.w);
    }
}

// centroid of every atom's ring, by atom id
// (FP64; the reference's is in its working precision).  No FMA contraction,
// so the centroids are those of the same operations in plain FP64 arithmetic.
__device__ __forceinline__ void ring_centroid_body(int blk, int nblk, const int *__restrict__ gate, int natoms, int P,
                                                   void *const *__restrict__ tab, const BoxDev *__restrict__ bx,
                                                   double4 *__restrict__ cen) {
#pragma clang fp contract(off)
    if (gate && *gate == 0) return;
    const int id = blk * blockDim.x + threadIdx.x;
    if (id >= natoms) return;
    const Geometry g = bx->geo;
    const double4 x0 = ((const double4 *)tab[0])[((const int *)tab[2])[id]];
    double s[3] = {0.0, 0.0, 0.0};
    for (int n = 1; n < P; ++n) {
        const double4 xn = ((const double4 *)tab[kRingTab * n])[((const int *)tab[kRingTab * n + 2])[id]];
        double dd[3] = {xn.x - x0.x, xn.y - x0.y, xn.z - x0.z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (g.periodic[d]) dd[d] -= g.L[d] * rint(dd[d] * g.invL[d]);
            s[d] += dd[d];
        }
    }
    double c[3] = {x0.x + s[0] / P, x0.y + s[1] / P, x0.z + s[2] / P};
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (g.periodic[d]) c[d] -= g.L[d] * floor((c[d] - g.lo[d]) * g.invL[d]);
    cen[id] = make_double4(c[0], c[1], c[2], 0.0);
}

// the centroid force (fixed point, by id) onto every bead of the atom; the
// by-id buffer is cleared for the next evaluation
__device__ __forceinline__ void ring_apply_body(int blk, int nblk, int natoms, int P, void *const *__restrict__ tab,
                                                    unsigned long long *__restrict__ cf) {
    const int id = blk * blockDim.x + threadIdx.x;
    if (id >= natoms) return;
    const unsigned long long f0 = cf[3 * id], f1 = cf[3 * id + 1], f2 = cf[3 * id + 2];
    if ((f0 | f1 | f2) == 0ull) return;
    cf[3 * id] = 0ull;
    cf[3 * id + 1] = 0ull;
    cf[3 * id + 2] = 0ull;
    for (int n = 0; n < P; ++n) {
        unsigned long long *f = (unsigned long long *)tab[kRingTab * n + 3];
        const int i = ((const int *)tab[kRingTab * n + 2])[id];
        atomicAdd(&f[3 * i + 0], f0);
        atomicAdd(&f[3 * i + 1], f1);
        atomicAdd(&f[3 * i + 2], f2);
    }
}

}  // namespace

void ring_centroids(const RingSpec &r, Ctx &c0, const int *gate) {
    launch<ring_centroid_body, 256>(c0, cdiv(r.natoms, 256), 0, gate, r.natoms, r.P, r.tab, c0.box, r.cen);
}

void ring_apply(const RingSpec &r, Ctx &c0) {
    launch<ring_apply_body, 256>(c0, cdiv(r.natoms, 256), 0, r.natoms, r.P, r.tab, (unsigned long long *)r.cf);
}

void ring_drift(const RingSpec &r, Ctx &c0) {
    const int teams = kRingThreads / r.P;
    const size_t shm = sizeof(double) * 6 * kRingThreads;
    launch<ring_body, kRingThreads>(c0, cdiv(r.natoms, teams), shm, r.natoms, r.P, r.tab, r.C, r.cw, c0.box);
}

}  // namespace mdx
