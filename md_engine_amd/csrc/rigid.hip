// Rigid water in FP64: SETTLE positions, RATTLE velocities, M-site placement
// and force distribution.  Replaces FixRigid's settlePositions
// (src/Fixes/FixRigid.cu:850-1136), settleVelocities (658-849), setMSite
// (560-599) and distributeMSite (475-558) as sequenced by postNVE_X and
// preStepFinal (1234-1279, 1916-1984).
//
// Design (gfx950): one thread per molecule does the whole half step of its
// four sites, so the velocity-Verlet kick, the drift, the constraint solve and
// the M placement are one pass over the molecule's data (no separate
// force-distribution or velocity-correction kernels).  The M-site force is
// distributed inside the kernel (F_O += gamma F_M, F_H += (1 - gamma)/2 F_M)
// from the raw fixed-point force buffer, which is left untouched.
// SETTLE follows Miyamoto & Kollman, J. Comput. Chem. 13, 952 (1992); the
// velocity constraint is the exact 3x3 solve of the three bond conditions.
#include "mdx_internal.h"

namespace mdx {

namespace {

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mk(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 unit(d3 a) { return (1.0 / sqrt(dot(a, a))) * a; }
__device__ __forceinline__ d3 xyz(double4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ double fxd(unsigned long long v) { return (double)(long long)v * kFxInv; }
__device__ __forceinline__ d3 force_of(const unsigned long long *f, int i) {
    return {fxd(f[3 * i]), fxd(f[3 * i + 1]), fxd(f[3 * i + 2])};
}
__device__ __forceinline__ d3 mi(d3 d, const Geometry &g) {
    if (g.periodic[0]) d.x -= g.L[0] * rint(d.x * g.invL[0]);
    if (g.periodic[1]) d.y -= g.L[1] * rint(d.y * g.invL[1]);
    if (g.periodic[2]) d.z -= g.L[2] * rint(d.z * g.invL[2]);
    return d;
}

struct RigidArgs {
    int nmol;
    const int4 *atoms;
    double mO, mH, dOH, dHH, gamma;
    double dt, hdtf;
    Geometry g;          // taken from bx in the kernels
    const BoxDev *bx;
};

// SETTLE: old positions relative to the old oxygen (A0 = 0, B0, C0) and
// unconstrained new positions A1, B1, C1 -> constrained positions in place.
__device__ void settle(d3 B0, d3 C0, d3 &A1, d3 &B1, d3 &C1, const RigidArgs &r) {
    const double mT = r.mO + 2.0 * r.mH, wO = r.mO / mT, wH = r.mH / mT;
    const double rc = 0.5 * r.dHH;
    const double dmid = sqrt(r.dOH * r.dOH - rc * rc);
    const double ra = 2.0 * wH * dmid, rb = dmid - ra;
    const d3 com = wO * A1 + wH * (B1 + C1);
    const d3 a1 = A1 - com, b1 = B1 - com, c1 = C1 - com;
    const d3 n0 = cross(B0, C0);
    const d3 xa = cross(a1, n0);
    const d3 ya = cross(n0, xa);
    const d3 ez = unit(n0), ex = unit(xa), ey = unit(ya);
    const double xb0 = dot(B0, ex), yb0 = dot(B0, ey), xc0 = dot(C0, ex), yc0 = dot(C0, ey);
    const double za1 = dot(a1, ez);
    const double xb1 = dot(b1, ex), yb1 = dot(b1, ey), zb1 = dot(b1, ez);
    const double xc1 = dot(c1, ex), yc1 = dot(c1, ey), zc1 = dot(c1, ez);
    const double sinphi = za1 / ra, cosphi = sqrt(1.0 - sinphi * sinphi);
    const double sinpsi = (zb1 - zc1) / (2.0 * rc * cosphi), cospsi = sqrt(1.0 - sinpsi * sinpsi);
    const double ya2 = ra * cosphi, xb2 = -rc * cospsi;
    const double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
    const double yb2 = t1 - t2, yc2 = t1 + t2;
    const double alpha = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
    const double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
    const double gam = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
    const double ab2 = alpha * alpha + beta * beta;
    const double sinth = (alpha * gam - beta * sqrt(ab2 - gam * gam)) / ab2;
    const double costh = sqrt(1.0 - sinth * sinth);
    A1 = com + (-ya2 * sinth) * ex + (ya2 * costh) * ey + za1 * ez;
    B1 = com + (xb2 * costh - yb2 * sinth) * ex + (xb2 * sinth + yb2 * costh) * ey + zb1 * ez;
    C1 = com + (-xb2 * costh - yc2 * sinth) * ex + (-xb2 * sinth + yc2 * costh) * ey + zc1 * ez;
}

__device__ __forceinline__ double det3(const double M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
           M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

// Velocity constraint: impulses along the three bonds making every
// (v_i - v_j).(r_i - r_j) zero; rO/rH1/rH2 relative positions.
__device__ void rattle(d3 rO, d3 rH1, d3 rH2, d3 &vO, d3 &vH1, d3 &vH2, double mO, double mH) {
    const d3 r[3] = {rO, rH1, rH2};
    d3 v[3] = {vO, vH1, vH2};
    const double im[3] = {1.0 / mO, 1.0 / mH, 1.0 / mH};
    const int bi[3] = {0, 0, 1}, bj[3] = {1, 2, 2};
    d3 e[3];
    double M[3][3], rhs[3];
    for (int k = 0; k < 3; ++k) e[k] = r[bi[k]] - r[bj[k]];
    for (int k = 0; k < 3; ++k) {
        rhs[k] = -dot(v[bi[k]] - v[bj[k]], e[k]);
        for (int l = 0; l < 3; ++l) {
            const double si = (bi[l] == bi[k]) ? 1.0 : (bj[l] == bi[k] ? -1.0 : 0.0);
            const double sj = (bi[l] == bj[k]) ? 1.0 : (bj[l] == bj[k] ? -1.0 : 0.0);
            M[k][l] = dot(e[k], e[l]) * (si * im[bi[k]] - sj * im[bj[k]]);
        }
    }
    const double det = det3(M);
    double lam[3];
    for (int c = 0; c < 3; ++c) {
        double A[3][3];
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) A[k][l] = (l == c) ? rhs[k] : M[k][l];
        lam[c] = det3(A) / det;
    }
    for (int k = 0; k < 3; ++k) {
        v[bi[k]] = v[bi[k]] + (lam[k] * im[bi[k]]) * e[k];
        v[bj[k]] = v[bj[k]] - (lam[k] * im[bj[k]]) * e[k];
    }
    vO = v[0];
    vH1 = v[1];
    vH2 = v[2];
}

__device__ __forceinline__ void dist_forces(const unsigned long long *f, int4 a, double gamma, d3 &fO,
                                            d3 &fH1, d3 &fH2) {
    fO = force_of(f, a.x);
    fH1 = force_of(f, a.y);
    fH2 = force_of(f, a.z);
    if (a.w >= 0) {
        const d3 fM = force_of(f, a.w);
        fO = fO + gamma * fM;
        const double h = 0.5 * (1.0 - gamma);
        fH1 = fH1 + h * fM;
        fH2 = fH2 + h * fM;
    }
}

// squared minimum-image displacement, the same expressions as the list's
// skin check (dist2 of the neighbor-list code), so the maxima agree bit for bit
__device__ __forceinline__ double disp2(const double4 &a, const double4 &b, const Geometry &g) {
    double d[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (g.periodic[k]) {
            double img = rint(d[k] * g.invL[k]);
            d[k] = d[k] - img * g.L[k];
        }
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// part != nullptr: also the skin check's per-workgroup maxima (k_check_part)
// of the sites' displacement since the last build and of the displacement
// plus 2 dt |v|, from the positions and velocities just written; valid when
// every site belongs to a rigid molecule.
// With xq set (captured all-rigid steps) the kick also does the force
// evaluation's position refresh for its sites - FP32 positions, the P3M mesh
// records of the charged ones, cleared fixed-point force rows (read above)
// and, in workgroup 0, the cleared energy / virial accumulators - so no
// separate refresh launch is needed (a rebuild's permutation rewrites xq and
// the records in the new order).
struct RefreshArgs {
    float4 *xq;
    float4 *rec;
    const float *q;
    unsigned long long *force;
    unsigned long long *acc;
    int nacc;
};

__device__ __forceinline__ void rigid_vv1_body(int blk, int nblk, RigidArgs r, double4 *__restrict__ pos, double4 *__restrict__ vel,
                                                   const unsigned long long *__restrict__ f,
                                                   const double4 *__restrict__ last, double dt2,
                                                   float *__restrict__ part, DecideArgs da, RefreshArgs ra) {
    r.g = r.bx->geo;
    const int m = blk * blockDim.x + threadIdx.x;
    float d2 = 0.f, p2 = 0.f;
    if (ra.xq && blk == 0)
        for (int k = threadIdx.x; k < ra.nacc; k += blockDim.x) ra.acc[k] = 0ull;
    if (m < r.nmol) {
    const int4 a = r.atoms[m];
    d3 fO, fH1, fH2;
    dist_forces(f, a, r.gamma, fO, fH1, fH2);
    const double4 pO = pos[a.x], pH1 = pos[a.y], pH2 = pos[a.z];
    const double4 wO = vel[a.x], wH1 = vel[a.y], wH2 = vel[a.z];
    d3 vO = xyz(wO) + (r.hdtf / r.mO) * fO;
    d3 vH1 = xyz(wH1) + (r.hdtf / r.mH) * fH1;
    d3 vH2 = xyz(wH2) + (r.hdtf / r.mH) * fH2;
    const d3 xO = xyz(pO);
    const d3 B0 = mi(xyz(pH1) - xO, r.g), C0 = mi(xyz(pH2) - xO, r.g);
    const d3 A1 = r.dt * vO, B1 = B0 + r.dt * vH1, C1 = C0 + r.dt * vH2;
    d3 A3 = A1, B3 = B1, C3 = C1;
    settle(B0, C0, A3, B3, C3, r);
    const double idt = 1.0 / r.dt;
    vO = vO + idt * (A3 - A1);
    vH1 = vH1 + idt * (B3 - B1);
    vH2 = vH2 + idt * (C3 - C1);
    const d3 nO = xO + A3;
    const d3 nH1 = nO + (B3 - A3), nH2 = nO + (C3 - A3);
    pos[a.x] = make_double4(nO.x, nO.y, nO.z, pO.w);
    pos[a.y] = make_double4(nH1.x, nH1.y, nH1.z, pH1.w);
    pos[a.z] = make_double4(nH2.x, nH2.y, nH2.z, pH2.w);
    vel[a.x] = make_double4(vO.x, vO.y, vO.z, wO.w);
    vel[a.y] = make_double4(vH1.x, vH1.y, vH1.z, wH1.w);
    vel[a.z] = make_double4(vH2.x, vH2.y, vH2.z, wH2.w);
    double4 sp[4], sv[4];
    sp[0] = make_double4(nO.x, nO.y, nO.z, pO.w);
    sp[1] = make_double4(nH1.x, nH1.y, nH1.z, pH1.w);
    sp[2] = make_double4(nH2.x, nH2.y, nH2.z, pH2.w);
    sv[0] = make_double4(vO.x, vO.y, vO.z, wO.w);
    sv[1] = make_double4(vH1.x, vH1.y, vH1.z, wH1.w);
    sv[2] = make_double4(vH2.x, vH2.y, vH2.z, wH2.w);
    if (a.w >= 0) {
        const d3 nM = nO + (0.5 * (1.0 - r.gamma)) * ((B3 - A3) + (C3 - A3));
        const double4 pM = pos[a.w];
        sp[3] = make_double4(nM.x, nM.y, nM.z, pM.w);
        sv[3] = make_double4(0.0, 0.0, 0.0, vel[a.w].w);
        pos[a.w] = sp[3];
        vel[a.w] = sv[3];
    }
    if (ra.xq) {
        const int site[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (site[s] < 0) continue;
            const double4 p = sp[s];
            const float qs = ra.q[site[s]];
            ra.xq[site[s]] = make_float4((float)(p.x - r.g.lo[0]), (float)(p.y - r.g.lo[1]), (float)(p.z - r.g.lo[2]), qs);
            if (ra.rec && qs != 0.0f) ra.rec[site[s]] = mesh_record(p, r.bx->mesh);
            ra.force[3 * (size_t)site[s] + 0] = 0ull;
            ra.force[3 * (size_t)site[s] + 1] = 0ull;
            ra.force[3 * (size_t)site[s] + 2] = 0ull;
        }
    }
    if (part) {
        const int site[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (site[s] < 0) continue;
            const double dd = disp2(sp[s], last[site[s]], r.g);
            const double4 v = sv[s];
            const double p = sqrt(dd) + dt2 * sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
            d2 = fmaxf(d2, (float)dd);
            p2 = fmaxf(p2, (float)(p * p));
        }
    }
    }
    if (part) {
        __shared__ float red[2][2];
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
            part[2 * blk] = fmaxf(red[0][0], red[0][1]);
            part[2 * blk + 1] = fmaxf(red[1][0], red[1][1]);
        }
        // the step's decisions in the last workgroup (one launch fewer)
        if (da.ticket && last_block(da.ticket, nblk)) decide_wg(nblk, da);
    }
}

// kick = false: only the velocity constraint and the M placement (start of a run)
// SUMS: on the steps their gates allow, also (doVir) the molecular virial
// correction and centre-of-mass M|V|^2 of the final velocities (as
// k_mol_virial) into acc[ACC_MOL], the barostat's pressure without a pass of
// its own, and (doTh) the sum of m v^2 over the sites (as k_mv2) into
// thermo[0] for the velocity rescale.
// the barostat step for the last workgroup of the rigid second half
// (ticket = nullptr: none)
struct BaroFuse {
    int *ticket;
    const int *gate;
    BoxDev *bx;
    BoxParams bp;
    BaroArgs ba;
};

template <bool SUMS>
__device__ __forceinline__ void rigid_vv2_body(int blk, int nblk, RigidArgs r, double4 *__restrict__ pos, double4 *__restrict__ vel,
                                                   const unsigned long long *__restrict__ f, int kick,
                                                   unsigned long long *__restrict__ acc,
                                                   const int *__restrict__ vgate, int doVir,
                                                   unsigned long long *__restrict__ thermo,
                                                   const int *__restrict__ tgate, int doTh, BaroFuse bf) {
    r.g = r.bx->geo;
    const int m = blk * blockDim.x + threadIdx.x;
    // uniform over the grid
    const bool vir = SUMS && doVir && (!vgate || *vgate);
    const bool th = SUMS && doTh && (!tgate || *tgate);
    double corr = 0.0, mv2 = 0.0, ke = 0.0;
    if (m < r.nmol) {
    const int4 a = r.atoms[m];
    const double4 wO = vel[a.x], wH1 = vel[a.y], wH2 = vel[a.z];
    d3 vO = xyz(wO), vH1 = xyz(wH1), vH2 = xyz(wH2);
    if (kick) {
        d3 fO, fH1, fH2;
        dist_forces(f, a, r.gamma, fO, fH1, fH2);
        vO = vO + (r.hdtf / r.mO) * fO;
        vH1 = vH1 + (r.hdtf / r.mH) * fH1;
        vH2 = vH2 + (r.hdtf / r.mH) * fH2;
    }
    const d3 xO = xyz(pos[a.x]);
    const d3 rH1 = mi(xyz(pos[a.y]) - xO, r.g), rH2 = mi(xyz(pos[a.z]) - xO, r.g);
    rattle(mk(0, 0, 0), rH1, rH2, vO, vH1, vH2, r.mO, r.mH);
    vel[a.x] = make_double4(vO.x, vO.y, vO.z, wO.w);
    vel[a.y] = make_double4(vH1.x, vH1.y, vH1.z, wH1.w);
    vel[a.z] = make_double4(vH2.x, vH2.y, vH2.z, wH2.w);
    if (!kick && a.w >= 0) {
        const d3 nM = xO + (0.5 * (1.0 - r.gamma)) * (rH1 + rH2);
        const double4 pM = pos[a.w];
        pos[a.w] = make_double4(nM.x, nM.y, nM.z, pM.w);
        vel[a.w] = make_double4(0.0, 0.0, 0.0, vel[a.w].w);
    }
    if (vir) {
        const double mT = r.mO + 2.0 * r.mH;
        const d3 h1 = xO + rH1, h2 = xO + rH2;
        const d3 R = (1.0 / mT) * (r.mO * xO + r.mH * (h1 + h2));
        corr = dot(force_of(f, a.x), xO - R) + dot(force_of(f, a.y), h1 - R) + dot(force_of(f, a.z), h2 - R);
        if (a.w >= 0) corr += dot(force_of(f, a.w), xO + mi(xyz(pos[a.w]) - xO, r.g) - R);
        const d3 V = (1.0 / mT) * (r.mO * vO + r.mH * (vH1 + vH2));
        mv2 = mT * dot(V, V);
    }
    if (th) {
        if (wO.w > 0.0) ke += dot(vO, vO) / wO.w;
        if (wH1.w > 0.0) ke += dot(vH1, vH1) / wH1.w;
        if (wH2.w > 0.0) ke += dot(vH2, vH2) / wH2.w;
        if (a.w >= 0) {
            const double4 v = vel[a.w];
            if (v.w > 0.0) ke += (v.x * v.x + v.y * v.y + v.z * v.z) / v.w;
        }
    }
    }
    if (vir || th) {
        __shared__ double red[3][2];
        corr = team_sum_d<64>(corr);
        mv2 = team_sum_d<64>(mv2);
        ke = team_sum_d<64>(ke);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            red[0][w] = corr;
            red[1][w] = mv2;
            red[2][w] = ke;
        }
        __syncthreads();
        if (vir && threadIdx.x < 2) {
            const double x = red[threadIdx.x][0] + red[threadIdx.x][1];
            if (x != 0.0) fx_add(&acc_rep(acc)[ACC_MOL + threadIdx.x], x);
        }
        if (th && threadIdx.x == 2) fx_add(&thermo[0], red[2][0] + red[2][1]);
    }
    // the last workgroup sees every workgroup's sums: the barostat step
    if (bf.ticket && last_block(bf.ticket, nblk) && threadIdx.x < 64) baro_wave(bf.gate, bf.bx, bf.bp, bf.ba, acc);
}

// molecule -> current site indices, and the per-site membership flag that
// keeps the plain integrator off these sites
__device__ __forceinline__ void rigid_map_body(int blk, int nblk, const int *__restrict__ gate, int nmol, const int4 *__restrict__ ids, const int *__restrict__ idToIdx,
                            int4 *__restrict__ atoms, int *__restrict__ member) {
    if (gate && *gate == 0) return;
    const int m = blk * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int4 i = ids[m];
    const int4 a = make_int4(idToIdx[i.x], idToIdx[i.y], idToIdx[i.z], i.w >= 0 ? idToIdx[i.w] : -1);
    atoms[m] = a;
    member[a.x] = 1;
    member[a.y] = 1;
    member[a.z] = 1;
    if (a.w >= 0) member[a.w] = 1;
}

RigidArgs make_args(const Ctx &c) {
    RigidArgs r;
    r.nmol = c.rigid.nmol;
    r.atoms = c.rigid.atoms;
    r.mO = c.rigid.mO;
    r.mH = c.rigid.mH;
    r.dOH = c.rigid.dOH;
    r.dHH = c.rigid.dHH;
    r.gamma = c.rigid.gamma;
    r.dt = c.dt;
    r.hdtf = 0.5 * c.dt * c.ftmToV;
    r.g = c.geo;
    r.bx = c.box;
    return r;
}

// Molecular virial correction and centre-of-mass kinetic term of each rigid
// molecule: corr = sum_s F_s . (r_s - R) over its four sites (R the centre of
// mass, hydrogens and M taken in O's image), mv2 = M |V|^2.  The pressure of
// rigid molecules is (M|V|^2 summed + W_atomic - corr) / 3V: the atomic virial
// minus sum_i F_i . (r_i - R_mol(i)) is the virial of the centre-of-mass
// separations, with no constraint forces needed.
// One thread per rigid molecule (m < nmol) and per site outside rigid
// molecules (i < n, member[i] == 0): the molecule's correction and
// centre-of-mass M|V|^2, or the site's m v^2, summed per workgroup into
// acc[slot] (corr) and acc[slot + 1] (M|V|^2 + m v^2).
__device__ __forceinline__ void mol_virial_body(int blk, int nblk, RigidArgs r, int n, const int *__restrict__ member,
                                                    const double4 *__restrict__ pos,
                                                    const double4 *__restrict__ vel,
                                                    const unsigned long long *__restrict__ f,
                                                    unsigned long long *__restrict__ acc, int slot,
                                                    const int *__restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ double red[2][4];
    r.g = r.bx->geo;
    const int t = blk * blockDim.x + threadIdx.x;
    double corr = 0.0, mv2 = 0.0;
    if (t < r.nmol) {
        const int4 a = r.atoms[t];
        const double mT = r.mO + 2.0 * r.mH;
        const d3 o = xyz(pos[a.x]);
        const d3 h1 = o + mi(xyz(pos[a.y]) - o, r.g);
        const d3 h2 = o + mi(xyz(pos[a.z]) - o, r.g);
        const d3 R = (1.0 / mT) * (r.mO * o + r.mH * (h1 + h2));
        corr = dot(force_of(f, a.x), o - R) + dot(force_of(f, a.y), h1 - R) + dot(force_of(f, a.z), h2 - R);
        if (a.w >= 0) corr += dot(force_of(f, a.w), o + mi(xyz(pos[a.w]) - o, r.g) - R);
        const d3 V = (1.0 / mT) * (r.mO * xyz(vel[a.x]) + r.mH * (xyz(vel[a.y]) + xyz(vel[a.z])));
        mv2 = mT * dot(V, V);
    }
    if (t < n && !(member && member[t])) {
        const double4 v = vel[t];
        if (v.w > 0.0) mv2 += (v.x * v.x + v.y * v.y + v.z * v.z) / v.w;
    }
    corr = team_sum_d<64>(corr);
    mv2 = team_sum_d<64>(mv2);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = corr;
        red[1][w] = mv2;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double x = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) x += red[threadIdx.x][k];
        if (x != 0.0) fx_add(&acc_rep(acc)[slot + threadIdx.x], x);
    }
}

__device__ __forceinline__ void dilate_body(int blk, int nblk, DilateArgs d, const BoxDev *__restrict__ bx,
                                                const int *__restrict__ gate) {
    if (gate && *gate == 0) return;
    dilate_thread(blk * blockDim.x + threadIdx.x, d, bx);
}

}  // namespace

void molecular_virial(Ctx &c, int slot, const int *gate) {
    RigidArgs r = make_args(c);
    const int *member = nullptr;
    if (c.rigid.on) {
        member = c.rigid.member;
    } else {
        r.nmol = 0;
        r.atoms = nullptr;
    }
    launch<mol_virial_body, 256>(c, cdiv(std::max(r.nmol, c.n), 256), 0, r, c.n, member, c.pos[c.cur], c.vel[c.cur], c.force, c.acc, slot, gate);
}

DilateArgs dilate_args(const Ctx &c) {
    DilateArgs d;
    d.nmol = c.rigid.on ? c.rigid.nmol : 0;
    d.n = c.n;
    d.atoms = c.rigid.on ? c.rigid.atoms : nullptr;
    d.member = c.rigid.on ? c.rigid.member : nullptr;
    d.mO = c.rigid.mO;
    d.mH = c.rigid.mH;
    d.pos = c.pos[c.cur];
    d.vblocks = 0;
    d.tgate = nullptr;
    d.vel = nullptr;
    d.thermo = nullptr;
    d.tconv = d.T0 = d.coupling = 0.0;
    d.tmode = 0;
    return d;
}

void dilate(Ctx &c, const int *gate) {
    const DilateArgs d = dilate_args(c);
    launch<dilate_body, 256>(c, cdiv(std::max(d.nmol, d.n), 256), 0, d, c.box, gate);
}

void rigid_map(Ctx &c) {
    if (!c.rigid.on) return;
    // captured list rebuilds clear member in k_copy_back and, with E3B on,
    // map the molecules in the E3B map kernel (fewer gated launches per step)
    if (c.capture && c.e3b.on) return;
    if (!c.capture) fill_int(c, c.rigid.member, c.n, 0, nullptr);
    launch<rigid_map_body, 128>(c, cdiv(c.rigid.nmol, 128), 0, rebuild_gate(c), c.rigid.nmol, c.rigid.ids, c.idToIdx, c.rigid.atoms, c.rigid.member);
}

int fused_check_blocks(const Ctx &c) {
    const int nb = cdiv(c.rigid.nmol, 128);
    return c.rigid.on && c.rigid.nsites == c.n && nb <= kCheckBlocks ? nb : 0;
}

void rigid_first_half(Ctx &c) {
    RigidArgs r = make_args(c);
    // captured steps of all-rigid systems: the skin check's partial maxima
    // here, and the step's decisions in the launch's last workgroup
    const bool chk = c.fuseCheck && fused_check_blocks(c) > 0;
    DecideArgs da{};
    RefreshArgs ra{};
    if (chk && c.fuseStep) {
        da = decide_args(c);
        da.ticket = c.dFlags + 10;
        c.decided = true;
        ra.xq = c.xq;
        ra.rec = c.pppm.on && c.hasCharges ? c.meshRec : nullptr;
        ra.q = c.q[c.cur];
        ra.force = c.force;
        ra.acc = c.acc;
        ra.nacc = kAccTotal;
        c.refreshed = true;
    }
    launch<rigid_vv1_body, 128>(c, cdiv(r.nmol, 128), 0, r, c.pos[c.cur], c.vel[c.cur], c.force, c.posLast, 2.0 * c.dt, chk ? c.checkPart : nullptr, da, ra);
}

void rigid_second_half(Ctx &c, bool kick) {
    RigidArgs r = make_args(c);
    BaroFuse bf{};
    // a captured all-rigid barostat step: the pressure, box and alpha in the
    // last workgroup (the molecular virial and M V^2 are summed here)
    if (kick && c.fuseMolVirial && c.fuseStep && c.baro.on && c.capture) {
        bf.ticket = c.dFlags + 9;
        bf.gate = baro_gate(c);
        bf.bx = c.box;
        bf.bp = box_params(c);
        bf.ba = baro_args(c);
        c.baroFused = true;
    }
    if (kick && (c.fuseMolVirial || c.fuseThermo))
        launch<rigid_vv2_body<true>, 128>(c, cdiv(r.nmol, 128), 0, r, c.pos[c.cur], c.vel[c.cur], c.force, 1, c.acc, baro_gate(c), c.fuseMolVirial ? 1 : 0, c.thermo, thermo_gate(c), c.fuseThermo ? 1 : 0, bf);
    else
        launch<rigid_vv2_body<false>, 128>(c, cdiv(r.nmol, 128), 0, r, c.pos[c.cur], c.vel[c.cur], c.force, kick ? 1 : 0, c.acc, nullptr, 0, nullptr, nullptr, 0, bf);
}

}  // namespace mdx
