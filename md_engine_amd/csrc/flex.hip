// Flexible water (q-TIP4P/F): the intramolecular potential of FixBondQuartic
// (src/Evaluators/BondEvaluatorQuartic.h:9-36: U = k2 dr^2 + k3 dr^3 + k4 dr^4
// per O-H bond) and FixAngleHarmonic (src/Evaluators/AngleEvaluatorHarmonic.h:
// 12-78: U = k/2 (theta - theta0)^2 per H-O-H angle), and the massless M site
// of FixTIP4PFlexible (src/Fixes/FixTIP4PFlexible.cu:92-300): placed at
// gamma r_O + (1 - gamma)/2 (r_H1 + r_H2) before the forces, its force spread
// as gamma F_M onto O and (1 - gamma)/2 F_M onto each H after them.
//
// One thread per molecule; FP64 arithmetic (a few hundred flops per molecule),
// forces into the fixed-point buffer.  Only the molecule's own atoms are
// written, so the M spread needs no atomics.
#include "mdx_internal.h"

namespace mdx {

namespace {

struct FlexArgs {
    int nmol;
    const int4 *atoms;
    double k2, k3, k4, r0, ktheta, theta0, gamma;
    const BoxDev *bx;
};

struct v3 {
    double x, y, z;
};
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 at(const double4 *p, int i) { return {p[i].x, p[i].y, p[i].z}; }
__device__ __forceinline__ v3 mimg(v3 d, const Geometry &g) {
    return {d.x - g.L[0] * rint(d.x * g.invL[0]) * g.periodic[0], d.y - g.L[1] * rint(d.y * g.invL[1]) * g.periodic[1],
            d.z - g.L[2] * rint(d.z * g.invL[2]) * g.periodic[2]};
}
__device__ __forceinline__ void fadd(unsigned long long *f, int i, v3 v) {
    fx_add(&f[3 * i + 0], v.x);
    fx_add(&f[3 * i + 1], v.y);
    fx_add(&f[3 * i + 2], v.z);
}

__device__ __forceinline__ void flex_map_body(int blk, int nblk, const int *__restrict__ gate, int nmol, const int4 *__restrict__ ids,
                           const int *__restrict__ idToIdx, int4 *__restrict__ atoms) {
    if (gate && *gate == 0) return;
    const int m = blk * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int4 d = ids[m];
    atoms[m] = make_int4(idToIdx[d.x], idToIdx[d.y], idToIdx[d.z], d.w >= 0 ? idToIdx[d.w] : -1);
}

// M = O + (1 - gamma)/2 ((H1 - O) + (H2 - O)), hydrogens in O's image
__device__ __forceinline__ void flex_place_m_body(int blk, int nblk, FlexArgs a, double4 *__restrict__ pos) {
    const int m = blk * blockDim.x + threadIdx.x;
    if (m >= a.nmol) return;
    const int4 t = a.atoms[m];
    if (t.w < 0) return;
    const Geometry g = a.bx->geo;
    const v3 o = at(pos, t.x);
    const v3 r1 = mimg(sub(at(pos, t.y), o), g), r2 = mimg(sub(at(pos, t.z), o), g);
    const double c = 0.5 * (1.0 - a.gamma);
    double4 pm = pos[t.w];
    pm.x = o.x + c * (r1.x + r2.x);
    pm.y = o.y + c * (r1.y + r2.y);
    pm.z = o.z + c * (r1.z + r2.z);
    pos[t.w] = pm;
}

// U = sum over the two O-H bonds of k2 dr^2 + k3 dr^3 + k4 dr^4 and
// k/2 (theta - theta0)^2 for H1-O-H2.  Energy into ACC_BOND, virial sum of
// r_H - r_O (x) F_H over the three atoms (the terms are translation invariant).
template <bool EV>
__device__ __forceinline__ void flex_forces_body(int blk, int nblk, FlexArgs a, const double4 *__restrict__ pos,
                                                     unsigned long long *__restrict__ force,
                                                     unsigned long long *__restrict__ acc) {
    const int m = blk * blockDim.x + threadIdx.x;
    double e = 0.0, w[6] = {0, 0, 0, 0, 0, 0};
    if (m < a.nmol) {
        const int4 t = a.atoms[m];
        const Geometry g = a.bx->geo;
        const v3 o = at(pos, t.x);
        const v3 d[2] = {mimg(sub(at(pos, t.y), o), g), mimg(sub(at(pos, t.z), o), g)};
        v3 fh[2] = {{0, 0, 0}, {0, 0, 0}};
        double rr[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const double r = sqrt(dot3(d[b], d[b]));
            rr[b] = r;
            const double dr = r - a.r0;
            const double dU = 2.0 * a.k2 * dr + 3.0 * a.k3 * dr * dr + 4.0 * a.k4 * dr * dr * dr;
            const double s = r > 0.0 ? -dU / r : 0.0;
            fh[b] = {s * d[b].x, s * d[b].y, s * d[b].z};
            if (EV) e += dr * dr * (a.k2 + dr * (a.k3 + dr * a.k4));
        }
        // angle: dU/dtheta = k (theta - theta0); grad_H1 theta = -(u2 - c u1) / (r1 sin)
        const double c = fmin(1.0, fmax(-1.0, dot3(d[0], d[1]) / (rr[0] * rr[1])));
        const double th = acos(c), sn = fmax(sqrt(1.0 - c * c), 1e-12);
        const double dth = th - a.theta0;
        const double pre = a.ktheta * dth / sn;  // -dU/dtheta / (-sin theta)
        const v3 u1 = {d[0].x / rr[0], d[0].y / rr[0], d[0].z / rr[0]}, u2 = {d[1].x / rr[1], d[1].y / rr[1], d[1].z / rr[1]};
        // F_H1 = -dU/dtheta dtheta/dr1 = k dth / sin (u2 - c u1) / r1
        const v3 a1 = {pre * (u2.x - c * u1.x) / rr[0], pre * (u2.y - c * u1.y) / rr[0], pre * (u2.z - c * u1.z) / rr[0]};
        const v3 a2 = {pre * (u1.x - c * u2.x) / rr[1], pre * (u1.y - c * u2.y) / rr[1], pre * (u1.z - c * u2.z) / rr[1]};
        fh[0] = {fh[0].x + a1.x, fh[0].y + a1.y, fh[0].z + a1.z};
        fh[1] = {fh[1].x + a2.x, fh[1].y + a2.y, fh[1].z + a2.z};
        if (EV) e += 0.5 * a.ktheta * dth * dth;
        fadd(force, t.y, fh[0]);
        fadd(force, t.z, fh[1]);
        fadd(force, t.x, v3{-fh[0].x - fh[1].x, -fh[0].y - fh[1].y, -fh[0].z - fh[1].z});
        if (EV) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                w[0] += d[b].x * fh[b].x;
                w[1] += d[b].y * fh[b].y;
                w[2] += d[b].z * fh[b].z;
                w[3] += d[b].x * fh[b].y;
                w[4] += d[b].x * fh[b].z;
                w[5] += d[b].y * fh[b].z;
            }
        }
    }
    if (EV) {
        e = team_sum_d<64>(e);
        if ((threadIdx.x & 63) == 0 && e != 0.0) fx_add(&acc_rep(acc)[ACC_BOND], e);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double x = team_sum_d<64>(w[k]);
            if ((threadIdx.x & 63) == 0 && x != 0.0) fx_add(&acc_rep(acc)[ACC_VIR + k], x);
        }
    }
}

// F_O += gamma F_M, F_H1 = F_H2 += (1 - gamma)/2 F_M, F_M = 0.  Integer
// arithmetic: the H shares are rounded, the O share takes the remainder, so
// the total force is conserved exactly.
__device__ __forceinline__ void flex_spread_m_body(int blk, int nblk, FlexArgs a, unsigned long long *__restrict__ force) {
    const int m = blk * blockDim.x + threadIdx.x;
    if (m >= a.nmol) return;
    const int4 t = a.atoms[m];
    if (t.w < 0) return;
    const double ch = 0.5 * (1.0 - a.gamma);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const long long fm = (long long)force[3 * t.w + k];
        const long long fhh = llrint((double)fm * ch);
        const long long fo = fm - 2 * fhh;
        force[3 * t.x + k] = (unsigned long long)((long long)force[3 * t.x + k] + fo);
        force[3 * t.y + k] = (unsigned long long)((long long)force[3 * t.y + k] + fhh);
        force[3 * t.z + k] = (unsigned long long)((long long)force[3 * t.z + k] + fhh);
        force[3 * t.w + k] = 0ull;
    }
}

FlexArgs make_args(const Ctx &c) {
    const FlexState &F = c.flex;
    FlexArgs a;
    a.nmol = F.nmol;
    a.atoms = F.atoms;
    a.k2 = F.k2;
    a.k3 = F.k3;
    a.k4 = F.k4;
    a.r0 = F.r0;
    a.ktheta = F.ktheta;
    a.theta0 = F.theta0;
    a.gamma = F.gamma;
    a.bx = c.box;
    return a;
}

}  // namespace

void flex_map(Ctx &c) {
    if (!c.flex.on) return;
    launch<flex_map_body, 256>(c, cdiv(c.flex.nmol, 256), 0, rebuild_gate(c), c.flex.nmol, c.flex.ids, c.idToIdx, c.flex.atoms);
}

void flex_place_m(Ctx &c) {
    launch<flex_place_m_body, 256>(c, cdiv(c.flex.nmol, 256), 0, make_args(c), c.pos[c.cur]);
}

void flex_forces(Ctx &c, bool energy, bool virial) {
    const dim3 grid(cdiv(c.flex.nmol, 256));
    if (energy || virial)
        launch<flex_forces_body<true>, 256>(c, grid.x, 0, make_args(c), c.pos[c.cur], c.force, c.acc);
    else
        launch<flex_forces_body<false>, 256>(c, grid.x, 0, make_args(c), c.pos[c.cur], c.force, c.acc);
}

void flex_spread_m(Ctx &c) {
    launch<flex_spread_m_body, 256>(c, cdiv(c.flex.nmol, 256), 0, make_args(c), c.force);
}

}  // namespace mdx
