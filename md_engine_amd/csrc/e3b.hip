// E3B three-body water forces.  Replaces FixE3B::compute (src/Fixes/FixE3B.cu:
// 85-347) and its kernels compute_E3B_force_twobody / _threebody
// (src/Evaluators/ThreeBodyE3B.h:18-549) with the evaluator functions of
// EvaluatorE3B.h:67-249.
//
// Design (gfx950): owner-computes in two passes, no float atomics and no
// per-slot scratch arrays (the reference stores four force vectors and four
// energies per list slot and clears seven arrays every step):
//   pass 1: a team per molecule walks its molecule list, adds the two-body
//           O-O force to its own oxygen, and sums its own totals
//           x = sum f(H_j -> O_i), y = sum f(H1_i -> O_j), z = sum f(H2_i -> O_j);
//   pass 2: a team per molecule recomputes the four hydrogen-bond weights of
//           each pair within r_max, reads the partner's totals, and applies
//           dE/df of every bond to its own O, H1, H2 (the fscale values of
//           threeBodyPairForce).  Each pair is evaluated by both owners, so
//           every atom is written by exactly one team.
// The pair test r_OO <= 6.1572 (RMAX_E3B, ThreeBodyE3B.h:15) is symmetric, so
// both owners agree on it.
#include "mdx_internal.h"

namespace mdx {

namespace {

struct E3BArgs {
    int nmol;
    const int4 *molAtoms;
    const float4 *xq;
    int cap;
    const int *nl;
    const int *nlCount;
    float L[3], invL[3];  // taken from bx in the kernels
    const BoxDev *bx;
    float E2, Ea, Eb, Ec, k2, k3, rs, rf, r2b2, rmax;
    float4 *tot;
    unsigned long long *force;
    unsigned long long *acc;
    int three;    // 0: two-body term only
    float *emol;  // optional: three-body energy owned by each molecule
    const int *vgate;  // EVM 2 inside a captured step: the barostat gate (0: no trace this step)
};

__device__ __forceinline__ float3 mi3(float4 a, float4 b, const E3BArgs &p) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    dx -= p.L[0] * rintf(dx * p.invL[0]);
    dy -= p.L[1] * rintf(dy * p.invL[1]);
    dz -= p.L[2] * rintf(dz * p.invL[2]);
    return make_float3(dx, dy, dz);
}

// f(r) = s(r) exp(-k3 r) for dr = r_H - r_O (threeBodyPairEnergy,
// EvaluatorE3B.h:119-125); g = dr (k3 e s - e s') / r = -grad_H f
// (threeBodyForce, 127-141).
__device__ __forceinline__ float hb(const float3 &dr, const E3BArgs &p, float3 &g) {
    const float r = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
    if (r > p.rf) {
        g = make_float3(0.f, 0.f, 0.f);
        return 0.f;
    }
    const float es = __expf(-p.k3 * r);
    const float inv = 1.0f / ((p.rf - p.rs) * (p.rf - p.rs) * (p.rf - p.rs));
    float s = 1.0f, fs = p.k3 * es / r;
    if (r > p.rs) {
        const float u = p.rf - r;
        s = u * u * (p.rf + 2.0f * r - 3.0f * p.rs) * inv;
        const float ds = 6.0f * (p.rf - r) * (p.rs - r) * inv;
        fs = fs * s - es * ds / r;
    }
    g = make_float3(dr.x * fs, dr.y * fs, dr.z * fs);
    return s * es;
}

__device__ __forceinline__ float hb_only(const float3 &dr, const E3BArgs &p) {
    const float r = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
    if (r > p.rf) return 0.f;
    const float es = __expf(-p.k3 * r);
    if (r <= p.rs) return es;
    const float u = p.rf - r;
    return u * u * (p.rf + 2.0f * r - 3.0f * p.rs) / ((p.rf - p.rs) * (p.rf - p.rs) * (p.rf - p.rs)) * es;
}

// energy (into eslot) and the six virial components: wave sums, then one
// atomic per workgroup and value into the workgroup's accumulator replica.
// Every thread of the workgroup calls it.
__device__ __forceinline__ void block_acc7(const double (&v)[7], int eslot, unsigned long long *acc) {
    __shared__ double red[4][7];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        const double x = team_sum_d<64>(v[s]);
        if (lane == 0) red[w][s] = x;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double x = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) x += red[k][threadIdx.x];
        if (x != 0.0) fx_add(&acc_rep(acc)[threadIdx.x == 0 ? eslot : ACC_VIR + threadIdx.x - 1], x);
    }
}

// one value: wave sums, one atomic per workgroup
__device__ __forceinline__ void block_acc1(double v, int slot, unsigned long long *acc) {
    __shared__ double red1[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double x = team_sum_d<64>(v);
    if (lane == 0) red1[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red1[k];
        if (t != 0.0) fx_add(&acc_rep(acc)[slot], t);
    }
}

__device__ __forceinline__ void box_into(E3BArgs &p) {
    for (int d = 0; d < 3; ++d) {
        p.L[d] = p.bx->Lf[d];
        p.invL[d] = p.bx->invLf[d];
    }
}

// EVM: 0 forces only; 1 energy and virial tensor; 2 the virial's trace only
template <int P, int EVM>
__device__ __forceinline__ void e3b_pass1_body(int blk, int nblk, E3BArgs p) {
    constexpr bool EV = EVM == 1, VT = EVM == 2;
    box_into(p);
    const int gt = blk * blockDim.x + threadIdx.x;
    const int i = gt / P;
    const int tl = threadIdx.x & (P - 1);
    float fx = 0.f, fy = 0.f, fz = 0.f, tx = 0.f, ty = 0.f, tz = 0.f, e2 = 0.f;
    float w0 = 0.f, w1 = 0.f, w2 = 0.f, w3 = 0.f, w4 = 0.f, w5 = 0.f;
    int4 ai = make_int4(0, 0, 0, 0);
    if (i < p.nmol) {
        ai = p.molAtoms[i];
        const float4 Oi = p.xq[ai.x], H1i = p.xq[ai.y], H2i = p.xq[ai.z];
        const int cnt = min(p.nlCount[i], p.cap);
        const int *row = p.nl + (size_t)i * p.cap;
        for (int k = tl; k < cnt; k += P) {
            const int4 aj = p.molAtoms[row[k]];
            // the hydrogens are loaded with the oxygen (one load level less
            // than loading them after the distance test)
            const float4 Oj = p.xq[aj.x], H1j = p.xq[aj.y], H2j = p.xq[aj.z];
            const float3 d = mi3(Oi, Oj, p);
            const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
            const float r = sqrtf(r2);
            if (r2 < p.r2b2) {
                const float ex = __expf(-p.k2 * r);
                const float s = p.k2 * p.E2 * ex / r;
                fx += d.x * s;
                fy += d.y * s;
                fz += d.z * s;
                if (VT) w0 += s * r2;
                if (EV) {
                    e2 += 0.5f * p.E2 * ex;
                    w0 += d.x * s * d.x;
                    w1 += d.y * s * d.y;
                    w2 += d.z * s * d.z;
                    w3 += d.x * s * d.y;
                    w4 += d.x * s * d.z;
                    w5 += d.y * s * d.z;
                }
            }
            if (p.three && r <= p.rmax) {
                tx += hb_only(mi3(H1j, Oi, p), p) + hb_only(mi3(H2j, Oi, p), p);
                ty += hb_only(mi3(H1i, Oj, p), p);
                tz += hb_only(mi3(H2i, Oj, p), p);
            }
        }
    }
    fx = team_sum<P>(fx);
    fy = team_sum<P>(fy);
    fz = team_sum<P>(fz);
    tx = team_sum<P>(tx);
    ty = team_sum<P>(ty);
    tz = team_sum<P>(tz);
    if (tl == 0 && i < p.nmol) {
        fx_add(&p.force[3 * ai.x + 0], fx);
        fx_add(&p.force[3 * ai.x + 1], fy);
        fx_add(&p.force[3 * ai.x + 2], fz);
        p.tot[i] = make_float4(tx, ty, tz, 0.f);
    }
    if (EV) {
        const double v[7] = {e2, 0.5 * w0, 0.5 * w1, 0.5 * w2, 0.5 * w3, 0.5 * w4, 0.5 * w5};
        block_acc7(v, ACC_E3B2, p.acc);
    }
    if (VT && (!p.vgate || *p.vgate)) block_acc1(0.5 * w0, ACC_VTRACE, p.acc);
}

template <int P, int EVM>
__device__ __forceinline__ void e3b_pass2_body(int blk, int nblk, E3BArgs p) {
    constexpr bool EV = EVM == 1, VT = EVM == 2;
    box_into(p);
    const int gt = blk * blockDim.x + threadIdx.x;
    const int i = gt / P;
    const int tl = threadIdx.x & (P - 1);
    float3 FO = make_float3(0.f, 0.f, 0.f), FH1 = FO, FH2 = FO;
    float ea = 0.f, eb = 0.f, ec = 0.f;
    float w0 = 0.f, w1 = 0.f, w2 = 0.f, w3 = 0.f, w4 = 0.f, w5 = 0.f;
    int4 ai = make_int4(0, 0, 0, 0);
    float4 ti = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < p.nmol) {
        ai = p.molAtoms[i];
        ti = p.tot[i];
        const float4 Oi = p.xq[ai.x], H1i = p.xq[ai.y], H2i = p.xq[ai.z];
        const int cnt = min(p.nlCount[i], p.cap);
        const int *row = p.nl + (size_t)i * p.cap;
        for (int k = tl; k < cnt; k += P) {
            const int mj = row[k];
            const int4 aj = p.molAtoms[mj];
            // every record of the neighbour is loaded at once (one load level)
            const float4 Oj = p.xq[aj.x], H1j = p.xq[aj.y], H2j = p.xq[aj.z];
            const float4 tj = p.tot[mj];
            const float3 d = mi3(Oi, Oj, p);
            const float r = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
            if (!(r <= p.rmax)) continue;
            float3 gA1, gA2, gB1, gB2;
            const float3 dA1 = mi3(H1i, Oj, p), dA2 = mi3(H2i, Oj, p);
            const float3 dB1 = mi3(H1j, Oi, p), dB2 = mi3(H2j, Oi, p);
            const float fA1 = hb(dA1, p, gA1), fA2 = hb(dA2, p, gA2);
            const float fB1 = hb(dB1, p, gB1), fB2 = hb(dB2, p, gB2);
            const float Dij = fA1 + fA2, Dji = fB1 + fB2;
            const float bI = p.Eb * (ti.x + tj.y + tj.z - 2.0f * Dji);
            const float bJ = p.Eb * (tj.x + ti.y + ti.z - 2.0f * Dij);
            const float cA = p.Ec * (tj.x - Dij);
            const float cB = p.Ec * (ti.x - Dji);
            const float cA1 = p.Ea * (ti.z - fA2) + bI + cA;
            const float cA2 = p.Ea * (ti.y - fA1) + bI + cA;
            const float cB1 = p.Ea * (tj.z - fB2) + bJ + cB;
            const float cB2 = p.Ea * (tj.y - fB1) + bJ + cB;
            FH1.x += cA1 * gA1.x;
            FH1.y += cA1 * gA1.y;
            FH1.z += cA1 * gA1.z;
            FH2.x += cA2 * gA2.x;
            FH2.y += cA2 * gA2.y;
            FH2.z += cA2 * gA2.z;
            FO.x -= cB1 * gB1.x + cB2 * gB2.x;
            FO.y -= cB1 * gB1.y + cB2 * gB2.y;
            FO.z -= cB1 * gB1.z + cB2 * gB2.z;
            if (VT)
                w0 += 0.5f * (cA1 * (gA1.x * dA1.x + gA1.y * dA1.y + gA1.z * dA1.z) +
                              cA2 * (gA2.x * dA2.x + gA2.y * dA2.y + gA2.z * dA2.z) +
                              cB1 * (gB1.x * dB1.x + gB1.y * dB1.y + gB1.z * dB1.z) +
                              cB2 * (gB2.x * dB2.x + gB2.y * dB2.y + gB2.z * dB2.z));
            if (EV) {
                ea += fA1 * fA2;
                eb += Dij * Dji;
                ec += Dji * Dji;
                // bond virial c g (x) (r_H - r_O), half per owner
                w0 += 0.5f * (cA1 * gA1.x * dA1.x + cA2 * gA2.x * dA2.x + cB1 * gB1.x * dB1.x + cB2 * gB2.x * dB2.x);
                w1 += 0.5f * (cA1 * gA1.y * dA1.y + cA2 * gA2.y * dA2.y + cB1 * gB1.y * dB1.y + cB2 * gB2.y * dB2.y);
                w2 += 0.5f * (cA1 * gA1.z * dA1.z + cA2 * gA2.z * dA2.z + cB1 * gB1.z * dB1.z + cB2 * gB2.z * dB2.z);
                w3 += 0.5f * (cA1 * gA1.x * dA1.y + cA2 * gA2.x * dA2.y + cB1 * gB1.x * dB1.y + cB2 * gB2.x * dB2.y);
                w4 += 0.5f * (cA1 * gA1.x * dA1.z + cA2 * gA2.x * dA2.z + cB1 * gB1.x * dB1.z + cB2 * gB2.x * dB2.z);
                w5 += 0.5f * (cA1 * gA1.y * dA1.z + cA2 * gA2.y * dA2.z + cB1 * gB1.y * dB1.z + cB2 * gB2.y * dB2.z);
            }
        }
    }
    FO.x = team_sum<P>(FO.x);
    FO.y = team_sum<P>(FO.y);
    FO.z = team_sum<P>(FO.z);
    FH1.x = team_sum<P>(FH1.x);
    FH1.y = team_sum<P>(FH1.y);
    FH1.z = team_sum<P>(FH1.z);
    FH2.x = team_sum<P>(FH2.x);
    FH2.y = team_sum<P>(FH2.y);
    FH2.z = team_sum<P>(FH2.z);
    if (tl == 0 && i < p.nmol) {
        fx_add(&p.force[3 * ai.x + 0], FO.x);
        fx_add(&p.force[3 * ai.x + 1], FO.y);
        fx_add(&p.force[3 * ai.x + 2], FO.z);
        fx_add(&p.force[3 * ai.y + 0], FH1.x);
        fx_add(&p.force[3 * ai.y + 1], FH1.y);
        fx_add(&p.force[3 * ai.y + 2], FH1.z);
        fx_add(&p.force[3 * ai.z + 0], FH2.x);
        fx_add(&p.force[3 * ai.z + 1], FH2.y);
        fx_add(&p.force[3 * ai.z + 2], FH2.z);
    }
    if (EV) {
        ea = team_sum<P>(ea);
        eb = team_sum<P>(eb);
        ec = team_sum<P>(ec);
        double e3 = 0.0;
        if (tl == 0 && i < p.nmol) {
            // E_A + E_B + E_C contributions owned by molecule i (donor for A,
            // centre for B, acceptor for C)
            e3 = (double)p.Ea * ((double)ti.y * ti.z - ea) +
                 (double)p.Eb * ((double)ti.x * (ti.y + ti.z) - eb) +
                 0.5 * (double)p.Ec * ((double)ti.x * ti.x - ec);
            if (p.emol) p.emol[i] = (float)e3;
        }
        const double v[7] = {e3, w0, w1, w2, w3, w4, w5};
        block_acc7(v, ACC_E3B3, p.acc);
    }
    if (VT && (!p.vgate || *p.vgate)) block_acc1(w0, ACC_VTRACE, p.acc);
}

template <int P>
void launch_e3b(Ctx &c, const E3BArgs &a, int evm) {
    const int threads = 256;
    const int blocks = cdiv((long long)a.nmol * P, threads);
    {
        ProfScope ps(c, PK_E3B1);
        if (evm == 1)
            launch<e3b_pass1_body<P, 1>, 256>(c, blocks, 0, a);
        else if (evm == 2)
            launch<e3b_pass1_body<P, 2>, 256>(c, blocks, 0, a);
        else
            launch<e3b_pass1_body<P, 0>, 256>(c, blocks, 0, a);
    }
    if (!a.three) return;
    ProfScope ps(c, PK_E3B2);
    if (evm == 1)
        launch<e3b_pass2_body<P, 1>, 256>(c, blocks, 0, a);
    else if (evm == 2)
        launch<e3b_pass2_body<P, 2>, 256>(c, blocks, 0, a);
    else
        launch<e3b_pass2_body<P, 0>, 256>(c, blocks, 0, a);
}

}  // namespace

void e3b_forces(Ctx &c, bool energy, bool virial, bool three_body, float *emol) {
    E3BState &E = c.e3b;
    if (!E.on || E.nmol == 0) return;
    E3BArgs a;
    a.nmol = E.nmol;
    a.molAtoms = E.molAtoms;
    a.xq = c.xq;
    a.cap = E.cap;
    a.nl = E.nl;
    a.nlCount = E.nlCount;
    for (int d = 0; d < 3; ++d) {
        a.L[d] = 0.0f;
        a.invL[d] = 0.0f;
    }
    a.bx = c.box;
    a.E2 = E.E2;
    a.Ea = E.Ea;
    a.Eb = E.Eb;
    a.Ec = E.Ec;
    a.k2 = E.k2;
    a.k3 = E.k3;
    a.rs = E.rs;
    a.rf = E.rf;
    a.r2b2 = E.r2b * E.r2b;
    a.rmax = E.rmax;
    a.tot = E.tot;
    a.force = c.force;
    a.acc = c.acc;
    a.three = three_body ? 1 : 0;
    a.emol = emol;
    a.vgate = baro_gate(c);
    const int evm = (energy || virial || emol) ? 1 : (c.vtrace ? 2 : 0);
    switch (E.team) {
        case 4: launch_e3b<4>(c, a, evm); break;
        case 8: launch_e3b<8>(c, a, evm); break;
        case 16: launch_e3b<16>(c, a, evm); break;
        default: launch_e3b<32>(c, a, evm); break;
    }
}

}  // namespace mdx
