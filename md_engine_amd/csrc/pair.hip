// Fused Lennard-Jones + Ewald real-space pair forces over the full neighbor
// list.  Replaces compute_force_iso / compute_energy_iso
// (src/Evaluators/PairEvaluateIso.h:8-358) instantiated with EvaluatorLJ
// (PairEvaluatorLJ.h:17-60) and ChargeEvaluatorEwald
// (ChargeEvaluatorEwald.h:11-63), the fusion the reference reaches through
// charge offloading (State.cpp:648-660, FixPair.cpp:89-97,
// EvaluatorWrapper.h:176-194).
//
// Design (gfx950): a team of P lanes (P = 4..32) owns one atom and strides
// through its row, so each row load is a P-wide coalesced segment and the
// neighbor positions (float4 x - lo, q) are gathered from cell-ordered memory
// that stays L2-resident.  Rows hold charged neighbors first and uncharged
// ones at the end, so an atom reads only the part that can interact with it:
// Coulomb needs q_i != 0 and charged neighbors, LJ needs a nonzero epsilon
// for the type pair (per-type flags say which part has such partners).  For
// TIP4P water an H or M site reads its charged part with Coulomb only and an
// O site its uncharged part with LJ only.  Arithmetic is FP32; the team's partial sums are
// combined with butterfly shuffles and added once per atom into the 64-bit
// fixed-point force buffer.  Energy and virial (0.5-weighted, full list) are
// reduced in FP64 per block and added with one fixed-point atomic per block.
#include "erfc_fit.h"
#include "mdx_internal.h"

namespace mdx {

namespace {

struct PairArgs {
    int n;
    const float4 *xq;
    const int *type;
    int cap;
    const uint32_t *nl;
    const int *nlCount;
    const int *nlCountA;
    const int *typeFlags;
    const float *params;
    int ntypes;
    const BoxDev *bx;             // box lengths and alpha
    float spec[3];
    float qqr, rq2;
    unsigned long long *force;
    unsigned long long *acc;
    const int *vgate;  // EVM 2 inside a captured step: the barostat gate (0: no trace this step)
    RowMaxArgs rm;     // a captured build's row maxima in one extra workgroup (rm.on)
};

// EVM: 0 forces only; 1 energy and virial tensor; 2 the virial's trace only
// (sum r F, what a barostat needs), into ACC_VTRACE, skipped on steps whose
// barostat gate is 0
template <int P, bool LJ, bool COUL, int EVM>
__device__ __forceinline__ void pair_body(int blk, int nblk, PairArgs a) {
    constexpr bool EV = EVM == 1, VT = EVM == 2;
    if (a.rm.on && blk == nblk - 1) {  // uniform over the workgroup
        if (!a.rm.gate || *a.rm.gate != 0) row_max_wg(a.rm.na, a.rm.a, a.rm.nb, a.rm.b, a.rm.flags);
        return;
    }
    // box values into locals
    const float bL0 = a.bx->Lf[0], bL1 = a.bx->Lf[1], bL2 = a.bx->Lf[2];
    const float sp1 = a.spec[0], sp2 = a.spec[1], sp3 = a.spec[2];
    const float bI0 = a.bx->invLf[0], bI1 = a.bx->invLf[1], bI2 = a.bx->invLf[2];
    const float alphaB = a.bx->alphaF;
    __shared__ float sp[3 * kMaxTypes * kMaxTypes];
    __shared__ int sfl[kMaxTypes];
    __shared__ double wred[4][7];
    const int T2 = a.ntypes * a.ntypes;
    if (LJ) {
        for (int t = threadIdx.x; t < 3 * T2; t += blockDim.x) sp[t] = a.params[t];
        for (int t = threadIdx.x; t < a.ntypes; t += blockDim.x) sfl[t] = a.typeFlags[t];
        __syncthreads();
    }
    const int gt = blk * blockDim.x + threadIdx.x;
    const int i = gt / P;
    const int tl = threadIdx.x & (P - 1);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    float e = 0.f, v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
    if (i < a.n) {
        const float4 pi = a.xq[i];
        const int ti = (LJ && a.ntypes > 1) ? a.type[i] : 0;
        const int fl = LJ ? sfl[ti] : 0;
        const int cnt = min(a.nlCount[i], a.cap);
        const int nA = min(a.nlCountA[i], cnt);
        const uint32_t *row = a.nl + (size_t)i * a.cap;
        const float twoAlphaRtPi = 2.0f * 0.5641895835477563f * alphaB;
        const float alpha2 = alphaB * alphaB;
        // one row entry: tag, neighbor record, neighbor type
        auto term = [&](uint32_t tag, const float4 &pj, int tj, bool lj, bool coul) {
            float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            dx -= bL0 * rintf(dx * bI0);
            dy -= bL1 * rintf(dy * bI1);
            dz -= bL2 * rintf(dz * bI2);
            const float r2 = dx * dx + dy * dy + dz * dz;
            // special-pair multiplier by selection (a dynamic index would
            // put the argument pack in scratch memory)
            const float m = tag == 0 ? 1.0f : (tag == 1 ? sp1 : (tag == 2 ? sp2 : sp3));
            float fs = 0.f;
            bool any = false;
            if (lj) {
                const int tp = a.ntypes > 1 ? ti * a.ntypes + tj : 0;
                if (r2 < sp[tp]) {
                    any = true;
                    if (m != 0.0f) {
                        const float eps24 = sp[T2 + tp], sig6 = sp[2 * T2 + tp];
                        const float r2inv = 1.0f / r2;
                        const float r6inv = r2inv * r2inv * r2inv;
                        const float p1 = eps24 * 2.0f * sig6 * sig6;
                        const float p2 = eps24 * sig6;
                        fs += r6inv * r2inv * (p1 * r6inv - p2) * m;
                        if (EV) {
                            const float rc2 = sp[tp];
                            const float s6r6 = sig6 * r6inv;
                            const float rc6 = rc2 * rc2 * rc2;
                            const float s6rc6 = sig6 / rc6;
                            e += 0.5f * (eps24 * (1.0f / 6.0f)) * (s6r6 * (s6r6 - 1.0f) - s6rc6 * (s6rc6 - 1.0f)) * m;
                        }
                    }
                }
            }
            if (coul && r2 < a.rq2) {
                any = true;
                const float ls = fmaxf(r2, 1e-10f);
                const float rinv = rsqrtf(ls);
                const float len = ls * rinv;
                const float qq = a.qqr * pi.w * pj.w;
                // erfc(ar) - (1 - m); for a full exclusion (m = 0) this is -erf(ar).
                // erfc shares the Gaussian's exponential (erfc_fit.h)
                const float gauss = __expf(-alpha2 * ls);
                const float ec = erfc_from_gauss(alphaB * len, gauss) - (1.0f - m);
                const float s = qq * (ec * rinv + twoAlphaRtPi * gauss);
                fs += s * rinv * rinv;
                if (EV) e += 0.5f * qq * rinv * ec;
            }
            if (any) {
                fx += dx * fs;
                fy += dy * fs;
                fz += dz * fs;
                if (VT) v0 += fs * r2;
                if (EV) {
                    v0 += dx * fs * dx;
                    v1 += dy * fs * dy;
                    v2 += dz * fs * dz;
                    v3 += dx * fs * dy;
                    v4 += dx * fs * dz;
                    v5 += dy * fs * dz;
                }
            }
        };
        // part 0: charged neighbors [0, nA); part 1: charge-free ones [nA, cnt).
        // A lane takes entries tl, tl + P, tl + 2P, ... in that order, U of them
        // per pass with all their loads in flight before any is used (one
        // dependent row-then-record load chain per entry left the lanes waiting).
        constexpr int U = 4;
        for (int part = 0; part < 2; ++part) {
            const bool lj = LJ && (fl & (1 << part));
            const bool coul = COUL && part == 0 && pi.w != 0.0f;
            if (!lj && !coul) continue;
            const int k0 = part == 0 ? 0 : nA, k1 = part == 0 ? nA : cnt;
            for (int kb = k0 + tl; kb < k1; kb += U * P) {
                uint32_t raw[U];
                bool ok[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    ok[u] = kb + u * P < k1;
                    raw[u] = ok[u] ? row[kb + u * P] : 0u;
                }
                float4 pj[U];
                int tj[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int j = ok[u] ? (int)(raw[u] & 0x3FFFFFFFu) : i;
                    pj[u] = a.xq[j];
                    tj[u] = (lj && a.ntypes > 1) ? a.type[j] : 0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (ok[u]) term(raw[u] >> 30, pj[u], tj[u], lj, coul);
            }
        }
    }
    fx = team_sum<P>(fx);
    fy = team_sum<P>(fy);
    fz = team_sum<P>(fz);
    if (tl == 0 && i < a.n) {
        fx_add(&a.force[3 * i + 0], fx);
        fx_add(&a.force[3 * i + 1], fy);
        fx_add(&a.force[3 * i + 2], fz);
    }
    if (VT && (!a.vgate || *a.vgate)) {  // one value for the whole workgroup
        double x = team_sum_d<64>(0.5 * (double)v0);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) wred[w][0] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            x = 0.0;
            for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) x += wred[ww][0];
            fx_add(&acc_rep(a.acc)[ACC_VTRACE], x);
        }
    }
    if (EV) {
        double vals[7] = {e, 0.5 * v0, 0.5 * v1, 0.5 * v2, 0.5 * v3, 0.5 * v4, 0.5 * v5};
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            double x = team_sum_d<64>(vals[s]);
            if (lane == 0) wred[w][s] = x;
        }
        __syncthreads();
        if (threadIdx.x < 7) {
            double x = 0.0;
            for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) x += wred[ww][threadIdx.x];
            fx_add(&acc_rep(a.acc)[threadIdx.x == 0 ? ACC_PAIR : ACC_VIR + threadIdx.x - 1], x);
        }
    }
}

template <int P, bool LJ, bool COUL>
void launch3(Ctx &c, const PairArgs &a, int evm) {
    const int blocks = cdiv((long long)a.n * P, 256) + a.rm.on;
    if (evm == 1) launch<pair_body<P, LJ, COUL, 1>, 256>(c, blocks, 0, a);
    else if (evm == 2) launch<pair_body<P, LJ, COUL, 2>, 256>(c, blocks, 0, a);
    else launch<pair_body<P, LJ, COUL, 0>, 256>(c, blocks, 0, a);
}

template <int P>
void launch2(Ctx &c, const PairArgs &a, bool lj, bool coul, int evm) {
    if (lj && coul) launch3<P, true, true>(c, a, evm);
    else if (lj) launch3<P, true, false>(c, a, evm);
    else launch3<P, false, true>(c, a, evm);
}

}  // namespace

// Per-type flags of the LJ table against the charge classes present: bit 0 =
// nonzero epsilon with some type that has charged atoms, bit 1 = with some type
// that has uncharged atoms.  Recomputed when the system or the table changes.
std::vector<int> pair_type_flag_table(const Ctx &c) {
    std::vector<int> charged(kMaxTypes, 0), neutral(kMaxTypes, 0), fl(kMaxTypes, 0);
    for (size_t i = 0; i < c.typeHost.size(); ++i) (c.qHost[i] != 0.0f ? charged : neutral)[c.typeHost[i]] = 1;
    const int T = c.ntypes;
    for (int a = 0; a < T && c.ljOn; ++a)
        for (int b = 0; b < T; ++b) {
            if (c.ljEps[a * T + b] == 0.0) continue;
            if (charged[b]) fl[a] |= 1;
            if (neutral[b]) fl[a] |= 2;
        }
    return fl;
}

static void pair_type_flags(Ctx &c) {
    if (!c.pairTypeFlags) c.pairTypeFlags = dalloc<int>(c, kMaxTypes);
    const std::vector<int> fl = pair_type_flag_table(c);
    MDX_REQUIRE(!c.capture, 3, "pair type flags must be set before a step is captured");
    MDX_HIP(hipMemcpyAsync(c.pairTypeFlags, fl.data(), sizeof(int) * kMaxTypes, hipMemcpyHostToDevice, c.stream));
    MDX_HIP(hipStreamSynchronize(c.stream));
    c.pairFlagsDirty = false;
}

void pair_forces(Ctx &c, bool energy, bool virial) {
    const bool coul = c.ewaldReal && c.hasCharges;
    if (!c.ljOn && !coul) return;
    if (c.pairFlagsDirty) pair_type_flags(c);
    PairArgs a;
    a.n = c.n;
    a.xq = c.xq;
    a.type = c.type[c.cur];
    a.cap = c.cap;
    a.nl = c.nl;
    a.nlCount = c.nlCount;
    a.nlCountA = c.nlCountA;
    a.typeFlags = c.pairTypeFlags;
    a.params = c.ljParams;
    a.ntypes = c.ntypes;
    for (int d = 0; d < 3; ++d) {
        a.spec[d] = c.special[d];
    }
    a.bx = c.box;
    a.qqr = c.qqr;
    a.rq2 = c.rq2;
    a.force = c.force;
    a.acc = c.acc;
    a.vgate = baro_gate(c);
    a.rm = row_max_args(c);
    const int evm = (energy || virial) ? 1 : (c.vtrace ? 2 : 0);
    ProfScope ps(c, PK_PAIR);
    switch (c.pairTeam) {
        case 4: launch2<4>(c, a, c.ljOn, coul, evm); break;
        case 8: launch2<8>(c, a, c.ljOn, coul, evm); break;
        case 16: launch2<16>(c, a, c.ljOn, coul, evm); break;
        default: launch2<32>(c, a, c.ljOn, coul, evm); break;
    }
}

}  // namespace mdx
