// Device-resident box and the Berendsen barostat step on the device.
//
// Kernels read every box-dependent quantity (lengths, inverse lengths, cell
// widths, culling and pre-test bounds, mesh spacing, Ewald alpha, P3M scale
// factors) from one BoxDev in device memory.  The host derives it with
// box_derive whenever it changes the box or the grid and uploads it; the
// barostat step recomputes it on the device, so captured graphs stay valid
// while the box breathes.
//
// The barostat step follows FixPressureBerendsen::stepFinal
// (src/Fixes/FixPressureBerendsen.cu:20-51): P from the molecular virial,
// mu = (1 - dt/period (P0 - P)/B)^(1/3) clamped to 1 -+ maxDilation, the box
// and the molecules' centres of mass scaled about the box centre; then, like
// FixChargeEwald::handleBoundsChange (src/Fixes/FixChargeEwald.cu:1200-1229),
// alpha is solved again and the Green function recomputed for the new box.
#include <cstring>

#include "mdx_internal.h"

namespace mdx {

namespace {

struct BaroArgs {
    double P0, period, bulk, maxDil, nktv2p, mvv, dt;
    // P3M alpha re-solve (alphaSolve = 0: alpha stays as uploaded)
    int alphaSolve, natoms, order;
    double q2, rcut;
    // cell edges below these lengths invalidate the grid
    double minL[3];
    // the centre-of-mass M V^2 was summed before a velocity rescale of the same
    // step (tgate open): scale it by that rescale's lambda^2 (k_vscale's formula)
    int rescaled, tmode;
    const int *tgate;
    const unsigned long long *thermo;
    double tconv, T0, coupling;
};

// The secant solve of pppm_solve_alpha for a fixed interpolation order, with
// the box-dependent factors taken out of the iteration and the powers
// unrolled into multiplications (one device thread runs this dependent
// chain; pow would cost an exp and a log per term).  Same functions
// DeltaF_k - DeltaF_real, start 4.79853 / rcut, tolerance 1e-5 and
// iteration limit as the host solve.
template <int ORDER>
__device__ double alpha_solve_dev(int natoms, double q2, const double *L, const int *sz, double rcut) {
    double h[3], cL[3], sL[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        h[d] = L[d] / sz[d];
        cL[d] = 1.0 / (L[d] * L[d]);
        sL[d] = L[d] / natoms * sqrt(2.0 * M_PI);
    }
    const double R0 = 2 * q2 / sqrt(natoms * rcut * L[0] * L[1] * L[2]), rc2 = rcut * rcut, q3 = q2 / 3.0;
    auto dfk = [&](double a) {
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double x = h[d] * a, x2 = x * x;
            double sum = 0.0, p = 1.0, po = 1.0;
#pragma unroll
            for (int mm = 0; mm < ORDER; ++mm) {
                sum += pppm_amp(ORDER, mm) * p;
                p *= x2;
                po *= x;
            }
            acc += cL[d] * po * sqrt(a * sL[d] * sum);
        }
        return q3 * acc;
    };
    auto dfr = [&](double a) { return R0 * exp(-a * a * rc2); };
    double xa = 0.0, xb = 4.79853 / rcut;
    double ya = dfk(xa) - dfr(xa);
    double fr = dfr(xb);
    double yb = dfk(xb) - fr;
    for (int it = 0; it < 100 && fabs(yb) / fr > 1e-5; ++it) {
        const double kinv = (xb - xa) / (yb - ya);
        ya = yb;
        xa = xb;
        xb = xa - ya * kinv;
        fr = dfr(xb);
        yb = dfk(xb) - fr;
    }
    return xb;
}

__device__ double alpha_solve_any(int order, int natoms, double q2, const double *L, const int *sz, double rcut) {
    switch (order) {
        case 1: return alpha_solve_dev<1>(natoms, q2, L, sz, rcut);
        case 2: return alpha_solve_dev<2>(natoms, q2, L, sz, rcut);
        case 3: return alpha_solve_dev<3>(natoms, q2, L, sz, rcut);
        case 4: return alpha_solve_dev<4>(natoms, q2, L, sz, rcut);
        case 5: return alpha_solve_dev<5>(natoms, q2, L, sz, rcut);
        case 6: return alpha_solve_dev<6>(natoms, q2, L, sz, rcut);
        default: return alpha_solve_dev<7>(natoms, q2, L, sz, rcut);
    }
}

// One wave: the accumulator replicas are summed across lanes (integer sums,
// exact in any order), then lane 0 computes the pressure, the dilation factor,
// the new box and, for P3M, alpha and the Green function's image range.
__device__ __forceinline__ void baro_body(int blk, int nblk, const int *__restrict__ gate, BoxDev *__restrict__ bx, BoxParams bp,
                                             BaroArgs ba, const unsigned long long *__restrict__ acc) {
    if (gate && *gate == 0) return;
    const int lane = threadIdx.x;
    constexpr int kSlots[6] = {ACC_VIR, ACC_VIR + 1, ACC_VIR + 2, ACC_MOL, ACC_MOL + 1, ACC_VTRACE};
    long long t[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        t[k] = lane < kAccRep ? (long long)acc[lane * ACC_N + kSlots[k]] : 0ll;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t[k] += __shfl_xor(t[k], o);
    }
    if (lane != 0) return;
    // trace of the virial: the tensor's diagonal plus the trace-only terms
    const double W = (double)t[0] * kFxInv + (double)t[1] * kFxInv + (double)t[2] * kFxInv + (double)t[5] * kFxInv;
    const double corr = (double)t[3] * kFxInv;
    double mv2 = (double)t[4] * kFxInv;
    if (ba.rescaled && (!ba.tgate || *ba.tgate)) {
        const double T = (double)(long long)ba.thermo[0] * kFxInv * ba.tconv;
        if (T > 0.0) {
            const double l2 = ba.tmode == 0 ? ba.T0 / T : 1.0 + ba.coupling * (ba.T0 / T - 1.0);
            mv2 *= fmax(l2, 0.0);
        }
    }
    const double P = (mv2 * ba.mvv + W - corr) / (3.0 * bx->vol) * ba.nktv2p;
    double mu = cbrt(1.0 - ba.dt / ba.period * (ba.P0 - P) / ba.bulk);
    mu = fmin(fmax(mu, 1.0 - ba.maxDil), 1.0 + ba.maxDil);
    bx->pressure = P;
    bx->mu = mu;
    double lo[3], L[3];
    for (int d = 0; d < 3; ++d) {
        const double ctr = bx->geo.lo[d] + 0.5 * bx->geo.L[d];
        L[d] = bx->geo.L[d] * mu;
        lo[d] = ctr - 0.5 * L[d];
        if (bp.periodic[d] && L[d] < ba.minL[d]) bx->regrid = 1;
    }
    // box_derive leaves alpha, alphaF and greenLim as they are
    box_derive(*bx, bp, lo, L);
    if (ba.alphaSolve) {
        const double a = alpha_solve_any(ba.order, ba.natoms, ba.q2, L, bp.meshSz, ba.rcut);
        bx->alpha = a;
        bx->alphaF = (float)a;
        const double hm = cbrt(bx->mesh.h[0] * bx->mesh.h[1] * bx->mesh.h[2]);
        bx->greenLim = green_lim(a, hm);
    }
}

}  // namespace

BoxParams box_params(const Ctx &c) {
    BoxParams p;
    for (int a = 0; a < 3; ++a) {
        p.periodic[a] = c.geo.periodic[a];
        p.nc[a] = c.nc[a];
        p.reach[a] = c.reach[a];
        p.meshSz[a] = c.pppm.on ? c.pppm.sz[a] : 0;
    }
    p.rlist = c.rlist;
    p.skin = c.skin;
    p.qqr = c.pppm.on ? c.pppm.qqr : 0.0;
    return p;
}

void box_upload(Ctx &c) {
    MDX_REQUIRE(!c.capture, 3, "the box cannot change while a step is captured");
    if (!c.box) c.box = dalloc<BoxDev>(c, 1);
    if (!c.boxLast) c.boxLast = dalloc<BoxDev>(c, 1);
    BoxDev &b = c.boxHost;
    std::memset(&b, 0, sizeof(BoxDev));
    box_derive(b, box_params(c), c.geo.lo, c.geo.L);
    b.alpha = c.pppm.on && c.pppm.alpha > 0 ? c.pppm.alpha : (double)c.alpha;
    b.alphaF = c.alpha;
    if (c.pppm.on && c.pppm.alpha > 0) {
        const double hm = std::pow(b.mesh.h[0] * b.mesh.h[1] * b.mesh.h[2], 1.0 / 3.0);
        b.greenLim = green_lim(c.pppm.alpha, hm);
    }
    b.mu = 1.0;
    b.pressure = c.baro.lastP;
    b.regrid = 0;
    MDX_HIP(hipMemcpyAsync(c.box, &b, sizeof(BoxDev), hipMemcpyHostToDevice, c.stream));
    // the copy reads c.boxHost when the stream reaches it
    MDX_HIP(hipStreamSynchronize(c.stream));
}

void box_pull(Ctx &c) {
    if (!c.box) return;
    MDX_HIP(hipMemcpyAsync(&c.boxHost, c.box, sizeof(BoxDev), hipMemcpyDeviceToHost, c.stream));
    MDX_HIP(hipStreamSynchronize(c.stream));
    const BoxDev &b = c.boxHost;
    for (int a = 0; a < 3; ++a) {
        c.geo.lo[a] = b.geo.lo[a];
        c.geo.L[a] = b.geo.L[a];
        c.geo.invL[a] = b.geo.invL[a];
    }
    c.alpha = b.alphaF;
    if (c.pppm.on && c.pppm.alpha > 0) {
        // alpha and G were refreshed on the device for this box
        c.pppm.alpha = b.alpha;
        for (int a = 0; a < 3; ++a) c.pppm.boxAtGreen[a] = b.geo.L[a];
    }
    c.baro.lastP = b.pressure;
}

// Kernels of one barostat application (gate: nullptr = always).  The forces
// of the step must have been computed with the virial.  With c.fuseMolVirial
// the rigid second half already summed the molecular terms into ACC_MOL.
void baro_step(Ctx &c, const int *gate, bool rescale) {
    const BarostatState &B = c.baro;
    // ACC_MOL starts at zero: every force evaluation clears the accumulators
    if (!c.fuseMolVirial) molecular_virial(c, ACC_MOL, gate);
    BaroArgs ba;
    ba.P0 = B.P0;
    ba.period = B.period;
    ba.bulk = B.bulk;
    ba.maxDil = B.maxDil;
    ba.nktv2p = B.nktv2p;
    ba.mvv = B.mvv;
    ba.dt = c.dt;
    ba.alphaSolve = c.pppm.on && c.pppm.alphaUser <= 0 ? 1 : 0;
    ba.natoms = c.n;
    ba.order = c.pppm.order;
    ba.q2 = c.pppm.totalQ2;
    ba.rcut = c.pppm.rcut;
    for (int a = 0; a < 3; ++a) ba.minL[a] = c.nc[a] * (c.rlist / c.reach[a] * (1.0 + 1e-9));
    ba.rescaled = c.fuseMolVirial && c.thermoOn ? 1 : 0;
    ba.tmode = c.thermoTau > 0 ? 1 : 0;
    ba.tgate = thermo_gate(c);
    ba.thermo = c.thermo;
    ba.tconv = c.tconv;
    ba.T0 = c.thermoT;
    ba.coupling = c.thermoTau > 0 ? c.thermoEvery * c.dt / c.thermoTau : 1.0;
    launch<baro_body, 64>(c, 1, 0, gate, c.box, box_params(c), ba, c.acc);
    // The dilation rides on the Green function's launch where it can, and
    // so does the thermostat's velocity rescale of the same step (rescale:
    // its m v^2 was summed by the second half kick; the pressure above
    // already scaled M V^2 by the rescale's lambda^2, and the dilation reads
    // positions only, so the velocities end as with the rescale first).
    DilateArgs dl = dilate_args(c);
    if (rescale) {
        dl.vblocks = cdiv(c.n, 256);
        dl.tgate = thermo_gate(c);
        dl.vel = c.vel[c.cur];
        dl.thermo = c.thermo;
        dl.tconv = c.tconv;
        dl.T0 = c.thermoT;
        dl.tmode = c.thermoTau > 0 ? 1 : 0;
        dl.coupling = c.thermoTau > 0 ? c.thermoEvery * c.dt / c.thermoTau : 1.0;
    }
    if (!(c.pppm.on && pppm_green_device(c, gate, &dl))) {
        dilate(c, gate);
        if (rescale) thermostat_apply(c);
    }
}

}  // namespace mdx
