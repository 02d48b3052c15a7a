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

// one wave: the barostat step (baro_wave, shared with the rigid second
// half, which can run it in its last workgroup)
__device__ __forceinline__ void baro_body(int blk, int nblk, const int *__restrict__ gate, BoxDev *__restrict__ bx, BoxParams bp,
                                          BaroArgs ba, const unsigned long long *__restrict__ acc) {
    baro_wave(gate, bx, bp, ba, acc);
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
BaroArgs baro_args(const Ctx &c) {
    const BarostatState &B = c.baro;
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
    return ba;
}

void baro_step(Ctx &c, const int *gate, bool rescale) {
    if (c.baroFused) {
        c.baroFused = false;  // the rigid second half ran the step in its last workgroup
    } else {
        // ACC_MOL starts at zero: every force evaluation clears the accumulators
        if (!c.fuseMolVirial) molecular_virial(c, ACC_MOL, gate);
        launch<baro_body, 64>(c, 1, 0, gate, c.box, box_params(c), baro_args(c), c.acc);
    }
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
