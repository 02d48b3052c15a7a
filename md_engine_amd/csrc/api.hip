// C ABI of libmdx: argument checking, host<->device marshalling by atom id and
// the step loop.  See include/mdx.h for the contract of each entry point.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <set>

#include "../../include/mdx.h"

#include "mdx_internal.h"

struct mdx_ctx {
    mdx::Ctx c;
};

namespace mdx {

void dfree(Ctx &c, void *p) {
    if (!p) return;
    MDX_REQUIRE(!c.noAlloc, 3, "this call would free device memory");
    ++c.layoutEpoch;
    auto it = std::find(c.allocs.begin(), c.allocs.end(), p);
    if (it != c.allocs.end()) {
        hipFree(p);
        c.allocs.erase(it);
    }
}

static hipEvent_t pool_event(Ctx &c) {
    if (!c.eventPool.empty()) {
        hipEvent_t e = c.eventPool.back();
        c.eventPool.pop_back();
        return e;
    }
    hipEvent_t e;
    MDX_HIP(hipEventCreate(&e));
    return e;
}

ProfScope::ProfScope(Ctx &c_, int k_) : c(c_), k(k_) {
    if (!c.prof) return;
    a = pool_event(c);
    b = pool_event(c);
    hipEventRecord(a, c.stream);
}

ProfScope::~ProfScope() {
    if (!a) return;
    hipEventRecord(b, c.stream);
    c.profRecs.push_back({k, a, b});
}

static void prof_clear(Ctx &c) {
    for (auto &r : c.profRecs) {
        c.eventPool.push_back(r.a);
        c.eventPool.push_back(r.b);
    }
    c.profRecs.clear();
}

namespace {

// Runs f with errors turned into return codes.  Every call through it may
// change the set-up, so cached step graphs of the context are retired;
// read-only entry points use guarded_ro.
template <class F>
int guarded_ro(mdx_ctx *ctx, F &&f) {
    if (!ctx) return MDX_ERR_ARG;
    try {
        f(ctx->c);
        ctx->c.err.clear();
        return MDX_OK;
    } catch (const Error &e) {
        ctx->c.err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        ctx->c.err = e.what();
        return MDX_ERR_INTERNAL;
    }
}

template <class F>
int guarded(mdx_ctx *ctx, F &&f) {
    if (!ctx) return MDX_ERR_ARG;
    ++ctx->c.apiEpoch;
    try {
        f(ctx->c);
        ctx->c.err.clear();
        return MDX_OK;
    } catch (const Error &e) {
        ctx->c.err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        ctx->c.err = e.what();
        return MDX_ERR_INTERNAL;
    }
}

void set_geometry(Ctx &c, const double lo[3], const double len[3]) {
    for (int a = 0; a < 3; ++a) {
        MDX_REQUIRE(len[a] > 0.0, MDX_ERR_ARG, "box lengths must be positive");
        c.geo.lo[a] = lo[a];
        c.geo.L[a] = len[a];
        c.geo.invL[a] = 1.0 / len[a];
    }
    if (c.n > 0) box_upload(c);
}

void require_system(const Ctx &c) { MDX_REQUIRE(c.n > 0, MDX_ERR_STATE, "no system set"); }

void ensure_list(Ctx &c) {
    if (!c.nlistConfigured) nlist_setup(c);
    if (c.centroidList && !c.cenById) c.listValid = false;  // a ring's centroid list, the ring gone
    if (!c.listValid) {
        nlist_rebuild(c, true);
    } else {
        nlist_check(c);
        nlist_rebuild(c, false);
    }
}

double degrees_of_freedom(const Ctx &c) {
    // massive sites x 3, minus centre-of-mass motion, minus 3 per rigid molecule
    double massive = 0;
    for (double im : c.invMassHost) massive += im > 0 ? 1 : 0;
    double ndf = 3.0 * massive - 3.0;
    if (c.rigid.on) ndf -= 3.0 * c.rigid.nmol;
    return ndf > 0 ? ndf : 1.0;
}

// E3B oxygens without charge are found in the uncharged part of the rows
void update_ox_flag(Ctx &c) {
    bool u = true;
    for (int o : c.e3b.oxIds) u = u && c.qHost[o] == 0.0f;
    c.e3b.oxUncharged = u;
}

static_assert(kFlagVirialTrace == MDX_VIRIAL_TRACE, "trace flag of the C ABI");
constexpr int kForceTerms = MDX_PAIR | MDX_PPPM | MDX_E3B | MDX_SURROGATE | MDX_BONDED;

// Instantaneous pressure (nktv_to_press units) from the virial of the last
// compute with MDX_VIRIAL: P = (sum M V^2 + W - corr) / 3V, W the atomic virial
// trace, corr the molecular correction (rigid molecules move as centres of
// mass; free sites count as their own molecules).  Synchronises.
double pressure_now(Ctx &c) {
    acc_clear(c, ACC_MOL, 2);
    molecular_virial(c, ACC_MOL, nullptr);
    double h[ACC_N];
    acc_read(c, h);
    auto v = [&](int k) { return h[k]; };
    const double W = v(ACC_VIR) + v(ACC_VIR + 1) + v(ACC_VIR + 2) + v(ACC_VTRACE);
    const double vol = c.geo.L[0] * c.geo.L[1] * c.geo.L[2];
    return (v(ACC_MOL + 1) * c.baro.mvv + W - v(ACC_MOL)) / (3.0 * vol) * c.baro.nktv2p;
}

// FixPressureBerendsen::stepFinal (src/Fixes/FixPressureBerendsen.cu): mu =
// (1 - dt/period (P0 - P)/B)^(1/3) clamped to 1 -+ maxDilation, applied to
// the box about its centre and to the centres of mass; the cell grid is set
// up again when its cell count would change, P3M refreshes alpha and G.
bool regrid_if_needed(Ctx &c);

void barostat_apply(Ctx &c) {
    acc_clear(c, ACC_MOL, 2);  // the captured step relies on the per-step clear instead
    baro_step(c, nullptr);
    box_pull(c);
    regrid_if_needed(c);
}

// After the box changed on the device: set the grid up again when a cell edge
// fell below r_list / reach (the list would miss pairs) or the box grew by a
// whole cell (finer cells are cheaper).
bool regrid_if_needed(Ctx &c) {
    bool regrid = c.boxHost.regrid != 0;
    for (int d = 0; d < 3; ++d) {
        if (c.geo.periodic[d]) MDX_REQUIRE(c.rlist < 0.5 * c.geo.L[d], 1, "box shrank below twice the list cutoff");
        const double edge = c.rlist / c.reach[d] * (1.0 + 1e-9);
        regrid = regrid || std::max(1, (int)std::floor(c.geo.L[d] * grid_headroom(c) / edge)) > c.nc[d];
    }
    if (regrid) {
        c.nlistConfigured = false;
        c.listValid = false;
    }
    return regrid;
}

void compute(Ctx &c, int flags) {
    const bool energy = flags & MDX_ENERGY;
    // a trace-only request still gets the full tensor from the terms that
    // have no trace mode (P3M, flexible water, surrogate)
    const bool virial = (flags & MDX_VIRIAL) || (flags & kFlagVirialTrace);
    c.vtrace = (flags & kFlagVirialTrace) && !(flags & MDX_VIRIAL) && !energy;
    if (c.flex.on) flex_place_m(c);  // the M sites follow O and H before any force term
    if (c.refreshed)
        c.refreshed = false;  // the first half kick of this captured step did the refresh
    else
        prepare_forces(c);
    const bool mesh = (flags & MDX_PPPM) && c.pppm.on;
    // In a captured step the P3M chain (mostly latency-bound kernels) runs
    // on the side stream beside the pair and E3B kernels: all three start
    // from the refreshed positions and add into the fixed-point force and
    // accumulator buffers, whose integer sums do not depend on the order.
    const bool fork = mesh && c.capture && !c.rec && c.side && c.meshStream;
    const bool forked = c.forked;  // recorded by this step's rebuild after its copy-back
    c.forked = false;
    if (fork) {
        if (!forked) MDX_HIP(hipEventRecord(c.evFork, c.stream));
        MDX_HIP(hipStreamWaitEvent(c.side, c.evFork, 0));
        std::swap(c.stream, c.side);
        try {
            pppm_step(c, energy, virial);
        } catch (...) {
            std::swap(c.stream, c.side);
            throw;
        }
        std::swap(c.stream, c.side);
        MDX_HIP(hipEventRecord(c.evJoin, c.side));
    }
    if ((flags & MDX_PAIR)) pair_forces(c, energy, c.vtrace ? false : virial);
    row_max_flush(c);  // a captured build's row maxima when no pair launch took them
    if (mesh && !fork) pppm_step(c, energy, virial);
    if ((flags & MDX_SURROGATE) && c.e3b.on) {
        MDX_REQUIRE(c.sur.on, MDX_ERR_STATE, "MDX_SURROGATE needs a model (mdx_surrogate_set)");
        surrogate_forces(c, energy, virial);  // the two-body term stays analytic, in the same launch
    } else if ((flags & MDX_E3B) && c.e3b.on) {
        e3b_forces(c, energy, c.vtrace ? false : virial);
    }
    if (fork) MDX_HIP(hipStreamWaitEvent(c.stream, c.evJoin, 0));
    if (c.flex.on) {
        if (flags & MDX_BONDED) flex_forces(c, energy, virial);
        if (!c.deferMSpread) flex_spread_m(c);  // every term's M force goes onto O and H before the kick
    }
    c.forceFlags = flags & kForceTerms;
}

void prepare_terms(Ctx &c) {
    if (c.pppm.on) {
        pppm_prepare(c);
        if (c.ewaldReal) c.alpha = (float)c.pppm.alpha;
    }
}

// Device copy of the per-atom state at a chunk boundary of mdx_run, restored
// when a neighbor-row overflow is found at the end of the chunk.
void snapshot_save(Ctx &c) {
    if (!c.snap.pos) {
        c.snap.pos = dalloc<double4>(c, c.n);
        c.snap.vel = dalloc<double4>(c, c.n);
        c.snap.q = dalloc<float>(c, c.n);
        c.snap.type = dalloc<int>(c, c.n);
        c.snap.id = dalloc<int>(c, c.n);
        c.snap.idToIdx = dalloc<int>(c, c.n);
    }
    const int k = c.cur;
    MDX_HIP(hipMemcpyAsync(c.snap.pos, c.pos[k], sizeof(double4) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.snap.vel, c.vel[k], sizeof(double4) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.snap.q, c.q[k], sizeof(float) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.snap.type, c.type[k], sizeof(int) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.snap.id, c.id[k], sizeof(int) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.snap.idToIdx, c.idToIdx, sizeof(int) * c.n, hipMemcpyDeviceToDevice, c.stream));
    c.snap.turn = c.turn;
    c.snap.rebuilds = c.nRebuilds;
    c.snap.geo = c.geo;
    // A list carried into the chunk was built earlier, from other positions:
    // its rows (their order sets the FP32 summation order of the pair sums)
    // and the skin check's reference positions are the build's.  Keep the
    // positions and box it was built from, so a replay starts from the same list.
    c.snap.listValid = c.listValid && c.nlistConfigured && c.posLast && c.box && c.boxLast;
    if (c.snap.listValid) {
        if (!c.snap.posLast) c.snap.posLast = dalloc<double4>(c, c.n);
        if (!c.snap.boxLast) c.snap.boxLast = dalloc<BoxDev>(c, 1);
        MDX_HIP(hipMemcpyAsync(c.snap.posLast, c.posLast, sizeof(double4) * c.n, hipMemcpyDeviceToDevice,
                               c.stream));
        MDX_HIP(hipMemcpyAsync(c.snap.boxLast, c.boxLast, sizeof(BoxDev), hipMemcpyDeviceToDevice, c.stream));
    }
    if (c.box) {
        if (!c.snap.box) c.snap.box = dalloc<BoxDev>(c, 1);
        MDX_HIP(hipMemcpyAsync(c.snap.box, c.box, sizeof(BoxDev), hipMemcpyDeviceToDevice, c.stream));
    }
    // the thermostat state advances with the particles: a replayed chunk must
    // start from the chain and the kinetic-energy sums of its first step
    if (c.nh.on && c.nh.dev) {
        const int len = 4 * c.nh.chain + 2;
        if (c.snap.nhLen != len) {
            if (c.snap.nh) dfree(c, c.snap.nh);
            c.snap.nh = dalloc<double>(c, len);
            c.snap.nhLen = len;
        }
        MDX_HIP(hipMemcpyAsync(c.snap.nh, c.nh.dev, sizeof(double) * len, hipMemcpyDeviceToDevice, c.stream));
    }
    if (!c.snap.thermo) c.snap.thermo = dalloc<unsigned long long>(c, 2);
    MDX_HIP(hipMemcpyAsync(c.snap.thermo, c.thermo, sizeof(unsigned long long) * 2, hipMemcpyDeviceToDevice,
                           c.stream));
    if (c.pppm.on && c.pppm.interval > 1 && c.pppm.stored) {
        if (!c.snap.pppmStored) c.snap.pppmStored = dalloc<long long>(c, 3 * (size_t)c.n);
        MDX_HIP(hipMemcpyAsync(c.snap.pppmStored, c.pppm.stored, sizeof(long long) * 3 * c.n,
                               hipMemcpyDeviceToDevice, c.stream));
        c.snap.pppmStoredValid = c.pppm.storedValid;
    }
}

void snapshot_restore(Ctx &c) {
    const int k = c.cur;
    MDX_HIP(hipMemcpyAsync(c.pos[k], c.snap.pos, sizeof(double4) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.vel[k], c.snap.vel, sizeof(double4) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.q[k], c.snap.q, sizeof(float) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.type[k], c.snap.type, sizeof(int) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.id[k], c.snap.id, sizeof(int) * c.n, hipMemcpyDeviceToDevice, c.stream));
    MDX_HIP(hipMemcpyAsync(c.idToIdx, c.snap.idToIdx, sizeof(int) * c.n, hipMemcpyDeviceToDevice, c.stream));
    if (c.nh.on && c.nh.dev && c.snap.nh && c.snap.nhLen == 4 * c.nh.chain + 2)
        MDX_HIP(hipMemcpyAsync(c.nh.dev, c.snap.nh, sizeof(double) * c.snap.nhLen, hipMemcpyDeviceToDevice,
                               c.stream));
    if (c.snap.thermo)
        MDX_HIP(hipMemcpyAsync(c.thermo, c.snap.thermo, sizeof(unsigned long long) * 2, hipMemcpyDeviceToDevice,
                               c.stream));
    if (c.pppm.on && c.pppm.interval > 1 && c.pppm.stored && c.snap.pppmStored) {
        MDX_HIP(hipMemcpyAsync(c.pppm.stored, c.snap.pppmStored, sizeof(long long) * 3 * c.n,
                               hipMemcpyDeviceToDevice, c.stream));
        c.pppm.storedValid = c.snap.pppmStoredValid;
    }
    c.turn = c.snap.turn;
    c.listValid = false;
    const bool relist = c.snap.listValid && c.nlistConfigured;
    if (relist) {
        // Build the chunk's starting list again from the positions and box of
        // its build: the records are already in that build's order, so the
        // permutation is the identity and the rows come out as they were.
        // Then the chunk-start positions and box go back in.
        MDX_HIP(hipMemcpyAsync(c.pos[c.cur], c.snap.posLast, sizeof(double4) * c.n, hipMemcpyDeviceToDevice,
                               c.stream));
        MDX_HIP(hipMemcpyAsync(c.box, c.snap.boxLast, sizeof(BoxDev), hipMemcpyDeviceToDevice, c.stream));
        nlist_rebuild(c, true);
        MDX_HIP(hipMemcpyAsync(c.pos[c.cur], c.snap.pos, sizeof(double4) * c.n, hipMemcpyDeviceToDevice,
                               c.stream));
        MDX_HIP(hipMemcpyAsync(c.box, c.snap.box, sizeof(BoxDev), hipMemcpyDeviceToDevice, c.stream));
        box_pull(c);
        refresh_xq(c);
    }
    c.nRebuilds = c.snap.rebuilds;
    bool moved = false;
    for (int d = 0; d < 3; ++d) moved = moved || c.snap.geo.L[d] != c.geo.L[d] || c.snap.geo.lo[d] != c.geo.lo[d];
    if (c.baro.on && c.snap.box) {
        // the barostat moved the device box inside the chunk (the host mirror
        // may not have seen it): the box, alpha and G go back too
        MDX_HIP(hipMemcpyAsync(c.box, c.snap.box, sizeof(BoxDev), hipMemcpyDeviceToDevice, c.stream));
        box_pull(c);
        if (c.pppm.on) pppm_green_device(c, nullptr);
        moved = true;
    }
    if (moved) {  // a barostat changed the box during the replayed chunk
        c.geo = c.snap.geo;
        // the grid that built the starting list was valid for the starting
        // box; without a carried list the grid is set up again
        if (!relist) c.nlistConfigured = false;
    }
}

// Reads the dMax slot written one step earlier (waits only for that step's
// check, while later work keeps the GPU busy).  Returns the rebuild decision
// for the current step and counts a violation when the list of the step
// that was checked was already older than the skin allows.
bool lagged_rebuild_decision(Ctx &c, int parity) {
    MDX_HIP(hipEventSynchronize(c.stepEvent[parity]));
    float d2, p2;
    std::memcpy(&d2, &c.pinned[2 + 2 * parity], sizeof(float));
    std::memcpy(&p2, &c.pinned[3 + 2 * parity], sizeof(float));
    const double half = 0.5 * c.skin;
    if (std::sqrt((double)d2) > half) c.listViolations++;
    return std::sqrt((double)p2) > half;
}

void check_and_copy(Ctx &c, int parity) {
    check_max(c, parity);
    MDX_HIP(hipMemcpyAsync(c.pinned + 2 + 2 * parity, c.dMax + 2 * parity, sizeof(unsigned) * 2,
                           hipMemcpyDeviceToHost, c.stream));
    MDX_HIP(hipEventRecord(c.stepEvent[parity], c.stream));
}

template <class T>
void upload(Ctx &c, T *dst, const T *src, size_t count) {
    MDX_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyHostToDevice, c.stream));
}

// ---------------------------------------------------------------------------
// Captured steps.  One MD step (kick-drift + constraints, device rebuild
// decision, gated list rebuild, forces, kick + constraints, gated thermostat)
// is recorded G times into one hipGraph for every context of a batch: each
// context's work stays on its own stream, forked from and joined to the first
// context's stream inside the capture, so one graph launch advances every
// replica by G steps.  Buffers are fixed while a graph lives; row overflow is
// found at the end of a chunk and the chunk is replayed with grown rows.
constexpr int kGraphSteps = 8;
constexpr int kGraphChunk = 64;

void captured_step(Ctx &c, int fflags) {
    // all-water systems without a Nose-Hoover chain: the rigid second half
    // sums the molecular virial terms and the thermostat's m v^2 (no separate
    // passes over the molecules; k_decide clears the m v^2 sum)
    const bool allRigid = c.rigid.on && c.rigid.nsites == c.n && !c.nh.on;
    c.fuseMolVirial = c.baro.on && allRigid;
    c.fuseThermo = c.thermoOn && allRigid;
    c.fuseCheck = c.rigid.on && c.rigid.nsites == c.n;
    integrate_first_half(c);
    decide_step(c);
    nlist_rebuild_async(c);  // every kernel gated by dFlags[3]
    // a barostat needs the virial; its steps are gated by dFlags[8]
    compute(c, fflags | (c.baro.on ? kFlagVirialTrace : 0));
    integrate_second_half(c);
    // with the barostat on, the rescale (gated by dFlags[5]) rides on its launch
    const bool rescaleWithBaro = c.thermoOn && c.fuseThermo && c.baro.on && c.fuseStep;
    if (c.thermoOn && !rescaleWithBaro) thermostat_apply(c);  // gated by dFlags[5]
    if (c.baro.on) baro_step(c, baro_gate(c), rescaleWithBaro);
    c.fuseMolVirial = false;
    c.fuseThermo = false;
    c.fuseCheck = false;
}

// How the operations of a captured step are issued.  each(): every context
// runs the same operations (one phase); one(): a single context runs an
// operation that serves them all (a ring-level launch).  Issued directly,
// the launches go to the contexts' streams as they are made; recorded, a
// phase becomes one launch per kernel across the contexts.
struct StepOps {
    virtual ~StepOps() = default;
    virtual void each(const std::vector<Ctx *> &cs, const std::function<void(Ctx &)> &f) = 0;
    virtual void one(Ctx &c, const std::function<void()> &f) = 0;
};

struct DirectOps : StepOps {
    void each(const std::vector<Ctx *> &cs, const std::function<void(Ctx &)> &f) override {
        for (Ctx *c : cs) f(*c);
    }
    void one(Ctx &, const std::function<void()> &f) override { f(); }
};

// Ring-polymer forces of every bead: per-bead terms, then (contraction) the
// mesh part once from the centroids onto every bead, then the M-site forces
// of flexible water onto O and H.
bool ring_contracts(const RingSpec &ring, int fflags) { return ring.contract && (fflags & MDX_PPPM); }

void ring_forces(const std::vector<Ctx *> &cs, const RingSpec &ring, int fflags, StepOps &ops) {
    const bool contract = ring_contracts(ring, fflags);
    ops.each(cs, [&](Ctx &c) {
        c.deferMSpread = contract;
        compute(c, contract ? (fflags & ~MDX_PPPM) : fflags);
        c.deferMSpread = false;
    });
    if (contract) {
        ops.one(*cs[0], [&] { ring_pppm(*cs[0], ring); });
        ops.each(cs, [&](Ctx &c) {
            if (c.flex.on) flex_spread_m(c);
        });
    }
    for (Ctx *cp : cs) cp->forceFlags = fflags & kForceTerms;
}

void ring_forces(const std::vector<Ctx *> &cs, const RingSpec &ring, int fflags) {
    DirectOps direct;
    ring_forces(cs, ring, fflags, direct);
}

// One step of a ring polymer whose beads share one stream: every bead's half
// kick, the free ring-polymer propagation across beads, then every bead's
// rebuild gate, forces, second half kick and thermostat.
void captured_ring_step(const std::vector<Ctx *> &cs, const RingSpec &ring, int fflags, StepOps &ops) {
    ops.each(cs, [](Ctx &c) { integrate_kick_first(c); });
    ops.one(*cs[0], [&] { ring_drift(ring, *cs[0]); });
    if (ring.centroidList) {
        // one decision for the whole ring from every bead's displacement
        // (GridGPU::setBuildFlag over all P n atoms), then the centroids of
        // a rebuild step, then every bead's build on them
        ops.each(cs, [](Ctx &c) { skin_check(c); });
        ops.each(cs, [](Ctx &c) { skin_decide(c); });
        ops.one(*cs[0], [&] { ring_centroids(ring, *cs[0], rebuild_gate(*cs[0])); });
    }
    if (ring_contracts(ring, fflags)) {
        ops.each(cs, [&](Ctx &c) {
            if (!ring.centroidList) decide_step(c);
            nlist_rebuild_async(c);
        });
        ring_forces(cs, ring, fflags, ops);
        ops.each(cs, [](Ctx &c) {
            integrate_second_half(c);
            if (c.thermoOn) thermostat_apply(c);  // gated by dFlags[5]
        });
        return;
    }
    ops.each(cs, [&](Ctx &c) {
        if (!ring.centroidList) decide_step(c);
        nlist_rebuild_async(c);
        compute(c, fflags);
        integrate_second_half(c);
        if (c.thermoOn) thermostat_apply(c);
    });
}

// Replicas of a merged batch whose P3M meshes have the same dimensions share
// one batched transform per direction: their work arrays (real mesh, half
// spectrum, field spectra, real field) are laid out back to back in one
// allocation while the batch is recorded and captured, so one plan of batch
// m (forward) and 3 m (inverse) covers all of them.  The replicas' own arrays
// come back afterwards; the graph keeps the shared ones.
struct FftGroup {
    std::vector<size_t> members;  // replica indices
    hipfftHandle r2c = 0, c2r = 0;
};

struct SavedMesh {
    float *real;
    float2 *rhoHat, *eHat;
    float *eReal;
};

std::vector<FftGroup> group_ffts(const std::vector<Ctx *> &cs, BatchKeep &keep, std::vector<SavedMesh> &saved,
                                 std::vector<int> &groupOf) {
    const size_t R = cs.size();
    std::vector<FftGroup> groups;
    groupOf.assign(R, -1);
    saved.resize(R);
    for (size_t k = 0; k < R; ++k) {
        const PppmState &P = cs[k]->pppm;
        saved[k] = {P.real, P.rhoHat, P.eHat, P.eReal};
        if (!P.on || !P.plansMade || pppm_lds_fft(P) || groupOf[k] >= 0) continue;
        FftGroup g;
        for (size_t j = k; j < R; ++j) {
            const PppmState &Q = cs[j]->pppm;
            if (groupOf[j] < 0 && Q.on && Q.plansMade && !pppm_lds_fft(Q) && Q.sz[0] == P.sz[0] && Q.sz[1] == P.sz[1] &&
                Q.sz[2] == P.sz[2]) {
                groupOf[j] = (int)groups.size();
                g.members.push_back(j);
            }
        }
        groups.push_back(g);
    }
    for (FftGroup &g : groups) {
        const int *sz = cs[g.members[0]]->pppm.sz;
        const int m = (int)g.members.size();
        const size_t ng = (size_t)sz[0] * sz[1] * sz[2], nh = (size_t)sz[0] * sz[1] * (sz[2] / 2 + 1);
        void *real = nullptr, *rho = nullptr, *eh = nullptr, *er = nullptr;
        MDX_HIP(hipMalloc(&real, sizeof(float) * ng * m));
        keep.mem.push_back(real);
        MDX_HIP(hipMalloc(&rho, sizeof(float2) * nh * m));
        keep.mem.push_back(rho);
        MDX_HIP(hipMalloc(&eh, sizeof(float2) * 3 * nh * m));
        keep.mem.push_back(eh);
        MDX_HIP(hipMalloc(&er, sizeof(float) * 3 * ng * m));
        keep.mem.push_back(er);
        for (int t = 0; t < m; ++t) {
            PppmState &P = cs[g.members[t]]->pppm;
            P.real = static_cast<float *>(real) + ng * t;
            P.rhoHat = static_cast<float2 *>(rho) + nh * t;
            P.eHat = static_cast<float2 *>(eh) + 3 * nh * t;
            P.eReal = static_cast<float *>(er) + 3 * ng * t;
        }
        int dims[3] = {sz[0], sz[1], sz[2]};
        MDX_REQUIRE(hipfftPlanMany(&g.r2c, 3, dims, nullptr, 1, (int)ng, nullptr, 1, (int)nh, HIPFFT_R2C, m) ==
                        HIPFFT_SUCCESS,
                    2, "hipfftPlanMany R2C (batch) failed");
        keep.plans.push_back(g.r2c);
        MDX_REQUIRE(hipfftPlanMany(&g.c2r, 3, dims, nullptr, 1, (int)nh, nullptr, 1, (int)ng, HIPFFT_C2R, 3 * m) ==
                        HIPFFT_SUCCESS,
                    2, "hipfftPlanMany C2R (batch) failed");
        keep.plans.push_back(g.c2r);
        MDX_REQUIRE(hipfftSetStream(g.r2c, cs[0]->stream) == HIPFFT_SUCCESS &&
                        hipfftSetStream(g.c2r, cs[0]->stream) == HIPFFT_SUCCESS,
                    2, "hipfftSetStream failed");
    }
    return groups;
}

void restore_meshes(const std::vector<Ctx *> &cs, const std::vector<SavedMesh> &saved) {
    for (size_t k = 0; k < cs.size() && k < saved.size(); ++k) {
        PppmState &P = cs[k]->pppm;
        P.real = saved[k].real;
        P.rhoHat = saved[k].rhoHat;
        P.eHat = saved[k].eHat;
        P.eReal = saved[k].eReal;
    }
}

// Records the operations of captured steps instead of issuing them.  While a
// context records, a stream that is being captured into a throw-away graph
// stands in for its stream, so a launch that bypassed the launch layer is
// caught (it lands in that graph) instead of running.
struct RecordOps : StepOps {
    struct Item {
        bool phase = false;                        // each() (else one())
        bool merged = false;                       // every context recorded the same kernels
        std::vector<std::vector<LaunchRec>> recs;  // per context (one() items: one)
    };
    std::vector<Item> items;
    hipStream_t dummy = nullptr;
    bool stray = false;

    RecordOps() { MDX_HIP(hipStreamCreateWithFlags(&dummy, hipStreamNonBlocking)); }
    ~RecordOps() { hipStreamDestroy(dummy); }

    void record(Ctx &c, std::vector<LaunchRec> &out, const std::function<void()> &f) {
        const hipStream_t own = c.stream;
        const bool prof = c.prof, cap = c.capture;
        hipGraph_t g = nullptr;
        MDX_HIP(hipStreamBeginCapture(dummy, hipStreamCaptureModeThreadLocal));
        c.stream = dummy;
        c.prof = false;
        c.capture = true;
        c.rec = &out;
        auto undo = [&] {
            c.rec = nullptr;
            c.capture = cap;
            c.prof = prof;
            c.stream = own;
        };
        try {
            f();
        } catch (...) {
            undo();
            hipStreamEndCapture(dummy, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        undo();
        MDX_HIP(hipStreamEndCapture(dummy, &g));
        size_t nodes = 0;
        if (g) {
            MDX_HIP(hipGraphGetNodes(g, nullptr, &nodes));
            MDX_HIP(hipGraphDestroy(g));
        }
        stray = stray || nodes > 0;
    }

    void each(const std::vector<Ctx *> &cs, const std::function<void(Ctx &)> &f) override {
        Item it;
        it.phase = true;
        it.recs.resize(cs.size());
        for (size_t k = 0; k < cs.size(); ++k) record(*cs[k], it.recs[k], [&] { f(*cs[k]); });
        it.merged = true;
        for (size_t k = 1; k < cs.size() && it.merged; ++k) {
            it.merged = it.recs[k].size() == it.recs[0].size();
            for (size_t i = 0; it.merged && i < it.recs[0].size(); ++i)
                it.merged = it.recs[k][i].kid == it.recs[0][i].kid &&
                            (it.recs[k][i].kid != nullptr || it.recs[k][i].host) &&
                            it.recs[k][i].fft == it.recs[0][i].fft &&
                            it.recs[k][i].args.size() == it.recs[0][i].args.size();
        }
        items.push_back(std::move(it));
    }

    void one(Ctx &c, const std::function<void()> &f) override {
        Item it;
        it.recs.resize(1);
        record(c, it.recs[0], f);
        items.push_back(std::move(it));
    }
};

// Captured steps of several contexts with one launch per kernel and phase.
// The steps are recorded first (RecordOps); each kernel of a merged phase
// becomes one launch over the concatenated grids of its contexts: an
// argument table (one packed argument set per context) and a block map
// ((context, block, blocks) per workgroup) in device memory, kept with the
// graph; single operations and unmerged phases become launches of a
// one-entry table, in recorded order.  P3M transforms of merged phases run
// once per group of equal meshes (group_ffts).  Returns nullptr when a launch
// bypassed the layer, or when a batch's replicas recorded different kernels
// (the caller then captures them as separate streams).
hipGraphExec_t capture_merged(const std::vector<Ctx *> &cs, int steps, int fflags, BatchKeep &keep,
                              const RingSpec *ring) {
    Ctx &c0 = *cs[0];
    const size_t R = cs.size();
    std::vector<SavedMesh> saved;
    std::vector<int> groupOf;
    std::vector<FftGroup> groups;
    try {
        groups = group_ffts(cs, keep, saved, groupOf);
    } catch (...) {
        restore_meshes(cs, saved);
        throw;
    }
    struct Restore {
        const std::vector<Ctx *> &cs;
        const std::vector<SavedMesh> &saved;
        ~Restore() { restore_meshes(cs, saved); }
    } restore{cs, saved};
    RecordOps rec;
    if (ring) {
        for (int s = 0; s < steps; ++s) captured_ring_step(cs, *ring, fflags, rec);
    } else {
        rec.each(cs, [&](Ctx &c) {
            for (int s = 0; s < steps; ++s) captured_step(c, fflags);
        });
    }
    bool allMerged = true;
    for (const auto &it : rec.items)
        if (it.phase) allMerged = allMerged && it.merged;
    c0.batchMerged = !rec.stray && allMerged ? 1 : 0;
    if (rec.stray || (!ring && !allMerged)) return nullptr;

    // every launch: its argument table and block map in one device block
    struct Emit {
        const LaunchRec *op;                   // a kernel (args, launcher) or a host operation
        std::vector<const LaunchRec *> parts;  // the contexts' records it covers (kernels)
        std::vector<const LaunchRec *> hosts;  // host operations to run, in order
        int fftKind = 0;                       // grouped transform of a merged phase
        size_t tab = 0, map = 0;
        int total = 0;
        size_t shm = 0;
    };
    std::vector<Emit> emits;
    for (const auto &it : rec.items) {
        if (it.merged && it.recs.size() == R && it.phase) {
            for (size_t i = 0; i < it.recs[0].size(); ++i) {
                Emit e;
                e.op = &it.recs[0][i];
                if (e.op->kid) {
                    for (size_t k = 0; k < R; ++k) e.parts.push_back(&it.recs[k][i]);
                } else if (e.op->fft) {
                    e.fftKind = e.op->fft;
                    for (size_t k = 0; k < R; ++k) e.hosts.push_back(&it.recs[k][i]);
                } else {
                    for (size_t k = 0; k < R; ++k) e.hosts.push_back(&it.recs[k][i]);
                }
                emits.push_back(std::move(e));
            }
        } else {
            for (const auto &r : it.recs)
                for (const LaunchRec &op : r) {
                    Emit e;
                    e.op = &op;
                    if (op.kid) e.parts.push_back(&op);
                    else e.hosts.push_back(&op);
                    emits.push_back(std::move(e));
                }
        }
    }
    auto align_to = [](size_t v, size_t a) { return (v + a - 1) / a * a; };
    size_t bytes = 0;
    for (Emit &e : emits) {
        if (e.parts.empty()) continue;
        const size_t stride = align_to(e.op->args.size(), e.op->align);
        bytes = align_to(bytes, 256);
        e.tab = bytes;
        bytes += stride * e.parts.size();
        for (const LaunchRec *p : e.parts) {
            e.total += std::max(p->blocks, 0);
            e.shm = std::max(e.shm, p->shm);
        }
        bytes = align_to(bytes, 256);
        e.map = bytes;
        bytes += sizeof(int4) * (size_t)std::max(e.total, 1);
    }
    std::vector<unsigned char> host(std::max<size_t>(bytes, 1), 0);
    for (const Emit &e : emits) {
        if (e.parts.empty()) continue;
        const size_t stride = align_to(e.op->args.size(), e.op->align);
        int4 *map = reinterpret_cast<int4 *>(host.data() + e.map);
        int b = 0;
        for (size_t k = 0; k < e.parts.size(); ++k) {
            const LaunchRec &p = *e.parts[k];
            std::memcpy(host.data() + e.tab + stride * k, p.args.data(), p.args.size());
            for (int j = 0; j < p.blocks; ++j) map[b++] = make_int4((int)k, j, p.blocks, 0);
        }
    }
    void *dev = nullptr;
    MDX_HIP(hipMalloc(&dev, host.size()));
    keep.mem.push_back(dev);
    MDX_HIP(hipMemcpy(dev, host.data(), host.size(), hipMemcpyHostToDevice));
    unsigned char *d = static_cast<unsigned char *>(dev);
    hipGraph_t graph = nullptr;
    MDX_HIP(hipStreamBeginCapture(c0.stream, hipStreamCaptureModeThreadLocal));
    try {
        for (const Emit &e : emits) {
            if (!e.parts.empty()) {
                if (e.total == 0) continue;
                e.op->many(c0.stream, d + e.tab, reinterpret_cast<const int4 *>(d + e.map), e.total, e.shm);
                MDX_HIP(hipGetLastError());
            } else if (e.fftKind) {
                // one transform per group of equal meshes, in context order
                std::vector<char> done(groups.size(), 0);
                for (size_t k = 0; k < e.hosts.size(); ++k) {
                    const int g = groupOf[k];
                    if (g < 0) {
                        e.hosts[k]->host(c0.stream);
                        continue;
                    }
                    if (done[g]) continue;
                    done[g] = 1;
                    const PppmState &P = cs[groups[g].members[0]]->pppm;
                    if (e.fftKind == 1)
                        MDX_REQUIRE(hipfftExecR2C(groups[g].r2c, P.real, (hipfftComplex *)P.rhoHat) == HIPFFT_SUCCESS,
                                    2, "hipfftExecR2C (batch) failed");
                    else
                        MDX_REQUIRE(hipfftExecC2R(groups[g].c2r, (hipfftComplex *)P.eHat, P.eReal) == HIPFFT_SUCCESS,
                                    2, "hipfftExecC2R (batch) failed");
                }
            } else {
                for (const LaunchRec *h : e.hosts) h->host(c0.stream);
            }
        }
    } catch (...) {
        hipStreamEndCapture(c0.stream, &graph);
        if (graph) hipGraphDestroy(graph);
        throw;
    }
    MDX_HIP(hipStreamEndCapture(c0.stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    MDX_HIP(ie);
    return exec;
}

hipGraphExec_t capture_steps(const std::vector<Ctx *> &cs, int steps, int fflags, const RingSpec *ring = nullptr,
                             BatchKeep *keep = nullptr) {
    if (cs.size() > 1 && keep && cs[0]->batchMerge) {
        hipGraphExec_t merged = capture_merged(cs, steps, fflags, *keep, ring);
        if (merged) return merged;
    }
    Ctx &c0 = *cs[0];
    for (Ctx *c : cs)
        if (c->pppm.on && c->meshStream && !c->side) {  // default priority, like the main stream
            MDX_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
            MDX_HIP(hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming));
            MDX_HIP(hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming));
        }
    std::vector<hipEvent_t> evs(cs.size() * 2, nullptr);
    for (auto &e : evs) MDX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    std::vector<bool> prof(cs.size());
    hipGraph_t graph = nullptr;
    MDX_HIP(hipStreamBeginCapture(c0.stream, hipStreamCaptureModeThreadLocal));
    try {
        MDX_HIP(hipEventRecord(evs[0], c0.stream));
        if (ring) {
            // the beads run on c0's stream, interleaved step by step
            for (size_t k = 0; k < cs.size(); ++k) {
                prof[k] = cs[k]->prof;
                cs[k]->prof = false;
                cs[k]->capture = true;
            }
            DirectOps direct;
            for (int s = 0; s < steps; ++s) captured_ring_step(cs, *ring, fflags, direct);
            for (size_t k = 0; k < cs.size(); ++k) {
                cs[k]->capture = false;
                cs[k]->prof = prof[k];
            }
        } else
        for (size_t k = 0; k < cs.size(); ++k) {
            Ctx &c = *cs[k];
            if (k) MDX_HIP(hipStreamWaitEvent(c.stream, evs[0], 0));
            prof[k] = c.prof;
            c.prof = false;
            c.capture = true;
            for (int s = 0; s < steps; ++s) captured_step(c, fflags);
            c.capture = false;
            c.prof = prof[k];
            if (k) {
                MDX_HIP(hipEventRecord(evs[cs.size() + k], c.stream));
                MDX_HIP(hipStreamWaitEvent(c0.stream, evs[cs.size() + k], 0));
            }
        }
    } catch (...) {
        for (size_t k = 0; k < cs.size(); ++k) {
            cs[k]->capture = false;
            if (k < prof.size()) cs[k]->prof = prof[k];
        }
        hipStreamEndCapture(c0.stream, &graph);
        if (graph) hipGraphDestroy(graph);
        for (auto e : evs) hipEventDestroy(e);
        throw;
    }
    MDX_HIP(hipStreamEndCapture(c0.stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    for (auto e : evs) hipEventDestroy(e);
    MDX_HIP(ie);
    return exec;
}

// Waits on `from` in stream `to` (cross-context ordering outside a graph).
void stream_after(hipStream_t to, hipStream_t from) {
    hipEvent_t e;
    MDX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    MDX_HIP(hipEventRecord(e, from));
    MDX_HIP(hipStreamWaitEvent(to, e, 0));
    MDX_HIP(hipEventDestroy(e));
}

// the ring's per-bead pointer table for the current buffers
void ring_table(const std::vector<Ctx *> &cs, const RingSpec &ring) {
    std::vector<void *> t(kRingTab * cs.size());
    for (size_t k = 0; k < cs.size(); ++k) {
        t[kRingTab * k] = cs[k]->pos[cs[k]->cur];
        t[kRingTab * k + 1] = cs[k]->vel[cs[k]->cur];
        t[kRingTab * k + 2] = cs[k]->idToIdx;
        t[kRingTab * k + 3] = cs[k]->force;
    }
    MDX_HIP(hipMemcpyAsync(ring.tab, t.data(), sizeof(void *) * t.size(), hipMemcpyHostToDevice, cs[0]->stream));
    MDX_HIP(hipStreamSynchronize(cs[0]->stream));
}

// Centroid lists outside the captured steps: when any bead lacks a list,
// every bead's list is built now from the centroids of this moment (one list
// for the ring, as the reference builds it).  Returns true if it built.
bool ring_lists(const std::vector<Ctx *> &cs, const RingSpec &ring) {
    if (!ring.centroidList) return false;
    bool any = false;
    for (Ctx *c : cs) any = any || !c->nlistConfigured || !c->listValid || !c->centroidList;
    if (!any) return false;
    for (Ctx *c : cs)
        if (!c->nlistConfigured) nlist_setup(*c);
    ring_table(cs, ring);
    ring_centroids(ring, *cs[0]);
    for (Ctx *c : cs) {
        nlist_rebuild(*c, true);
        c->forceFlags = -1;  // forces again on the new lists
    }
    return true;
}

// The E3B judge of a run with the learned three-body term (one context):
// force terms after a fallback, steps to the next judge point, and the check
// at a judge point.
bool judging(const Ctx &c, int fflags) {
    return (fflags & MDX_SURROGATE) && c.sur.judgeEvery > 0 && c.e3b.on && c.sur.on;
}

int judged_flags(const Ctx &c, int fflags) {
    return (c.sur.fellBack && (fflags & MDX_SURROGATE)) ? ((fflags & ~MDX_SURROGATE) | MDX_E3B) : fflags;
}

int steps_to_judge(const Ctx &c) {
    const long long e = c.sur.judgeEvery;
    return (int)(e - c.turn % e);
}

void judge_if_due(Ctx &c, int fflags) {
    if (!judging(c, fflags) || c.sur.fellBack || c.turn % c.sur.judgeEvery != 0) return;
    const double r = surrogate_judge(c);
    SurrogateState &S = c.sur;
    ++S.judgeChecks;
    S.judgeLast = r;
    S.judgeWorst = std::max(S.judgeWorst, r);
    if (!(r <= S.judgeMax)) S.fellBack = true;
}

void run_captured(const std::vector<Ctx *> &cs, int nsteps, int flags, const RingSpec *ring = nullptr) {
    const int flags0 = flags & kForceTerms;
    int fflags = judged_flags(*cs[0], flags0);
    const bool judge = cs.size() == 1 && !ring && judging(*cs[0], flags0);
    for (Ctx *cp : cs) {
        Ctx &c = *cp;
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        MDX_REQUIRE(c.device == cs[0]->device, MDX_ERR_ARG, "batched contexts must share one device");
        MDX_REQUIRE(!c.pppm.on || c.pppm.interval <= 1, MDX_ERR_STATE,
                    "captured steps evaluate the mesh every step: long-range interval must be 1");
        MDX_HIP(hipSetDevice(c.device));
        prepare_terms(c);
        if (c.rigid.on && c.rigid.fresh) {
            rigid_second_half(c, false);
            c.rigid.fresh = false;
            c.listValid = false;
        }
        if (c.nh.on && c.nh.fresh) {
            c.nh.ndf = degrees_of_freedom(c);
            nose_hoover_start(c);
        }
    }
    // Without a ring, the step graphs stay with the first context after the
    // call: the next call replays them as long as nothing they bake in (the
    // contexts, their set-up, buffers, cell grid, buffer parity, the force
    // terms) changed.  A ring's bead table is made per call.
    Ctx::GraphCache *cache = ring ? nullptr : &cs[0]->graphs;
    hipGraphExec_t big = nullptr, one = nullptr;
    BatchKeep localKeep;
    BatchKeep &keep = cache ? cache->keep : localKeep;
    if (cache) {
        big = cache->big;
        one = cache->one;
    }
    std::vector<int> ringCur;
    auto drop = [&] {
        if (big) hipGraphExecDestroy(big);
        if (one) hipGraphExecDestroy(one);
        big = one = nullptr;
        if (!keep.empty()) {
            // what merged launches use goes with their graphs
            hipStreamSynchronize(cs[0]->stream);
            keep.release();
        }
        if (cache) {
            cache->big = cache->one = nullptr;
            cache->key.clear();
        }
    };
    auto key_now = [&] {
        std::vector<uint64_t> k{(uint64_t)fflags, (uint64_t)cs.size()};
        for (Ctx *cp : cs) {
            k.push_back(cp->serial);
            k.push_back(cp->apiEpoch);
            k.push_back(cp->layoutEpoch);
            k.push_back((uint64_t)cp->cur);
        }
        return k;
    };
    try {
        int done = 0, chunkNo = 0;
        bool forcedDone = false;
        while (done < nsteps) {
            int len = std::min(kGraphChunk, nsteps - done);
            if (judge && !cs[0]->sur.fellBack) len = std::min(len, steps_to_judge(*cs[0]));
            // a contracted ring evaluates every bead's forces together, below
            const bool ringForces = ring && ring_contracts(*ring, fflags);
            bool ringStale = false;
            if (ring) ring_lists(cs, *ring);  // centroid lists: every bead's at once
            for (Ctx *cp : cs) {
                Ctx &c = *cp;
                snapshot_save(c);
                if (!c.listValid || c.forceFlags != fflags) {
                    ensure_list(c);
                    if (ringForces) ringStale = true;
                    else compute(c, fflags);
                }
                MDX_HIP(hipStreamSynchronize(c.stream));
                if (!capacity_ok(c)) {
                    nlist_rebuild(c, true);
                    if (ringForces) ringStale = true;
                    else compute(c, fflags);
                }
                // chunk state: displacement slots, running row maxima, counters, turn
                fill_int(c, (int *)c.dMax, 4, 0);
                fill_int(c, c.dFlags + 1, 5, 0);
                // device turn counter: the thermostat and barostat gates take it
                // modulo their intervals
                const long long period = (long long)(c.thermoOn ? c.thermoEvery : 1) * (c.baro.on ? c.baro.every : 1);
                fill_int(c, c.dFlags + 6, 1, (int)(c.turn % period));
                if (cp != cs[0]) stream_after(cs[0]->stream, c.stream);
            }
            if (ring) {
                // a bead's buffers may have flipped in the set-up above: a new
                // table and new graphs
                std::vector<int> curs;
                for (Ctx *cp : cs) curs.push_back(cp->cur);
                if (curs != ringCur) drop();
                ringCur = curs;
                ring_table(cs, *ring);
                if (ringStale) {
                    ring_forces(cs, *ring, fflags);
                    MDX_HIP(hipStreamSynchronize(cs[0]->stream));
                }
            }
            if (cache) {
                const std::vector<uint64_t> k = key_now();
                if (k != cache->key) {
                    drop();
                    cache->key = k;
                }
            }
            if (!big) big = capture_steps(cs, kGraphSteps, fflags, ring, &keep);
            const int nbig = len / kGraphSteps, nrest = len % kGraphSteps;
            if (nrest && !one) one = capture_steps(cs, 1, fflags, ring, &keep);
            if (cache) {
                cache->big = big;
                cache->one = one;
            }
            for (int k = 0; k < nbig; ++k) MDX_HIP(hipGraphLaunch(big, cs[0]->stream));
            for (int k = 0; k < nrest; ++k) MDX_HIP(hipGraphLaunch(one, cs[0]->stream));
            MDX_HIP(hipStreamSynchronize(cs[0]->stream));
            // Every replica advanced by len steps in the graph.  First read
            // what each one found (row overflow, a grid its box outgrew); only
            // then either replay the chunk for all of them or accept it for
            // all of them, so no replica is left one chunk ahead.
            std::vector<std::array<int, 7>> hs(cs.size());
            std::vector<char> overflow(cs.size(), 0), badGrid(cs.size(), 0);
            bool regrow = false, forcedNow = false;
            for (size_t k = 0; k < cs.size(); ++k) {
                Ctx &c = *cs[k];
                int *h = hs[k].data();
                MDX_HIP(hipMemcpy(h, c.dFlags + 1, sizeof(int) * 7, hipMemcpyDeviceToHost));
                c.pinned[0] = h[0];
                c.pinned[1] = h[1];
                c.pinned[6] = h[6];
                flag7_check(h[6]);
                overflow[k] = h[0] > c.cap || (c.e3b.on && h[1] > c.e3b.cap);
                if (c.baro.on) {
                    box_pull(c);
                    // a cell edge fell below r_list / reach inside the chunk:
                    // replay it on a grid set up for the starting box
                    badGrid[k] = c.boxHost.regrid != 0;
                }
                regrow = regrow || overflow[k] || badGrid[k];
                forcedNow = forcedNow || (!forcedDone && c.replayChunk == chunkNo + 1);
            }
            forcedDone = forcedDone || forcedNow;
            regrow = regrow || forcedNow;
            if (regrow) {
                for (size_t k = 0; k < cs.size(); ++k) {
                    Ctx &c = *cs[k];
                    // rows grow first: the restore builds the starting list into them
                    if (overflow[k]) grow_if_needed(c);
                    snapshot_restore(c);
                    if (badGrid[k]) c.nlistConfigured = false;
                    c.forceFlags = -1;
                }
                drop();
                continue;
            }
            ++chunkNo;
            bool regrid = false;
            for (size_t k = 0; k < cs.size(); ++k) {
                Ctx &c = *cs[k];
                c.nRebuilds += hs[k][3];
                c.turn += len;
                // the box grew by a cell: the next chunk takes a finer grid
                if (c.baro.on && regrid_if_needed(c)) regrid = true;
            }
            if (regrid) drop();
            done += len;
            if (judge) {
                judge_if_due(*cs[0], flags0);
                fflags = judged_flags(*cs[0], flags0);  // a fallback recaptures with E3B
            }
        }
    } catch (...) {
        drop();
        throw;
    }
    if (!cache) drop();
    for (Ctx *cp : cs) MDX_HIP(hipStreamSynchronize(cp->stream));
}

}  // namespace
}  // namespace mdx

namespace mdx {
namespace {

// Beads of one ring (validated), their ring description on the first bead's
// device and stream, and the work `f` between set-up and tear-down.  Every
// bead runs on the first bead's stream for the step-by-step coupling.
// Contraction (unless MDX_PPPM_PER_BEAD): every bead takes alpha from the
// secant solve over all P x n ring atoms with one bead's Q^2, as the
// reference's single State of P n atoms does (FixChargeEwald.cu:866-900).
template <class F>
void with_ring(mdx_ctx *const *beads, int P, int flags, double omegaP, F &&f) {
    Ctx &c0 = beads[0]->c;
    MDX_REQUIRE(P >= 2 && P <= 64 && (P & (P - 1)) == 0, MDX_ERR_ARG, "beads: a power of two from 2 to 64");
    std::vector<Ctx *> cs;
    for (int k = 0; k < P; ++k) {
        Ctx &c = beads[k]->c;
        for (Ctx *o : cs) MDX_REQUIRE(o != &c, MDX_ERR_ARG, "a bead appears twice");
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        MDX_REQUIRE(c.n == c0.n && c.device == c0.device && c.dt == c0.dt, MDX_ERR_ARG,
                    "beads must hold the same atoms, device and time step");
        MDX_REQUIRE(!c.rigid.on && !c.baro.on, MDX_ERR_STATE,
                    "ring polymers take flexible molecules at constant volume");
        for (int d = 0; d < 3; ++d)
            MDX_REQUIRE(c.geo.L[d] == c0.geo.L[d] && c.geo.lo[d] == c0.geo.lo[d], MDX_ERR_ARG,
                        "beads must share the box");
        cs.push_back(&c);
    }
    RingSpec ring;
    ring.P = P;
    ring.natoms = c0.n;
    ring.contract = (flags & MDX_PPPM) && !(flags & MDX_PPPM_PER_BEAD) && c0.pppm.on;
    ring.centroidList = !(flags & MDX_RING_BEAD_LISTS);
    for (Ctx *c : cs) {
        MDX_REQUIRE(!ring.centroidList || !c->e3b.on, MDX_ERR_STATE,
                    "a ring's centroid list takes no E3B (use MDX_RING_BEAD_LISTS)");
        // a list of the other kind is built again
        if (c->centroidList != ring.centroidList) c->listValid = false;
    }
    if (ring.contract) {
        for (Ctx *c : cs) {
            MDX_REQUIRE(c->pppm.on && c->pppm.alphaUser == c0.pppm.alphaUser && c->pppm.rcut == c0.pppm.rcut,
                        MDX_ERR_ARG, "contracted beads must share the P3M set-up");
            MDX_REQUIRE(c->pppm.interval <= 1, MDX_ERR_STATE, "ring contraction evaluates the mesh every step");
            MDX_REQUIRE(c->qHost == c0.qHost, MDX_ERR_ARG, "beads must carry the same charges");
        }
    }
    const int want = ring.contract ? P * c0.n : 0;
    for (Ctx *c : cs)
        if (c->pppm.on && c->pppm.alphaAtoms != want) {
            c->pppm.alphaAtoms = want;
            c->pppm.alpha = 0.0;  // solved again at the next preparation
        }
    if (omegaP > 0) {
        // normal-mode matrix and per-mode rotation (nve_xPIMD_cu), in FP64 on the host
        std::vector<double> C((size_t)P * P), cw(3 * (size_t)P);
        const int h = P / 2;
        for (int nb = 0; nb < P; ++nb) {
            const int n = nb + 1;
            for (int k = 0; k < P; ++k) {
                double v;
                if (k == 0) v = std::sqrt(1.0 / P);
                else if (k < h) v = std::sqrt(2.0 / P) * std::cos(2.0 * M_PI * k * n / P);
                else if (k == h) v = std::sqrt(1.0 / P) * ((n & 1) ? -1.0 : 1.0);
                else v = std::sqrt(2.0 / P) * std::sin(2.0 * M_PI * k * n / P);
                C[(size_t)nb * P + k] = v;
            }
        }
        const double dt = c0.dt;
        for (int k = 0; k < P; ++k) {
            if (k == 0) {
                cw[0] = 1.0;
                cw[1] = dt;
                cw[2] = 0.0;
            } else {
                const double w = 2.0 * omegaP * std::sin(k * M_PI / P);
                cw[3 * k] = std::cos(w * dt);
                cw[3 * k + 1] = std::sin(w * dt) / w;
                cw[3 * k + 2] = w * std::sin(w * dt);
            }
        }
        ring.C = dalloc<double>(c0, C.size());
        ring.cw = dalloc<double>(c0, cw.size());
        upload(c0, ring.C, C.data(), C.size());
        upload(c0, ring.cw, cw.data(), cw.size());
    }
    ring.tab = (void **)dalloc<char>(c0, sizeof(void *) * kRingTab * P);
    if (ring.contract || ring.centroidList) ring.cen = dalloc<double4>(c0, c0.n);
    if (ring.centroidList) {
        ring.part = dalloc<float>(c0, (size_t)P * 2 * kCheckBlocks);
        MDX_HIP(hipMemsetAsync(ring.part, 0, sizeof(float) * P * 2 * kCheckBlocks, c0.stream));
        for (int k = 0; k < P; ++k) {
            Ctx &c = *cs[k];
            ring.buildPos.push_back(dalloc<double4>(c0, c0.n));
            ring.ownPart.push_back(c.checkPart);
            c.checkPart = ring.part + (size_t)k * 2 * kCheckBlocks;
            c.ringPart = ring.part;
            c.ringPartNb = P * kCheckBlocks;
            c.cenById = ring.cen;
            c.buildPos = ring.buildPos.back();
        }
    }
    if (ring.contract) {
        ring.q = dalloc<float>(c0, c0.n);
        ring.cf = dalloc<long long>(c0, 3 * (size_t)c0.n);
        upload(c0, ring.q, c0.qHost.data(), c0.n);
        MDX_HIP(hipMemsetAsync(ring.cf, 0, sizeof(long long) * 3 * c0.n, c0.stream));
    }
    std::vector<hipStream_t> own;
    for (Ctx *c : cs) {
        MDX_HIP(hipStreamSynchronize(c->stream));
        own.push_back(c->stream);
        c->stream = c0.stream;
    }
    auto restore = [&] {
        hipStreamSynchronize(c0.stream);
        for (size_t k = 0; k < cs.size(); ++k) cs[k]->stream = own[k];
        for (size_t k = 0; k < ring.ownPart.size(); ++k) {
            Ctx &c = *cs[k];
            c.checkPart = ring.ownPart[k];
            c.ringPart = nullptr;
            c.ringPartNb = 0;
            c.cenById = nullptr;
            c.buildPos = nullptr;
            c.forceFlags = -1;  // a later run on its own builds its own list first
            dfree(c0, ring.buildPos[k]);
        }
        dfree(c0, ring.part);
        dfree(c0, ring.C);
        dfree(c0, ring.cw);
        dfree(c0, (char *)ring.tab);
        dfree(c0, ring.cen);
        dfree(c0, ring.q);
        dfree(c0, ring.cf);
    };
    try {
        f(cs, ring);
    } catch (...) {
        restore();
        throw;
    }
    MDX_HIP(hipStreamSynchronize(c0.stream));
    restore();
}

}  // namespace
}  // namespace mdx

using namespace mdx;

extern "C" {

int mdx_version(void) { return 1; }

int mdx_create(int device, mdx_ctx **out) {
    if (!out) return MDX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MDX_ERR_HIP;
    if (device < 0 || device >= ndev) return MDX_ERR_ARG;
    static std::atomic<uint64_t> serials{0};
    mdx_ctx *ctx = new mdx_ctx();
    ctx->c.device = device;
    ctx->c.serial = ++serials;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->c.stream) != hipSuccess) {
        delete ctx;
        return MDX_ERR_HIP;
    }
    int rc = guarded(ctx, [](Ctx &c) {
        pppm_scratch_init(c.device);  // device address of the no-allocation entry point's scratch
        c.acc = dalloc<unsigned long long>(c, kAccTotal);
        c.dFlags = dalloc<int>(c, 12);
        c.dMax = dalloc<unsigned>(c, 4);
        c.checkPart = dalloc<float>(c, 2 * kCheckBlocks);
        c.thermo = dalloc<unsigned long long>(c, 2);
        MDX_HIP(hipHostMalloc((void **)&c.pinned, sizeof(int) * 8, hipHostMallocDefault));
        std::memset(c.pinned, 0, sizeof(int) * 8);
        MDX_HIP(hipEventCreateWithFlags(&c.capEvent, hipEventDisableTiming));
        for (int k = 0; k < 2; ++k) MDX_HIP(hipEventCreateWithFlags(&c.stepEvent[k], hipEventDisableTiming));
        MDX_HIP(hipMemsetAsync(c.dMax, 0, sizeof(unsigned) * 4, c.stream));
        MDX_HIP(hipMemsetAsync(c.dFlags, 0, sizeof(int) * 12, c.stream));
        MDX_HIP(hipMemsetAsync(c.acc, 0, sizeof(unsigned long long) * kAccTotal, c.stream));
    });
    if (rc != MDX_OK) {
        mdx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return MDX_OK;
}

int mdx_destroy(mdx_ctx *ctx) {
    if (!ctx) return MDX_OK;
    Ctx &c = ctx->c;
    hipSetDevice(c.device);
    if (c.stream) hipStreamSynchronize(c.stream);
    if (c.graphs.big) hipGraphExecDestroy(c.graphs.big);
    if (c.graphs.one) hipGraphExecDestroy(c.graphs.one);
    c.graphs.keep.release();
    pppm_release(c);
    prof_clear(c);
    if (c.pinned) hipHostFree(c.pinned);
    if (c.capEvent) hipEventDestroy(c.capEvent);
    for (int k = 0; k < 2; ++k)
        if (c.stepEvent[k]) hipEventDestroy(c.stepEvent[k]);
    for (hipEvent_t e : c.eventPool) hipEventDestroy(e);
    c.eventPool.clear();
    for (void *p : c.allocs) hipFree(p);
    c.allocs.clear();
    if (c.evFork) hipEventDestroy(c.evFork);
    if (c.evJoin) hipEventDestroy(c.evJoin);
    if (c.side) hipStreamDestroy(c.side);
    if (c.stream) hipStreamDestroy(c.stream);
    delete ctx;
    return MDX_OK;
}

const char *mdx_last_error(const mdx_ctx *ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

int mdx_system_set(mdx_ctx *ctx, int n, const double *x, const double *v, const double *inv_mass,
                   const double *q, const int32_t *type, const double lo[3], const double len[3],
                   const int32_t periodic[3]) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(n > 0 && n < (1 << 30), MDX_ERR_ARG, "atom count must be in [1, 2^30)");
        MDX_REQUIRE(x && lo && len, MDX_ERR_ARG, "positions and box are required");
        MDX_REQUIRE(c.n == 0 || c.n == n, MDX_ERR_STATE, "atom count cannot change on a context");
        hipSetDevice(c.device);
        set_geometry(c, lo, len);
        for (int a = 0; a < 3; ++a) c.geo.periodic[a] = periodic ? (periodic[a] != 0) : 1;
        const bool fresh = c.n == 0;
        c.n = n;
        if (fresh) {
            for (int b = 0; b < 2; ++b) {
                c.pos[b] = dalloc<double4>(c, n);
                c.vel[b] = dalloc<double4>(c, n);
                c.q[b] = dalloc<float>(c, n);
                c.type[b] = dalloc<int>(c, n);
                c.id[b] = dalloc<int>(c, n);
            }
            c.idToIdx = dalloc<int>(c, n);
            c.xq = dalloc<float4>(c, n);
            c.meshRec = dalloc<float4>(c, n);
            c.force = dalloc<unsigned long long>(c, 3 * (size_t)n);
            c.posLast = dalloc<double4>(c, n);
        }
        std::vector<double4> hp(n), hv(n);
        std::vector<float> hq(n);
        std::vector<int> ht(n), hid(n);
        c.invMassHost.assign(n, 1.0);
        c.qHost.assign(n, 0.0f);
        c.typeHost.assign(n, 0);
        c.pairFlagsDirty = true;
        c.hasCharges = false;
        c.nCharged = 0;
        c.maxType = 0;
        for (int i = 0; i < n; ++i) {
            hp[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0);
            double im = inv_mass ? inv_mass[i] : 1.0;
            c.invMassHost[i] = im;
            hv[i] = v ? make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], im) : make_double4(0, 0, 0, im);
            hq[i] = q ? (float)q[i] : 0.0f;
            c.qHost[i] = hq[i];
            if (hq[i] != 0.0f) {
                c.hasCharges = true;
                ++c.nCharged;
            }
            ht[i] = type ? type[i] : 0;
            c.typeHost[i] = ht[i];
            MDX_REQUIRE(ht[i] >= 0 && ht[i] < kMaxTypes, MDX_ERR_ARG, "atom type out of range [0, 16)");
            c.maxType = std::max(c.maxType, ht[i]);
            hid[i] = i;
        }
        c.cur = 0;
        MDX_REQUIRE(!c.ljOn || c.ntypes > c.maxType, MDX_ERR_ARG, "LJ table does not cover every atom type");
        upload(c, c.pos[0], hp.data(), n);
        upload(c, c.vel[0], hv.data(), n);
        upload(c, c.q[0], hq.data(), n);
        upload(c, c.type[0], ht.data(), n);
        upload(c, c.id[0], hid.data(), n);
        upload(c, c.idToIdx, hid.data(), n);
        MDX_HIP(hipStreamSynchronize(c.stream));
        c.listValid = false;
        c.nlistConfigured = false;
        c.nRebuilds = 0;
        update_ox_flag(c);
        if (c.rigid.on) {
            rigid_map(c);
            c.rigid.fresh = true;
        }
        flex_map(c);
        if (c.pppm.on) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += (double)hq[i] * hq[i];
            c.pppm.totalQ2 = c.pppm.qqr * s;
            c.pppm.alpha = 0.0;
        }
        box_upload(c);  // periodic flags and the atom count are set now
    });
}

int mdx_positions_set(mdx_ctx *ctx, const double *x) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(x, MDX_ERR_ARG, "positions required");
        std::vector<int> idx(c.n);
        MDX_HIP(hipMemcpyAsync(idx.data(), c.idToIdx, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        std::vector<double4> hp(c.n);
        for (int i = 0; i < c.n; ++i) hp[idx[i]] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0);
        upload(c, c.pos[c.cur], hp.data(), c.n);
        MDX_HIP(hipStreamSynchronize(c.stream));
        c.listValid = false;
    });
}

int mdx_system_get(mdx_ctx *ctx, double *x, double *v) {
    return guarded_ro(ctx, [&](Ctx &c) {
        require_system(c);
        std::vector<int> idx(c.n);
        std::vector<double4> hp(c.n), hv(c.n);
        MDX_HIP(hipMemcpyAsync(idx.data(), c.idToIdx, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(hp.data(), c.pos[c.cur], sizeof(double4) * c.n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(hv.data(), c.vel[c.cur], sizeof(double4) * c.n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        for (int i = 0; i < c.n; ++i) {
            const double4 &p = hp[idx[i]];
            const double4 &w = hv[idx[i]];
            if (x) {
                x[3 * i] = p.x;
                x[3 * i + 1] = p.y;
                x[3 * i + 2] = p.z;
            }
            if (v) {
                v[3 * i] = w.x;
                v[3 * i + 1] = w.y;
                v[3 * i + 2] = w.z;
            }
        }
    });
}

int mdx_box_set(mdx_ctx *ctx, const double lo[3], const double len[3]) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(lo && len, MDX_ERR_ARG, "box required");
        set_geometry(c, lo, len);
        c.listValid = false;
        c.nlistConfigured = false;
    });
}

int mdx_box_get(mdx_ctx *ctx, double lo[3], double len[3]) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(lo && len, MDX_ERR_ARG, "box output required");
        for (int d = 0; d < 3; ++d) {
            lo[d] = c.geo.lo[d];
            len[d] = c.geo.L[d];
        }
    });
}

int mdx_nlist_config(mdx_ctx *ctx, double r_list, double skin) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(r_list > 0.0 && skin >= 0.0 && skin < r_list, MDX_ERR_ARG,
                    "need r_list > 0 and 0 <= skin < r_list");
        if (r_list != c.rlist && c.nl) {
            dfree(c, c.nl);
            dfree(c, c.nlCount);
            dfree(c, c.nlCountA);
            c.nl = nullptr;
            c.nlCount = nullptr;
            c.nlCountA = nullptr;
            c.cap = 0;
        }
        c.rlist = r_list;
        c.skin = skin;
        c.nlistConfigured = false;
        c.listValid = false;
    });
}

int mdx_exclusions_set(mdx_ctx *ctx, int n_ids, const int32_t *offsets, const uint32_t *tagged_ids) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(n_ids == c.n && offsets, MDX_ERR_ARG, "exclusion offsets must cover every id");
        MDX_REQUIRE(offsets[0] == 0, MDX_ERR_ARG, "offsets[0] must be 0");
        const int total = offsets[n_ids];
        MDX_REQUIRE(total == 0 || tagged_ids, MDX_ERR_ARG, "missing excluded ids");
        int maxe = 0, span = 0;
        for (int i = 0; i < n_ids; ++i) {
            MDX_REQUIRE(offsets[i + 1] >= offsets[i], MDX_ERR_ARG, "offsets must be non-decreasing");
            maxe = std::max(maxe, offsets[i + 1] - offsets[i]);
            for (int k = offsets[i]; k < offsets[i + 1]; ++k) {
                const int j = (int)(tagged_ids[k] & 0x3FFFFFFFu);
                MDX_REQUIRE(j < n_ids && j != i, MDX_ERR_ARG, "excluded id out of range");
                span = std::max(span, std::abs(j - i));
            }
        }
        if (c.exclOff) {
            dfree(c, c.exclOff);
            dfree(c, c.exclIds);
        }
        c.exclOff = dalloc<int>(c, n_ids + 1);
        c.exclIds = dalloc<uint32_t>(c, std::max(total, 1));
        upload(c, c.exclOff, offsets, n_ids + 1);
        if (total) upload(c, c.exclIds, tagged_ids, total);
        MDX_HIP(hipStreamSynchronize(c.stream));
        c.nExclIds = n_ids;
        c.maxExcl = maxe;
        c.exclSpan = span;
        c.listValid = false;
    });
}

int mdx_exclusions_from_bonds(mdx_ctx *ctx, int nbonds, const int32_t *bonds, int max_depth) {
    if (!ctx) return MDX_ERR_ARG;
    std::vector<int32_t> off;
    std::vector<uint32_t> ids;
    int rc = guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(max_depth >= 1 && max_depth <= 3, MDX_ERR_ARG, "max_depth must be 1, 2 or 3");
        const int n = c.n;
        std::vector<std::set<int>> adj(n);
        for (int b = 0; b < nbonds; ++b) {
            int a0 = bonds[2 * b], a1 = bonds[2 * b + 1];
            MDX_REQUIRE(a0 >= 0 && a0 < n && a1 >= 0 && a1 < n, MDX_ERR_ARG, "bond id out of range");
            adj[a0].insert(a1);
            adj[a1].insert(a0);
        }
        off.assign(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            // levels[d]: ids at bond distance d+1 not closer (generateExclusionList)
            std::vector<std::set<int>> levels(1, adj[i]);
            for (int d = 1; d < max_depth; ++d) {
                std::set<int> nxt;
                for (int k : levels[d - 1]) nxt.insert(adj[k].begin(), adj[k].end());
                for (auto it = nxt.begin(); it != nxt.end();) {
                    bool closer = (*it == i);
                    for (int e = 0; e < d && !closer; ++e) closer = levels[e].count(*it) > 0;
                    it = closer ? nxt.erase(it) : std::next(it);
                }
                levels.push_back(nxt);
            }
            for (int d = 0; d < (int)levels.size(); ++d)
                for (int j : levels[d]) ids.push_back((uint32_t)j | ((uint32_t)(d + 1) << 30));
            off[i + 1] = (int32_t)ids.size();
        }
    });
    if (rc != MDX_OK) return rc;
    return mdx_exclusions_set(ctx, ctx->c.n, off.data(), ids.data());
}

int mdx_nlist_update(mdx_ctx *ctx, int force, int *rebuilt) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        long long before = c.nRebuilds;
        if (force || !c.listValid) {
            if (!c.nlistConfigured) nlist_setup(c);
            nlist_rebuild(c, true);
        } else {
            ensure_list(c);
        }
        if (rebuilt) *rebuilt = c.nRebuilds != before;
    });
}

int mdx_nlist_get(mdx_ctx *ctx, int32_t *counts, uint32_t *tagged_ids, int row_cap, int *max_count) {
    return guarded_ro(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.listValid, MDX_ERR_STATE, "no neighbor list built");
        const int n = c.n;
        std::vector<int> cnt(n), ids(n);
        std::vector<uint32_t> nl((size_t)n * c.cap);
        MDX_HIP(hipMemcpyAsync(cnt.data(), c.nlCount, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(ids.data(), c.id[c.cur], sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(nl.data(), c.nl, sizeof(uint32_t) * nl.size(), hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        int mx = 0;
        std::vector<uint32_t> row;
        for (int k = 0; k < n; ++k) {
            const int myid = ids[k];
            const int cn = std::min(cnt[k], c.cap);
            mx = std::max(mx, cnt[k]);
            row.resize(cn);
            for (int s = 0; s < cn; ++s) {
                const uint32_t e = nl[(size_t)k * c.cap + s];
                row[s] = (uint32_t)ids[e & 0x3FFFFFFFu] | (e & 0xC0000000u);
            }
            std::sort(row.begin(), row.end(),
                      [](uint32_t a, uint32_t b) { return (a & 0x3FFFFFFFu) < (b & 0x3FFFFFFFu); });
            if (counts) counts[myid] = cnt[k];
            if (tagged_ids)
                for (int s = 0; s < std::min(cn, row_cap); ++s) tagged_ids[(size_t)myid * row_cap + s] = row[s];
        }
        if (max_count) *max_count = mx;
    });
}

int mdx_nlist_stats(mdx_ctx *ctx, int64_t *rebuilds, int *row_cap, int32_t cells[3],
                    int64_t *violations) {
    return guarded_ro(ctx, [&](Ctx &c) {
        if (rebuilds) *rebuilds = c.nRebuilds;
        if (violations) *violations = c.listViolations;
        if (row_cap) *row_cap = c.cap;
        if (cells)
            for (int a = 0; a < 3; ++a) cells[a] = c.nc[a];
    });
}

int mdx_lj_set(mdx_ctx *ctx, int ntypes, const double *rc2, const double *eps24, const double *sig6) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(ntypes >= 1 && ntypes <= kMaxTypes, MDX_ERR_ARG, "ntypes must be in [1, 16]");
        MDX_REQUIRE(c.n == 0 || ntypes > c.maxType, MDX_ERR_ARG, "LJ table does not cover every atom type");
        const int T2 = ntypes * ntypes;
        std::vector<float> p(3 * T2);
        for (int t = 0; t < T2; ++t) {
            p[t] = (float)rc2[t];
            p[T2 + t] = (float)eps24[t];
            p[2 * T2 + t] = (float)sig6[t];
        }
        if (!c.ljParams) c.ljParams = dalloc<float>(c, 3 * kMaxTypes * kMaxTypes);
        upload(c, c.ljParams, p.data(), p.size());
        MDX_HIP(hipStreamSynchronize(c.stream));
        c.ntypes = ntypes;
        c.ljOn = true;
        c.ljEps.assign(eps24, eps24 + T2);
        c.pairFlagsDirty = true;
    });
}

int mdx_special_set(mdx_ctx *ctx, double c12, double c13, double c14) {
    return guarded(ctx, [&](Ctx &c) {
        c.special[0] = (float)c12;
        c.special[1] = (float)c13;
        c.special[2] = (float)c14;
    });
}

int mdx_ewald_real_set(mdx_ctx *ctx, int enable, double alpha, double qqr, double r_cut) {
    return guarded(ctx, [&](Ctx &c) {
        c.ewaldReal = enable != 0;
        c.alpha = (float)alpha;
        c.qqr = (float)qqr;
        c.rq2 = (float)(r_cut * r_cut);
        if (c.n > 0) box_upload(c);
    });
}

int mdx_pppm_set(mdx_ctx *ctx, int nx, int ny, int nz, int order, double r_cut, double qqr,
                 double alpha) {
    return guarded(ctx, [&](Ctx &c) {
        auto pow2 = [](int v) { return v >= 32 && v <= 1024 && (v & (v - 1)) == 0; };
        MDX_REQUIRE(pow2(nx) && pow2(ny) && pow2(nz), MDX_ERR_ARG,
                    "P3M mesh sizes must be powers of two in [32, 1024] (as FixChargeEwald::setParameters)");
        MDX_REQUIRE(order == 1 || order == 3 || order == 5 || order == 7, MDX_ERR_ARG,
                    "interpolation order must be 1, 3, 5 or 7");
        MDX_REQUIRE(r_cut > 0, MDX_ERR_ARG, "r_cut must be positive");
        const int interval = c.pppm.on ? c.pppm.interval : 1;
        pppm_release(c);
        c.pppm = PppmState();
        c.pppm.on = true;
        c.pppm.interval = interval;
        c.pppm.sz[0] = nx;
        c.pppm.sz[1] = ny;
        c.pppm.sz[2] = nz;
        c.pppm.order = order;
        c.pppm.rcut = r_cut;
        c.pppm.qqr = qqr;
        c.pppm.alphaUser = alpha;
        if (c.n > 0) {
            // qqr sum q^2 over the stored (FP32) charges in FP64, in id order
            // (setTotalQ2, FixChargeEwald.cu:903-915)
            double s = 0.0;
            for (int i = 0; i < c.n; ++i) s += (double)c.qHost[i] * c.qHost[i];
            c.pppm.totalQ2 = qqr * s;
        }
    });
}

int mdx_pppm_set_error(mdx_ctx *ctx, double error, double r_cut, int order, double qqr) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(error > 0 && r_cut > 0, MDX_ERR_ARG, "error and r_cut must be positive");
        MDX_REQUIRE(order == 1 || order == 3 || order == 5 || order == 7, MDX_ERR_ARG,
                    "interpolation order must be 1, 3, 5 or 7");
        // the search starts from the mesh set before, else FixChargeEwald's 64^3
        int sz[3] = {64, 64, 64};
        if (c.pppm.on)
            for (int a = 0; a < 3; ++a) sz[a] = c.pppm.sz[a];
        const int interval = c.pppm.on ? c.pppm.interval : 1;
        pppm_release(c);
        c.pppm = PppmState();
        c.pppm.on = true;
        for (int a = 0; a < 3; ++a) c.pppm.sz[a] = sz[a];
        c.pppm.order = order;
        c.pppm.rcut = r_cut;
        c.pppm.qqr = qqr;
        c.pppm.errorTarget = error;
        c.pppm.interval = interval;
        if (c.n > 0) {
            double s = 0.0;
            for (int i = 0; i < c.n; ++i) s += (double)c.qHost[i] * c.qHost[i];
            c.pppm.totalQ2 = qqr * s;
        }
    });
}

int mdx_pppm_interval_set(mdx_ctx *ctx, int interval) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(c.pppm.on, MDX_ERR_STATE, "P3M not configured");
        MDX_REQUIRE(interval >= 0, MDX_ERR_ARG, "interval must be >= 0");
        if (interval) c.pppm.interval = interval;  // 0 keeps it, as the reference does
        c.pppm.storedValid = false;
    });
}

int mdx_pppm_mesh_get(mdx_ctx *ctx, int32_t sz[3], int *interval) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(c.pppm.on, MDX_ERR_STATE, "P3M not configured");
        require_system(c);
        pppm_prepare(c);
        if (sz)
            for (int a = 0; a < 3; ++a) sz[a] = c.pppm.sz[a];
        if (interval) *interval = c.pppm.interval;
    });
}

int mdx_pppm_info(mdx_ctx *ctx, double *alpha, double *rms_error) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(c.pppm.on, MDX_ERR_STATE, "P3M not configured");
        require_system(c);
        pppm_prepare(c);
        if (alpha) *alpha = c.pppm.alpha;
        if (rms_error) *rms_error = c.pppm.err;
    });
}

int mdx_pppm_paths(mdx_ctx *ctx, int *spread, int *gather, int *fft) {
    return guarded_ro(ctx, [&](Ctx &c) {
        if (spread) *spread = c.pppm.lastSpread;
        if (gather) *gather = c.pppm.lastGather;
        if (fft) *fft = c.pppm.lastFft;
    });
}

int mdx_e3b_set(mdx_ctx *ctx, int nmol, const int32_t *mol_ids, const double params[8], double r_two_body,
                double r_max) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(nmol > 0 && mol_ids && params, MDX_ERR_ARG, "molecules and parameters required");
        E3BState &E = c.e3b;
        if (E.molIds) {
            dfree(c, E.molIds);
            dfree(c, E.molAtoms);
            dfree(c, E.atomMolO);
            dfree(c, E.nl);
            dfree(c, E.nlCount);
            dfree(c, E.tot);
        }
        std::vector<int4> ids(nmol);
        std::vector<int> seen(c.n, 0);
        for (int m = 0; m < nmol; ++m) {
            ids[m] = make_int4(mol_ids[4 * m], mol_ids[4 * m + 1], mol_ids[4 * m + 2], mol_ids[4 * m + 3]);
            for (int k = 0; k < 4; ++k) {
                int a = mol_ids[4 * m + k];
                MDX_REQUIRE(a >= 0 && a < c.n, MDX_ERR_ARG, "E3B molecule id out of range");
                MDX_REQUIRE(!seen[a]++, MDX_ERR_ARG, "atom listed in two E3B molecules");
            }
        }
        E.on = true;
        E.nmol = nmol;
        E.E2 = (float)params[0];
        E.Ea = (float)params[1];
        E.Eb = (float)params[2];
        E.Ec = (float)params[3];
        E.k2 = (float)params[4];
        E.k3 = (float)params[5];
        E.rs = (float)params[6];
        E.rf = (float)params[7];
        E.r2b = (float)r_two_body;
        E.rmax = (float)r_max;
        E.molIds = dalloc<int4>(c, nmol);
        E.molAtoms = dalloc<int4>(c, nmol);
        E.atomMolO = dalloc<int>(c, c.n);
        E.tot = dalloc<float4>(c, nmol);
        E.nlCount = dalloc<int>(c, nmol);
        E.cap = 64;
        E.nl = dalloc<int>(c, (size_t)nmol * E.cap);
        upload(c, E.molIds, ids.data(), nmol);
        fill_int(c, E.atomMolO, c.n, -1);
        MDX_HIP(hipStreamSynchronize(c.stream));
        E.oxIds.resize(nmol);
        for (int m = 0; m < nmol; ++m) E.oxIds[m] = ids[m].x;
        update_ox_flag(c);
        c.listValid = false;
    });
}

int mdx_rigid_set(mdx_ctx *ctx, int nmol, const int32_t *mol_ids, double m_O, double m_H, double d_OH,
                  double d_HH, double gamma) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(nmol > 0 && mol_ids, MDX_ERR_ARG, "molecules required");
        MDX_REQUIRE(m_O > 0 && m_H > 0 && d_OH > 0 && d_HH > 0 && d_HH < 2 * d_OH, MDX_ERR_ARG,
                    "invalid rigid water geometry");
        RigidState &R = c.rigid;
        if (R.ids) {
            dfree(c, R.ids);
            dfree(c, R.atoms);
            dfree(c, R.member);
        }
        std::vector<int4> ids(nmol);
        std::vector<int> seen(c.n, 0);
        for (int m = 0; m < nmol; ++m) {
            const int *p = mol_ids + 4 * m;
            for (int k = 0; k < 3; ++k) {
                MDX_REQUIRE(p[k] >= 0 && p[k] < c.n, MDX_ERR_ARG, "rigid molecule id out of range");
                MDX_REQUIRE(!seen[p[k]]++, MDX_ERR_ARG, "atom listed in two rigid molecules");
            }
            MDX_REQUIRE(p[3] < c.n, MDX_ERR_ARG, "M-site id out of range");
            if (p[3] >= 0) MDX_REQUIRE(!seen[p[3]]++, MDX_ERR_ARG, "atom listed in two rigid molecules");
            // masses ordered O > H1 == H2 > M, as FixRigid::createRigid demands (FixRigid.cu:1646-1658)
            const double iO = c.invMassHost[p[0]], i1 = c.invMassHost[p[1]], i2 = c.invMassHost[p[2]];
            MDX_REQUIRE(iO > 0 && i1 > 0 && i1 == i2 && iO < i1, MDX_ERR_ARG,
                        "ids of a rigid water must be ordered O, H1, H2, M");
            if (p[3] >= 0) MDX_REQUIRE(c.invMassHost[p[3]] == 0.0, MDX_ERR_ARG, "M site must be massless");
            ids[m] = make_int4(p[0], p[1], p[2], p[3]);
        }
        R.on = true;
        R.fresh = true;
        R.nmol = nmol;
        R.nsites = 0;
        for (int m = 0; m < nmol; ++m) R.nsites += ids[m].w >= 0 ? 4 : 3;
        R.mO = m_O;
        R.mH = m_H;
        R.dOH = d_OH;
        R.dHH = d_HH;
        R.gamma = gamma;
        R.ids = dalloc<int4>(c, nmol);
        R.atoms = dalloc<int4>(c, nmol);
        R.member = dalloc<int>(c, c.n);
        upload(c, R.ids, ids.data(), nmol);
        rigid_map(c);
        MDX_HIP(hipStreamSynchronize(c.stream));
    });
}

int mdx_flexible_water_set(mdx_ctx *ctx, int nmol, const int32_t *mol_ids, double k2, double k3, double k4,
                           double r0, double ktheta, double theta0, double gamma) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(nmol > 0 && mol_ids, MDX_ERR_ARG, "molecules required");
        MDX_REQUIRE(!c.rigid.on, MDX_ERR_STATE, "flexible and rigid water cannot be combined");
        MDX_REQUIRE(r0 > 0 && theta0 > 0 && theta0 < M_PI && gamma > 0 && gamma <= 1, MDX_ERR_ARG,
                    "invalid flexible water parameters");
        FlexState &F = c.flex;
        if (F.ids) {
            dfree(c, F.ids);
            dfree(c, F.atoms);
        }
        std::vector<int4> ids(nmol);
        std::vector<int> seen(c.n, 0);
        for (int m = 0; m < nmol; ++m) {
            const int *p = mol_ids + 4 * m;
            for (int k = 0; k < 3; ++k) {
                MDX_REQUIRE(p[k] >= 0 && p[k] < c.n, MDX_ERR_ARG, "molecule id out of range");
                MDX_REQUIRE(!seen[p[k]]++, MDX_ERR_ARG, "atom listed in two molecules");
            }
            MDX_REQUIRE(p[3] < c.n, MDX_ERR_ARG, "M-site id out of range");
            if (p[3] >= 0) {
                MDX_REQUIRE(!seen[p[3]]++, MDX_ERR_ARG, "atom listed in two molecules");
                MDX_REQUIRE(c.invMassHost[p[3]] == 0.0, MDX_ERR_ARG, "M site must be massless");
            }
            ids[m] = make_int4(p[0], p[1], p[2], p[3]);
        }
        F.on = true;
        F.nmol = nmol;
        F.k2 = k2;
        F.k3 = k3;
        F.k4 = k4;
        F.r0 = r0;
        F.ktheta = ktheta;
        F.theta0 = theta0;
        F.gamma = gamma;
        F.ids = dalloc<int4>(c, nmol);
        F.atoms = dalloc<int4>(c, nmol);
        upload(c, F.ids, ids.data(), nmol);
        flex_map(c);
        MDX_HIP(hipStreamSynchronize(c.stream));
        c.forceFlags = -1;
    });
}

int mdx_bonded_energy_get(mdx_ctx *ctx, double *e) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(e, MDX_ERR_ARG, "output required");
        double h[ACC_N];
        acc_read(c, h);
        *e = h[ACC_BOND];
    });
}

int mdx_forces_compute(mdx_ctx *ctx, int flags) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        prepare_terms(c);
        ensure_list(c);
        compute(c, flags);
        MDX_HIP(hipStreamSynchronize(c.stream));
    });
}

int mdx_forces_compute_noalloc(mdx_ctx *ctx, int flags) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        MDX_REQUIRE(c.nlistConfigured && c.nl && c.waveMax, MDX_ERR_STATE,
                    "no neighbor list yet: call mdx_forces_compute once to set the context up");
        const bool mesh = (flags & MDX_PPPM) && c.pppm.on;
        MDX_REQUIRE(!mesh || c.pppm.plansMade, MDX_ERR_STATE,
                    "no P3M plans yet: call mdx_forces_compute once to set the context up");
        struct NoAlloc {
            Ctx &c;
            explicit NoAlloc(Ctx &c_) : c(c_) { c.noAlloc = true; }
            ~NoAlloc() { c.noAlloc = false; }
        } guard(c);
        prepare_terms(c);
        if (mesh) {
            PppmScratchScope scr(c);
            ensure_list(c);
            compute(c, flags);
            MDX_HIP(hipStreamSynchronize(c.stream));
        } else {
            ensure_list(c);
            compute(c, flags);
            MDX_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int mdx_forces_get(mdx_ctx *ctx, double *f) {
    return guarded_ro(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(f, MDX_ERR_ARG, "output required");
        std::vector<unsigned long long> hf(3 * (size_t)c.n);
        std::vector<int> ids(c.n);
        MDX_HIP(hipMemcpyAsync(hf.data(), c.force, sizeof(unsigned long long) * hf.size(), hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(ids.data(), c.id[c.cur], sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        for (int k = 0; k < c.n; ++k)
            for (int a = 0; a < 3; ++a) f[3 * ids[k] + a] = (double)(long long)hf[3 * k + a] * kFxInv;
    });
}

int mdx_energy_get(mdx_ctx *ctx, double e[4], double virial[6]) {
    return guarded_ro(ctx, [&](Ctx &c) {
        double h[ACC_N];
        acc_read(c, h);
        auto d = [&](int s) { return h[s]; };
        if (e) {
            e[0] = d(ACC_PAIR);
            e[1] = d(ACC_KSPACE);
            if (c.pppm.on) e[1] -= c.pppm.alpha / std::sqrt(M_PI) * c.pppm.totalQ2;
            e[2] = d(ACC_E3B2);
            e[3] = d(ACC_E3B3);
        }
        if (virial)
            for (int a = 0; a < 6; ++a) virial[a] = d(ACC_VIR + a);
    });
}

int mdx_virial_trace_get(mdx_ctx *ctx, double *trace) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(trace, MDX_ERR_ARG, "output required");
        double h[ACC_N];
        acc_read(c, h);
        *trace = h[ACC_VIR] + h[ACC_VIR + 1] + h[ACC_VIR + 2] + h[ACC_VTRACE];
    });
}

int mdx_integrator_set(mdx_ctx *ctx, double dt, double ftm_to_v) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(dt > 0, MDX_ERR_ARG, "dt must be positive");
        c.dt = dt;
        c.ftmToV = ftm_to_v;
    });
}

int mdx_run(mdx_ctx *ctx, int nsteps, int flags) {
    return guarded_ro(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        MDX_REQUIRE(nsteps >= 0, MDX_ERR_ARG, "nsteps must be >= 0");
        c.pppm.turnInit = c.turn;  // the long-range interval counts from a run's first step
        // a long-range interval > 1 runs the steps eagerly (the host picks the
        // mesh steps); captured steps evaluate the mesh every step
        if (c.graphSteps && c.pppm.interval <= 1) {
            run_captured({&c}, nsteps, flags);
            return;
        }
        const int flags0 = flags & kForceTerms;
        int fflags = judged_flags(c, flags0);
        const bool judge = judging(c, flags0);
        prepare_terms(c);
        if (c.rigid.on && c.rigid.fresh) {
            // constrain the starting velocities and place the M sites
            rigid_second_half(c, false);
            c.rigid.fresh = false;
            c.listValid = false;
        }
        if (c.nh.on && c.nh.fresh) {
            c.nh.ndf = degrees_of_freedom(c);
            nose_hoover_start(c);
        }
        struct InRun {
            Ctx &c;
            explicit InRun(Ctx &c_) : c(c_) { c.inRun = true; }
            ~InRun() { c.inRun = false; }
        } inRun(c);
        constexpr int kChunk = 64;
        int done = 0, chunkNo = 0;
        bool forcedDone = false;
        while (done < nsteps) {
            int len = std::min(kChunk, nsteps - done);
            if (judge && !c.sur.fellBack) len = std::min(len, steps_to_judge(c));
            snapshot_save(c);
            if (!c.listValid || c.forceFlags != fflags) {
                ensure_list(c);
                compute(c, fflags);
            }
            // the "previous step" check for step 0 of the chunk, from cleared slots
            fill_int(c, (int *)c.dMax, 4, 0);
            check_and_copy(c, 1);
            for (int s = 0; s < len; ++s) {
                const int par = s & 1;
                const bool baro = c.baro.on && (c.turn + 1) % c.baro.every == 0;
                integrate_first_half(c);
                if (lagged_rebuild_decision(c, 1 - par)) nlist_rebuild_async(c);
                check_and_copy(c, par);
                compute(c, fflags | (baro ? kFlagVirialTrace : 0));
                integrate_second_half(c);
                ++c.turn;
                if (c.thermoOn && c.turn % c.thermoEvery == 0) thermostat_apply(c);
                if (baro) {
                    barostat_apply(c);
                    if (!c.listValid) {
                        ensure_list(c);
                        compute(c, fflags);
                        fill_int(c, (int *)c.dMax, 4, 0);
                        check_and_copy(c, par);
                    }
                }
            }
            MDX_HIP(hipStreamSynchronize(c.stream));
            lagged_rebuild_decision(c, (len - 1) & 1);  // count a violation in the last step
            const bool forced = !forcedDone && c.replayChunk == chunkNo + 1;
            forcedDone = forcedDone || forced;
            if (!capacity_ok(c) || forced) {
                // a build overflowed its rows: redo the chunk with the grown rows
                snapshot_restore(c);
                c.forceFlags = -1;
                continue;
            }
            ++chunkNo;
            done += len;
            if (judge) {
                judge_if_due(c, flags0);
                fflags = judged_flags(c, flags0);
            }
        }
        MDX_HIP(hipStreamSynchronize(c.stream));
    });
}

int mdx_run_batch(mdx_ctx *const *ctxs, int n, int nsteps, int flags) {
    if (!ctxs || n <= 0) return MDX_ERR_ARG;
    for (int k = 0; k < n; ++k)
        if (!ctxs[k]) return MDX_ERR_ARG;
    return guarded_ro(ctxs[0], [&](Ctx &) {
        MDX_REQUIRE(nsteps >= 0, MDX_ERR_ARG, "nsteps must be >= 0");
        std::vector<Ctx *> cs;
        for (int k = 0; k < n; ++k) {
            for (Ctx *o : cs) MDX_REQUIRE(o != &ctxs[k]->c, MDX_ERR_ARG, "a context appears twice in the batch");
            cs.push_back(&ctxs[k]->c);
        }
        run_captured(cs, nsteps, flags);
    });
}

int mdx_tune_batch(mdx_ctx *const *ctxs, int n, int flags, int steps, int32_t *chosen, double *ms_per_step) {
    if (!ctxs || n <= 0) return MDX_ERR_ARG;
    for (int k = 0; k < n; ++k)
        if (!ctxs[k]) return MDX_ERR_ARG;
    return guarded_ro(ctxs[0], [&](Ctx &) {
        MDX_REQUIRE(steps > 0 && steps <= 10000, MDX_ERR_ARG, "steps must be in 1..10000");
        std::vector<Ctx *> cs;
        for (int k = 0; k < n; ++k) {
            for (Ctx *o : cs) MDX_REQUIRE(o != &ctxs[k]->c, MDX_ERR_ARG, "a context appears twice in the batch");
            cs.push_back(&ctxs[k]->c);
        }
        // one value of a knob on every replica; the step graphs are captured again
        auto apply = [&](int build, int pair, int e3b) {
            for (Ctx *c : cs) {
                c->buildTeam = build;
                c->buildTeamUser = true;
                c->pairTeam = pair;
                c->e3b.team = e3b;
                ++c->apiEpoch;
            }
        };
        auto timed = [&](int build, int pair, int e3b) {
            apply(build, pair, e3b);
            run_captured(cs, steps, flags);  // captures the graphs, warms the caches
            const auto t0 = std::chrono::steady_clock::now();
            run_captured(cs, steps, flags);
            const auto t1 = std::chrono::steady_clock::now();
            return std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
        };
        int build = cs[0]->buildTeam, pair = cs[0]->pairTeam, e3b = cs[0]->e3b.team;
        double best = timed(build, pair, e3b);
        auto knob = [&](int &v, std::initializer_list<int> cands, auto make) {
            for (int cand : cands) {
                if (cand == v) continue;
                const double t = make(cand);
                if (t < best) {
                    best = t;
                    v = cand;
                }
            }
        };
        if (flags & MDX_PAIR) knob(pair, {8, 16, 32}, [&](int t) { return timed(build, t, e3b); });
        if ((flags & (MDX_E3B | MDX_SURROGATE)) && cs[0]->e3b.on)
            knob(e3b, {16, 32}, [&](int t) { return timed(build, pair, t); });
        knob(build, {64, 8, 16}, [&](int t) { return timed(t, pair, e3b); });
        apply(build, pair, e3b);
        if (chosen) {
            chosen[0] = build;
            chosen[1] = pair;
            chosen[2] = e3b;
        }
        if (ms_per_step) *ms_per_step = best;
    });
}

int mdx_batch_merged(mdx_ctx *ctx, int *merged) {
    return guarded_ro(ctx, [&](Ctx &c) {
        if (merged) *merged = c.batchMerged;
    });
}

int mdx_ring_run(mdx_ctx *const *beads, int P, int nsteps, int flags, double omegaP) {
    if (!beads || P <= 0) return MDX_ERR_ARG;
    for (int k = 0; k < P; ++k)
        if (!beads[k]) return MDX_ERR_ARG;
    return guarded(beads[0], [&](Ctx &) {
        MDX_REQUIRE(nsteps >= 0, MDX_ERR_ARG, "nsteps must be >= 0");
        MDX_REQUIRE(omegaP > 0, MDX_ERR_ARG, "ring frequency must be positive");
        with_ring(beads, P, flags, omegaP, [&](std::vector<Ctx *> &cs, RingSpec &ring) {
            run_captured(cs, nsteps, flags & ~(MDX_PPPM_PER_BEAD | MDX_RING_BEAD_LISTS), &ring);
        });
    });
}

int mdx_ring_forces_compute(mdx_ctx *const *beads, int P, int flags) {
    if (!beads || P <= 0) return MDX_ERR_ARG;
    for (int k = 0; k < P; ++k)
        if (!beads[k]) return MDX_ERR_ARG;
    return guarded(beads[0], [&](Ctx &) {
        with_ring(beads, P, flags, 0.0, [&](std::vector<Ctx *> &cs, RingSpec &ring) {
            const int fflags = flags & ~(MDX_PPPM_PER_BEAD | MDX_RING_BEAD_LISTS);
            for (Ctx *c : cs) prepare_terms(*c);
            ring_lists(cs, ring);
            for (Ctx *c : cs) ensure_list(*c);
            ring_table(cs, ring);
            ring_forces(cs, ring, fflags);
        });
    });
}

int mdx_force_steps(mdx_ctx *ctx, int nsteps, int flags, int rebuild_every) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        MDX_REQUIRE(nsteps >= 0 && rebuild_every >= 0, MDX_ERR_ARG, "counts must be >= 0");
        prepare_terms(c);
        for (;;) {
            for (int s = 0; s < nsteps; ++s) {
                if (!c.listValid) {
                    ensure_list(c);
                } else if (rebuild_every > 0 && s % rebuild_every == 0) {
                    nlist_rebuild_async(c);
                }
                compute(c, flags & kForceTerms);
            }
            MDX_HIP(hipStreamSynchronize(c.stream));
            if (capacity_ok(c)) break;
            c.listValid = false;  // rows grew: evaluate again with complete rows
        }
    });
}

int mdx_teams_set(mdx_ctx *ctx, int build_team, int pair_team, int e3b_team) {
    return guarded(ctx, [&](Ctx &c) {
        auto ok = [](int v, int lo, int hi) { return v == 0 || (v >= lo && v <= hi && (v & (v - 1)) == 0); };
        MDX_REQUIRE(ok(build_team, 8, 64) && ok(pair_team, 4, 32) && ok(e3b_team, 4, 32), MDX_ERR_ARG,
                    "team sizes: build 8..64, pair 4..32, e3b 4..32, powers of two");
        if (build_team) {
            c.buildTeam = build_team;
            c.buildTeamUser = true;
        }
        if (pair_team) c.pairTeam = pair_team;
        if (e3b_team) c.e3b.team = e3b_team;
    });
}

// Grid search of the launch knobs by timing, as Integrator::tune does for
// nThreadPerBlock x nThreadPerAtom (src/Integrators/Integrator.cu:272-376):
// each candidate runs `evals` times between HIP events on the library stream
// and the fastest average wins.  The list build is timed as a forced rebuild
// (both kernels, every team width); pair and E3B kernels at the current list.
int mdx_tune(mdx_ctx *ctx, int flags, int evals, int32_t *chosen, double *avg_ms) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.rlist > 0, MDX_ERR_STATE, "mdx_nlist_config not called");
        MDX_REQUIRE(evals > 0 && evals <= 1000, MDX_ERR_ARG, "evals must be in 1..1000");
        prepare_terms(c);
        ensure_list(c);
        MDX_HIP(hipStreamSynchronize(c.stream));
        if (!capacity_ok(c)) nlist_rebuild(c, true);
        hipEvent_t e0, e1;
        MDX_HIP(hipEventCreate(&e0));
        MDX_HIP(hipEventCreate(&e1));
        auto timed = [&](auto &&body) {
            body();  // warm
            MDX_HIP(hipEventRecord(e0, c.stream));
            for (int k = 0; k < evals; ++k) body();
            MDX_HIP(hipEventRecord(e1, c.stream));
            MDX_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            MDX_HIP(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / evals;
        };
        const bool prof = c.prof;
        c.prof = false;
        // list build: (mode, team) candidates; rows regrow if a candidate overflows
        int bestMode = 2, bestTeam = c.buildTeam;
        double bestBuild = 1e30;
        const int modes[4][2] = {{2, 8}, {2, 16}, {2, 32}, {2, 64}};
        for (auto &mt : modes) {
            c.buildMode = mt[0];
            if (mt[1]) c.buildTeam = mt[1];
            const double t = timed([&] { nlist_rebuild_async(c); });
            MDX_HIP(hipStreamSynchronize(c.stream));
            MDX_REQUIRE(capacity_ok(c), MDX_ERR_STATE, "list rows changed size while tuning");
            if (t < bestBuild) {
                bestBuild = t;
                bestMode = mt[0];
                bestTeam = mt[1] ? mt[1] : bestTeam;
            }
        }
        c.buildMode = bestMode;
        c.buildTeam = bestTeam;
        c.buildTeamUser = true;
        nlist_rebuild(c, true);
        refresh_xq(c);
        // pair kernel
        int bestPair = c.pairTeam;
        double bestPairT = 1e30;
        if (flags & MDX_PAIR) {
            for (int t : {4, 8, 16, 32}) {
                c.pairTeam = t;
                const double ms = timed([&] { pair_forces(c, false, false); });
                if (ms < bestPairT) {
                    bestPairT = ms;
                    bestPair = t;
                }
            }
            c.pairTeam = bestPair;
        }
        // E3B
        int bestE3b = c.e3b.team;
        double bestE3bT = 1e30;
        if ((flags & MDX_E3B) && c.e3b.on) {
            for (int t : {4, 8, 16, 32}) {
                c.e3b.team = t;
                const double ms = timed([&] { e3b_forces(c, false, false); });
                if (ms < bestE3bT) {
                    bestE3bT = ms;
                    bestE3b = t;
                }
            }
            c.e3b.team = bestE3b;
        }
        c.prof = prof;
        MDX_HIP(hipEventDestroy(e0));
        MDX_HIP(hipEventDestroy(e1));
        // leave forces consistent with the current positions and flags
        compute(c, c.forceFlags > 0 ? c.forceFlags : (flags & kForceTerms));
        MDX_HIP(hipStreamSynchronize(c.stream));
        if (chosen) {
            chosen[0] = bestMode;
            chosen[1] = bestTeam;
            chosen[2] = bestPair;
            chosen[3] = bestE3b;
        }
        if (avg_ms) {
            avg_ms[0] = bestBuild;
            avg_ms[1] = bestPairT < 1e30 ? bestPairT : 0.0;
            avg_ms[2] = bestE3bT < 1e30 ? bestE3bT : 0.0;
        }
    });
}

static void surrogate_alloc(Ctx &c) {
    SurrogateState &S = c.sur;
    if (!S.params) S.params = dalloc<float>(c, kSurNParam);
    if (S.nmolAlloc != c.e3b.nmol) {
        if (S.feat) {
            dfree(c, S.feat);
            dfree(c, S.emol);
            dfree(c, S.wgt);
        }
        S.nmolAlloc = c.e3b.nmol;
        S.feat = dalloc<float>(c, (size_t)S.nmolAlloc * kSurNR);
        S.emol = dalloc<float>(c, S.nmolAlloc);
        S.wgt = dalloc<float>(c, (size_t)S.nmolAlloc * kSurNW);
    }
}

int mdx_surrogate_set(mdx_ctx *ctx, const float *params, int nparams) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(c.e3b.on, MDX_ERR_STATE, "the surrogate replaces E3B: call mdx_e3b_set first");
        MDX_REQUIRE(params && nparams == kSurNParam, MDX_ERR_ARG, "surrogate parameter count mismatch");
        surrogate_alloc(c);
        upload(c, c.sur.params, params, kSurNParam);
        surrogate_pack(c);
        MDX_HIP(hipStreamSynchronize(c.stream));
        c.sur.on = true;
        c.forceFlags = -1;
    });
}

int mdx_surrogate_get(mdx_ctx *ctx, float *params, int nparams) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(c.sur.params, MDX_ERR_STATE, "no surrogate model");
        MDX_REQUIRE(params && nparams == kSurNParam, MDX_ERR_ARG, "surrogate parameter count mismatch");
        MDX_HIP(hipMemcpyAsync(params, c.sur.params, sizeof(float) * kSurNParam, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
    });
}

int mdx_surrogate_record(mdx_ctx *ctx, int reset, int capacity) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.e3b.on, MDX_ERR_STATE, "training targets come from E3B: call mdx_e3b_set first");
        SurrogateState &S = c.sur;
        surrogate_alloc(c);
        if (reset) S.nsamp = 0;
        if (capacity > S.capSamp) {
            MDX_REQUIRE(S.nsamp == 0, MDX_ERR_STATE, "set the sample capacity on the first (reset) record");
            if (S.trainX) {
                dfree(c, S.trainX);
                dfree(c, S.trainT);
            }
            S.capSamp = capacity;
            S.trainX = dalloc<float>(c, (size_t)capacity * kSurNR);
            S.trainT = dalloc<float>(c, capacity);
        }
        MDX_REQUIRE(S.nsamp + c.e3b.nmol <= S.capSamp, MDX_ERR_ARG, "sample capacity exhausted");
        prepare_terms(c);
        ensure_list(c);
        refresh_xq(c);
        surrogate_features(c, S.trainX + (size_t)S.nsamp * kSurNR);
        zero_forces(c, true);
        e3b_forces(c, false, false, true, S.trainT + S.nsamp);
        MDX_HIP(hipStreamSynchronize(c.stream));
        MDX_REQUIRE(capacity_ok(c), MDX_ERR_STATE, "list rows grew while recording: record again");
        S.nsamp += c.e3b.nmol;
        c.forceFlags = -1;  // the force buffer now holds E3B only
    });
}

int mdx_surrogate_fit(mdx_ctx *ctx, int epochs, int batch, double lr, int seed, int normalize, double *loss) {
    return guarded(ctx, [&](Ctx &c) {
        SurrogateState &S = c.sur;
        MDX_REQUIRE(S.params && S.nsamp > 0, MDX_ERR_STATE, "no model or no recorded samples");
        MDX_REQUIRE(epochs >= 0 && batch >= 16 && lr > 0, MDX_ERR_ARG, "epochs >= 0, batch >= 16, lr > 0");
        if (normalize) surrogate_normalize(c);
        const double l = surrogate_fit(c, epochs, batch, lr, (unsigned)seed);
        MDX_HIP(hipStreamSynchronize(c.stream));
        if (loss) *loss = l;
        c.forceFlags = -1;
    });
}

int mdx_surrogate_grad(mdx_ctx *ctx, int bstart, int bsize, uint32_t mul, uint32_t add, float *grad, int n) {
    return guarded(ctx, [&](Ctx &c) {
        SurrogateState &S = c.sur;
        const int nt = surrogate_trained_count();
        MDX_REQUIRE(S.params && S.nsamp > 0, MDX_ERR_STATE, "no model or no recorded samples");
        MDX_REQUIRE(grad && n >= nt + 1, MDX_ERR_ARG, "grad needs the trained parameter count + 1 floats");
        MDX_REQUIRE(bstart >= 0 && bsize >= 1 && bsize <= S.nsamp && mul >= 1, MDX_ERR_ARG,
                    "0 <= bstart, 1 <= bsize <= recorded samples, mul >= 1");
        float *d = dalloc<float>(c, nt + 1);
        try {
            surrogate_grad(c, bstart, bsize, mul, add, d);
            MDX_HIP(hipMemcpyAsync(grad, d, sizeof(float) * (nt + 1), hipMemcpyDeviceToHost, c.stream));
            MDX_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            dfree(c, d);
            throw;
        }
        dfree(c, d);
    });
}

int mdx_surrogate_judge_set(mdx_ctx *ctx, int every, double max_rmse) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(every >= 0 && (every == 0 || max_rmse > 0), MDX_ERR_ARG, "judge: every >= 0, max_rmse > 0");
        SurrogateState &S = c.sur;
        S.judgeEvery = every;
        S.judgeMax = max_rmse;
        S.judgeChecks = 0;
        S.judgeLast = S.judgeWorst = 0.0;
        S.fellBack = false;
        if (every > 0 && S.judgeN != c.n) {
            if (S.fSave) {
                dfree(c, S.fSave);
                dfree(c, S.fE3B);
            }
            S.fSave = dalloc<unsigned long long>(c, 3 * (size_t)c.n);
            S.fE3B = dalloc<unsigned long long>(c, 3 * (size_t)c.n);
            if (!S.judgeSum) S.judgeSum = dalloc<double>(c, 1);
            S.judgeN = c.n;
        }
    });
}

int mdx_surrogate_judge_stats(mdx_ctx *ctx, int64_t *checks, double *last_rmse, double *worst_rmse, int *fell_back) {
    return guarded_ro(ctx, [&](Ctx &c) {
        const SurrogateState &S = c.sur;
        if (checks) *checks = S.judgeChecks;
        if (last_rmse) *last_rmse = S.judgeLast;
        if (worst_rmse) *worst_rmse = S.judgeWorst;
        if (fell_back) *fell_back = S.fellBack ? 1 : 0;
    });
}

int mdx_surrogate_eval(mdx_ctx *ctx, float *emol, float *feat) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.sur.on, MDX_ERR_STATE, "no surrogate model");
        prepare_terms(c);
        ensure_list(c);
        refresh_xq(c);
        zero_forces(c, true);
        surrogate_forces(c, false, false, feat != nullptr);
        const size_t m = c.e3b.nmol;
        if (emol) MDX_HIP(hipMemcpyAsync(emol, c.sur.emol, sizeof(float) * m, hipMemcpyDeviceToHost, c.stream));
        if (feat)
            MDX_HIP(hipMemcpyAsync(feat, c.sur.feat, sizeof(float) * m * kSurNR, hipMemcpyDeviceToHost,
                                   c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        c.forceFlags = -1;
    });
}

int mdx_e3b_molecule_energies(mdx_ctx *ctx, float *emol) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.e3b.on && emol, MDX_ERR_ARG, "E3B not set or no output");
        prepare_terms(c);
        ensure_list(c);
        refresh_xq(c);
        zero_forces(c, true);
        float *d = dalloc<float>(c, c.e3b.nmol);
        e3b_forces(c, false, false, true, d);
        MDX_HIP(hipMemcpyAsync(emol, d, sizeof(float) * c.e3b.nmol, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        dfree(c, d);
        c.forceFlags = -1;
    });
}

int mdx_profile_enable(mdx_ctx *ctx, int enable) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_HIP(hipStreamSynchronize(c.stream));
        prof_clear(c);
        c.prof = enable != 0;
    });
}

int mdx_profile_get(mdx_ctx *ctx, int kernel, int *count, double *total_ms) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(kernel >= 0 && kernel < PK_N, MDX_ERR_ARG, "unknown kernel id");
        MDX_HIP(hipStreamSynchronize(c.stream));
        int n = 0;
        double t = 0.0;
        for (auto &r : c.profRecs) {
            if (r.kernel != kernel) continue;
            float ms = 0.f;
            MDX_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            t += ms;
            ++n;
        }
        if (count) *count = n;
        if (total_ms) *total_ms = t;
    });
}

int mdx_pair_entries_read(mdx_ctx *ctx, int64_t *entries, int64_t *atoms) {
    return guarded_ro(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.listValid, MDX_ERR_STATE, "no neighbor list built");
        const int n = c.n;
        std::vector<int> cnt(n), cntA(n), type(n);
        std::vector<float> q(n);
        MDX_HIP(hipMemcpyAsync(cnt.data(), c.nlCount, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(cntA.data(), c.nlCountA, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(type.data(), c.type[c.cur], sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipMemcpyAsync(q.data(), c.q[c.cur], sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        const std::vector<int> fl = pair_type_flag_table(c);
        const bool coul = c.ewaldReal && c.hasCharges;
        int64_t e = 0, a = 0;
        for (int i = 0; i < n; ++i) {
            const int tot = std::min(cnt[i], c.cap), na = std::min(cntA[i], tot);
            const int f = c.ljOn ? fl[type[i]] : 0;
            const bool p0 = (f & 1) || (coul && q[i] != 0.0f), p1 = (f & 2) != 0;
            e += (p0 ? na : 0) + (p1 ? tot - na : 0);
            a += (p0 || p1) ? 1 : 0;
        }
        if (entries) *entries = e;
        if (atoms) *atoms = a;
    });
}

int mdx_nlist_totals(mdx_ctx *ctx, int64_t *pair_entries, int64_t *mol_entries) {
    return guarded_ro(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(c.listValid, MDX_ERR_STATE, "no neighbor list built");
        std::vector<int> cnt(c.n);
        MDX_HIP(hipMemcpyAsync(cnt.data(), c.nlCount, sizeof(int) * c.n, hipMemcpyDeviceToHost, c.stream));
        std::vector<int> mc(c.e3b.on ? c.e3b.nmol : 0);
        if (c.e3b.on)
            MDX_HIP(hipMemcpyAsync(mc.data(), c.e3b.nlCount, sizeof(int) * mc.size(), hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        int64_t s = 0, m = 0;
        for (int v : cnt) s += std::min(v, c.cap);
        for (int v : mc) m += std::min(v, c.e3b.cap);
        if (pair_entries) *pair_entries = s;
        if (mol_entries) *mol_entries = m;
    });
}

int mdx_option_set(mdx_ctx *ctx, const char *name, int value) {
    return guarded(ctx, [&](Ctx &c) {
        MDX_REQUIRE(name, MDX_ERR_ARG, "option name required");
        std::string k(name);
        if (k == "tiled_spread") c.mdxDisableTiledSpread = (value == 0);
        else if (k == "graph") {
            MDX_REQUIRE(value == 0 || value == 1, MDX_ERR_ARG, "graph: 0 eager steps, 1 captured steps");
            c.graphSteps = value;
        }
        else if (k == "build_mode") {
            MDX_REQUIRE(value == 0 || value == 2, MDX_ERR_ARG, "build_mode: 0 or 2 (per-atom teams)");
            c.buildMode = value;
        }
        else if (k == "fuse_step") {
            // captured steps: 1 (default) the step decisions, the rebuild's
            // molecule map and row maxima and the thermostat's rescale ride on
            // other launches; 0 each in a launch of its own (same results)
            c.fuseStep = value != 0;
            if (c.graphs.big || c.graphs.one) {
                MDX_HIP(hipStreamSynchronize(c.stream));
                if (c.graphs.big) hipGraphExecDestroy(c.graphs.big);
                if (c.graphs.one) hipGraphExecDestroy(c.graphs.one);
                c.graphs.keep.release();
                c.graphs = Ctx::GraphCache{};
            }
        }
        else if (k == "mesh_stream") {
            // captured steps: 1 (default) the P3M chain on a second stream beside the pair and E3B kernels
            c.meshStream = value != 0;
            if (c.graphs.big || c.graphs.one) {
                MDX_HIP(hipStreamSynchronize(c.stream));
                if (c.graphs.big) hipGraphExecDestroy(c.graphs.big);
                if (c.graphs.one) hipGraphExecDestroy(c.graphs.one);
                c.graphs.keep.release();
                c.graphs = Ctx::GraphCache{};
            }
        }
        else if (k == "p3m_spread") {
            MDX_REQUIRE(value >= 0 && value <= 2, MDX_ERR_ARG,
                        "p3m_spread: 2 owner-computes tiles, 1 LDS tiles with global atomics, 0 global atomics");
            c.pppm.spreadMode = value;
        }
        else if (k == "p3m_gather") {
            MDX_REQUIRE(value >= 0 && value <= 2, MDX_ERR_ARG,
                        "p3m_gather: 2 LDS tiles of the field, 1 row teams from the mesh records, 0 row teams from the positions");
            c.pppm.gatherMode = value;
        }
        else if (k == "gather_tile_kb") {
            MDX_REQUIRE(value >= 8 && value <= 64, MDX_ERR_ARG, "gather_tile_kb: 8..64 KB of LDS per tile");
            c.pppm.tileKB = value;
        }
        else if (k == "gather_tile_team") {
            MDX_REQUIRE(value == 4 || value == 8 || value == 16, MDX_ERR_ARG, "gather_tile_team: 4, 8 or 16");
            c.pppm.tileTeam = value;
        }
        else if (k == "gather_tile_planar") {
            c.pppm.tilePlanar = value != 0;
        }
        else if (k == "spread_tile_kb") {
            MDX_REQUIRE(value >= 8 && value <= 64, MDX_ERR_ARG, "spread_tile_kb: 8 to 64 KB of LDS a spread tile");
            c.pppm.spreadKB = value;
        }
        else if (k == "p3m_tile_cols") {
            MDX_REQUIRE(value == 1 || value == 2, MDX_ERR_ARG, "p3m_tile_cols: 1 or 2 cell columns per tile side");
            c.pppm.tileCols = value;
        }
        else if (k == "p3m_fft") {
            MDX_REQUIRE(value == 0 || value == 1, MDX_ERR_ARG, "p3m_fft: 1 LDS passes, 0 hipFFT");
            c.pppm.fftMode = value;
        }
        else if (k == "gather_team") {
            MDX_REQUIRE(value == 8 || value == 16, MDX_ERR_ARG, "gather_team: 8 or 16 lanes per charged atom");
            c.pppm.gatherTeam = value;
        }
        else if (k == "cell_reach_xy" || k == "cell_reach_z") {
            // cell edge >= rlist / reach: finer cells cut the candidate volume,
            // the xy reach also widens the build kernel's column table
            const bool xy = k == "cell_reach_xy";
            MDX_REQUIRE(value >= 1 && value <= (xy ? 2 : 16), MDX_ERR_ARG,
                        "cell_reach_xy: 1..2, cell_reach_z: 1..16");
            if (xy)
                c.reach[0] = c.reach[1] = value;
            else
                c.reach[2] = value;
            c.nlistConfigured = false;
            c.listValid = false;
        }
        else if (k == "sur_force_lanes") {
            MDX_REQUIRE(value == 8 || value == 16 || value == 32 || value == 64, MDX_ERR_ARG,
                        "sur_force_lanes: 8, 16, 32 or 64 lanes per molecule");
            c.sur.forceLanes = value;
        }
        else if (k == "batch_merge") {
            // batches led by this context: one launch per kernel across the
            // replicas (1, default) or one stream per replica (0)
            c.batchMerge = value != 0;
            if (c.graphs.big || c.graphs.one) {
                MDX_HIP(hipStreamSynchronize(c.stream));
                if (c.graphs.big) hipGraphExecDestroy(c.graphs.big);
                if (c.graphs.one) hipGraphExecDestroy(c.graphs.one);
                c.graphs.keep.release();
                c.graphs = Ctx::GraphCache{};
            }
        }
        else if (k == "replay_chunk") {
            MDX_REQUIRE(value >= 0, MDX_ERR_ARG, "replay_chunk: chunk number >= 1, 0 for off");
            c.replayChunk = value;
        }
        else if (k == "row_cap") {
            // neighbor-row capacity now (rows still grow on overflow, the
            // step loop replays a chunk whose rebuild overflowed)
            MDX_REQUIRE(value >= 16 && c.nlCount, MDX_ERR_ARG, "row_cap: >= 16, after the first list build");
            MDX_HIP(hipStreamSynchronize(c.stream));
            dfree(c, c.nl);
            c.cap = value;
            c.nl = dalloc<uint32_t>(c, (size_t)c.n * c.cap);
            c.listValid = false;
        }
        else throw Error{MDX_ERR_ARG, "unknown option " + k};
    });
}

int mdx_thermostat_set(mdx_ctx *ctx, int enable, double temperature, int every, double tau,
                       double boltz, double mvv_to_eng) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(!enable || (temperature > 0 && every >= 1 && boltz > 0 && mvv_to_eng > 0), MDX_ERR_ARG,
                    "thermostat needs T > 0, every >= 1, boltz > 0, mvv_to_eng > 0");
        MDX_REQUIRE(!(enable && c.nh.on), MDX_ERR_STATE, "Nose-Hoover thermostat is on");
        c.thermoOn = enable != 0;
        c.thermoT = temperature;
        c.thermoEvery = every;
        c.thermoTau = tau;
        c.tconv = mvv_to_eng / (degrees_of_freedom(c) * boltz);
    });
}

int mdx_nose_hoover_set(mdx_ctx *ctx, int enable, double temperature, double time_constant, int chain_length,
                        double boltz, double mvv_to_eng) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(!enable || (temperature > 0 && time_constant > 0 && chain_length >= 2 &&
                                chain_length <= 64 && boltz > 0 && mvv_to_eng > 0),
                    MDX_ERR_ARG, "Nose-Hoover needs T > 0, time constant > 0, 2 <= chain <= 64, units > 0");
        MDX_REQUIRE(!(enable && c.thermoOn), MDX_ERR_STATE, "velocity-rescale thermostat is on");
        NoseHooverState &H = c.nh;
        if (enable && (!H.dev || H.chain != chain_length)) {
            if (H.dev) dfree(c, H.dev);
            H.dev = dalloc<double>(c, 4 * chain_length + 2);
        }
        H.on = enable != 0;
        H.fresh = true;
        H.chain = chain_length;
        H.T = temperature;
        H.freq = 1.0 / time_constant;
        H.boltz = boltz;
        H.mvv = mvv_to_eng;
        H.ndf = degrees_of_freedom(c);
    });
}

int mdx_nose_hoover_energy(mdx_ctx *ctx, double *chain_energy) {
    return guarded_ro(ctx, [&](Ctx &c) {
        NoseHooverState &H = c.nh;
        MDX_REQUIRE(H.on && H.dev && chain_energy, MDX_ERR_STATE, "Nose-Hoover is off");
        std::vector<double> st(4 * H.chain + 2);
        MDX_HIP(hipMemcpyAsync(st.data(), H.dev, sizeof(double) * st.size(), hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        const int M = H.chain;
        const double kT = H.boltz * H.T;
        double e = 0.0;
        for (int k = 0; k < M; ++k) {
            e += 0.5 * st[2 * M + k] * st[k] * st[k];
            e += (k == 0 ? H.ndf : 1.0) * kT * st[3 * M + k];
        }
        *chain_energy = H.fresh ? 0.0 : e;
    });
}

int mdx_nose_hoover_state(mdx_ctx *ctx, double *state, int n) {
    return guarded_ro(ctx, [&](Ctx &c) {
        NoseHooverState &H = c.nh;
        MDX_REQUIRE(H.on && H.dev && state, MDX_ERR_STATE, "Nose-Hoover is off");
        MDX_REQUIRE(n >= 4 * H.chain + 2, MDX_ERR_ARG, "state needs 4 chain_length + 2 doubles");
        MDX_HIP(hipMemcpyAsync(state, H.dev, sizeof(double) * (4 * H.chain + 2), hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
    });
}

int mdx_barostat_set(mdx_ctx *ctx, int enable, double pressure, double period, int every, double bulk_modulus,
                     double max_dilation, double nktv_to_press, double mvv_to_eng) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(!enable || (period > 0 && every >= 1 && bulk_modulus > 0 && max_dilation > 0 &&
                                max_dilation < 0.1 && nktv_to_press > 0 && mvv_to_eng > 0),
                    MDX_ERR_ARG, "barostat needs period > 0, every >= 1, bulk modulus > 0, 0 < max dilation < 0.1");
        BarostatState &B = c.baro;
        B.on = enable != 0;
        B.P0 = pressure;
        B.period = period;
        B.every = every;
        B.bulk = bulk_modulus;
        B.maxDil = max_dilation;
        B.nktv2p = nktv_to_press;
        B.mvv = mvv_to_eng;
    });
}

int mdx_pressure(mdx_ctx *ctx, int flags, double nktv_to_press, double mvv_to_eng, double *pressure) {
    return guarded(ctx, [&](Ctx &c) {
        require_system(c);
        MDX_REQUIRE(pressure, MDX_ERR_ARG, "no output");
        prepare_terms(c);
        ensure_list(c);
        MDX_HIP(hipStreamSynchronize(c.stream));
        if (!capacity_ok(c)) nlist_rebuild(c, true);
        compute(c, (flags & kForceTerms) | MDX_VIRIAL);
        const double k = c.baro.nktv2p, m = c.baro.mvv;
        c.baro.nktv2p = nktv_to_press;
        c.baro.mvv = mvv_to_eng;
        *pressure = pressure_now(c);
        c.baro.nktv2p = k;
        c.baro.mvv = m;
    });
}

int mdx_barostat_last(mdx_ctx *ctx, double *pressure) {
    return guarded_ro(ctx, [&](Ctx &c) {
        MDX_REQUIRE(pressure, MDX_ERR_ARG, "no output");
        *pressure = c.baro.lastP;
    });
}

int mdx_temperature(mdx_ctx *ctx, double boltz, double mvv_to_eng, double *temperature,
                    double *kinetic) {
    return guarded_ro(ctx, [&](Ctx &c) {
        require_system(c);
        thermo_measure(c);
        unsigned long long h = 0;
        MDX_HIP(hipMemcpyAsync(&h, c.thermo, sizeof(h), hipMemcpyDeviceToHost, c.stream));
        MDX_HIP(hipStreamSynchronize(c.stream));
        const double mv2 = (double)(long long)h * kFxInv;
        if (kinetic) *kinetic = 0.5 * mv2 * mvv_to_eng;
        if (temperature) *temperature = mv2 * mvv_to_eng / (degrees_of_freedom(c) * boltz);
    });
}

int mdx_sync(mdx_ctx *ctx) {
    return guarded_ro(ctx, [&](Ctx &c) { MDX_HIP(hipStreamSynchronize(c.stream)); });
}

}  // extern "C"
