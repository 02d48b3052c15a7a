// Internal state and helpers of libmdx (not part of the C ABI).
//
// Data layout in HBM (one context = one system on one GPU):
//   * per-atom arrays are kept in cell order (columns of cells ranked along a
//     2-D Morton curve, z contiguous inside a column); the order is rebuilt at
//     every neighbor-list rebuild and idToIdx maps atom id -> current index;
//   * positions and velocities are FP64 (double4, w unused / inverse mass);
//   * xq is an FP32 copy of (x - lo, q) refreshed every step for the pair,
//     E3B and mesh kernels;
//   * forces, energies and virials accumulate as 64-bit fixed point
//     (value * 2^32) through integer atomics, so the result does not depend on
//     the order in which kernels or lanes add their contributions.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdint>
#include <string>
#include <cmath>
#include <numeric>
#include <vector>

#include "mdx_launch.h"

namespace mdx {

constexpr double kFxScale = 4294967296.0;         // 2^32, forces / energies / virials
constexpr double kFxInv = 1.0 / 4294967296.0;
constexpr double kMeshScale = 1099511627776.0;    // 2^40, P3M charge mesh
constexpr double kMeshInv = 1.0 / 1099511627776.0;
constexpr int kMaxTypes = 16;

// accumulator slots (fixed point, kFxScale)
enum AccSlot {
    ACC_PAIR = 0,     // LJ + Ewald real-space energy
    ACC_KSPACE = 1,   // P3M reciprocal-space energy (without self term)
    ACC_E3B2 = 2,     // E3B two-body energy
    ACC_E3B3 = 3,     // E3B three-body energy
    ACC_VIR = 4,      // 6 slots: xx yy zz xy xz yz
    ACC_MOL = 10,     // 2 slots: molecular virial correction, sum M V^2 of centres of mass
    ACC_BOND = 12,    // intramolecular (quartic bond + harmonic angle) energy
    ACC_VTRACE = 13,  // virial trace of terms evaluated in trace-only mode (barostat steps)
    ACC_N = 16
};
// The accumulator block is replicated kAccRep times; a workgroup adds into
// replica (blockIdx mod kAccRep) and readers sum the replicas (integer sums
// commute, so the total is exact and order-free).  Same-address device
// atomics serialise across the XCDs; spreading them keeps the energy/virial
// kernels from queueing on one address.
constexpr int kAccRep = 32;
constexpr int kAccTotal = ACC_N * kAccRep;

struct Geometry {
    double lo[3];
    double L[3];
    double invL[3];
    int periodic[3];
};

// ---- P3M splitting parameter (find_optimal_parameters, FixChargeEwald.cu:
// 174-217 and 935-967): secant solve of DeltaF_k(alpha) = DeltaF_real(alpha)
// in FP64, with the reference's start 4.79853 / rcut, tolerance 1e-5 and
// 100-iteration limit.  Host and device run the same code (the device one
// after a barostat changed the box inside a captured step).
__host__ __device__ inline double pppm_amp(int order, int m) {
    constexpr double t[7][7] = {
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
    return t[order - 1][m];
}

__host__ __device__ inline double deltaf_k(double a, int natoms, double q2, const double *L, const int *sz,
                                           int order) {
    const double h[3] = {L[0] / sz[0], L[1] / sz[1], L[2] / sz[2]};
    double sum[3] = {0, 0, 0};
    for (int mm = 0; mm < order; ++mm) {
        const double amp = pppm_amp(order, mm);
        for (int d = 0; d < 3; ++d) sum[d] += amp * pow(h[d] * a, (double)(2 * mm));
    }
    double acc = 0.0;
    for (int d = 0; d < 3; ++d)
        acc += 1.0 / (L[d] * L[d]) * pow(a * h[d], (double)order) *
               sqrt(a * L[d] / natoms * sqrt(2.0 * M_PI) * sum[d]);
    return q2 / 3.0 * acc;
}

__host__ __device__ inline double deltaf_real(double a, int natoms, double q2, const double *L, double rcut) {
    return 2 * q2 / sqrt(natoms * rcut * L[0] * L[1] * L[2]) * exp(-a * a * rcut * rcut);
}

__host__ __device__ inline double pppm_solve_alpha(int natoms, double q2, const double *L, const int *sz,
                                                   int order, double rcut, double *err) {
    double xa = 0.0, xb = 4.79853 / rcut;
    double ya = deltaf_k(xa, natoms, q2, L, sz, order) - deltaf_real(xa, natoms, q2, L, rcut);
    double yb = deltaf_k(xb, natoms, q2, L, sz, order) - deltaf_real(xb, natoms, q2, L, rcut);
    const double tol = 1e-5;
    int it = 0;
    while ((fabs(yb) / deltaf_real(xb, natoms, q2, L, rcut) > tol) && it < 100) {
        const double kinv = (xb - xa) / (yb - ya);
        ya = yb;
        xa = xb;
        xb = xa - ya * kinv;
        yb = deltaf_k(xb, natoms, q2, L, sz, order) - deltaf_real(xb, natoms, q2, L, rcut);
        ++it;
    }
    if (err) *err = deltaf_k(xb, natoms, q2, L, sz, order) + deltaf_real(xb, natoms, q2, L, rcut);
    return xb;
}

// alias range of the Green function's image sums (calc_Green_function,
// FixChargeEwald.cu:314-434)
__host__ __device__ inline int green_lim(double alpha, double hmean) {
    return (int)(alpha * hmean / 3.14159 * (sqrt(-log(10E-7)))) + 1;
}

// ---- box-dependent launch data, kept in device memory ----------------------
// Every kernel reads what depends on the box from one BoxDev in device
// memory, so a captured graph stays valid while a barostat changes the box on
// the device.  box_derive computes all of it from (lo, L) and the parameters
// fixed while a cell grid lives, with the same expressions on host and device.
struct CellGrid {
    int nc[3];
    double cellInv[3];  // nc / L
    int reach[3];       // stencil half-width in cells per axis
};

// FP32 pre-test on xq (x - lo in FP32): candidates with r^2 clearly inside or
// outside r_list^2 (by more than `band`, a bound on the FP32 rounding of the
// positions and of the distance) are decided in FP32; only the few inside the
// band take the canonical FP64 test, so membership is unchanged.
struct Prefilter {
    float L[3], invL[3];
    float lo2, hi2;  // rlist^2 -+ band
};

// Per-atom culling of the stencil: column (a, b) is skipped when its xy
// distance from the atom exceeds r_list, and its z-run is cut to the cells
// within sqrt(r_list^2 - d_xy^2) of the atom.  Distances are measured to the
// unwrapped cell (cx - reach + a) * w: with nc >= 2 reach + 1 each wrapped
// cell appears once in the stencil and every neighbour's minimum image lies in
// its unwrapped cell.  The margin in r2 covers the FP32 rounding of the
// coordinates and of the cell bounds.  Axes that are open or hold fewer than
// 2 reach + 1 cells are not culled.
struct Cull {
    float w[3];
    float r2;
    int on[3];
};

struct MeshGeo {
    int sz[3];
    double lo[3], L[3], invL[3], h[3];
    int periodic[3];
};

// Nearest P3M mesh point and offset, FP64 like the oracle: the coordinate is
// wrapped into [lo, lo + L) first, then ngp = int((pos + h/2) / h) and
// d = pos / h - ngp (FixChargeEwald.cu:178-190).  ngp lies in [0, sz].
__host__ __device__ __forceinline__ void mesh_point(double x, int a, const MeshGeo &m, int &ngp, float &d) {
#pragma clang fp contract(off)
    double img = floor((x - m.lo[a]) * m.invL[a]);
    double xw = x - img * m.L[a];
    double p = xw - m.lo[a];
    ngp = (int)((p + 0.5 * m.h[a]) / m.h[a]);
    d = (float)(p / m.h[a] - ngp);
}

// Per charged site, written with the force pass's position refresh: the three
// offsets d and the nearest mesh point wrapped into [0, sz) per axis, packed
// 10 bits an axis in w (meshes up to 1024 points an axis).
__device__ __forceinline__ float4 mesh_record(const double4 &p, const MeshGeo &m) {
    int g[3];
    float d[3];
    mesh_point(p.x, 0, m, g[0], d[0]);
    mesh_point(p.y, 1, m, g[1], d[1]);
    mesh_point(p.z, 2, m, g[2], d[2]);
    unsigned packed = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) packed |= (unsigned)(g[a] >= m.sz[a] ? g[a] - m.sz[a] : g[a]) << (10 * a);
    return make_float4(d[0], d[1], d[2], __uint_as_float(packed));
}

__device__ __forceinline__ int rec_point(const float4 &r, int a) { return (int)((__float_as_uint(r.w) >> (10 * a)) & 1023u); }

struct BoxParams {  // fixed while a cell grid and a mesh live
    int periodic[3];
    int nc[3], reach[3];
    int meshSz[3];
    double rlist, skin, qqr;
};

struct BoxDev {
    Geometry geo;
    CellGrid cg;
    Prefilter pf;
    Cull cu;
    MeshGeo mesh;
    float Lf[3], invLf[3];  // FP32 minimum image (invLf = 0 on open axes)
    double vol;
    double alpha;           // Ewald splitting parameter (FP64) and its FP32 copy
    float alphaF;
    float gatherScale;      // -sqrt(qqr) / V, the P3M field-to-force factor
    double tileCellW[3], tileMargin;
    int greenLim;           // alias range of the Green function sums
    double mu;              // last barostat dilation factor
    double pressure;        // last barostat pressure
    int regrid;             // a cell edge fell below rlist / reach (grid must be rebuilt)
};

__host__ __device__ inline void box_derive(BoxDev &b, const BoxParams &p, const double lo[3], const double L[3]) {
    double lmax = 0.0;
    bool open = false;
    for (int a = 0; a < 3; ++a) {
        b.geo.lo[a] = lo[a];
        b.geo.L[a] = L[a];
        b.geo.invL[a] = 1.0 / L[a];
        b.geo.periodic[a] = p.periodic[a];
        lmax = lmax > L[a] ? lmax : L[a];
        open = open || !p.periodic[a];
    }
    for (int a = 0; a < 3; ++a) {
        b.cg.nc[a] = p.nc[a];
        b.cg.reach[a] = p.reach[a];
        b.cg.cellInv[a] = p.nc[a] > 0 ? (double)p.nc[a] / L[a] : 0.0;
        b.Lf[a] = (float)L[a];
        b.invLf[a] = p.periodic[a] ? (float)(1.0 / L[a]) : 0.0f;
        b.pf.L[a] = b.Lf[a];
        b.pf.invL[a] = b.invLf[a];
        b.cu.w[a] = p.nc[a] > 0 ? (float)(L[a] / p.nc[a]) : 0.0f;
        b.cu.on[a] = p.periodic[a] && p.nc[a] >= 2 * p.reach[a] + 1;
        b.mesh.sz[a] = p.meshSz[a];
        b.mesh.lo[a] = lo[a];
        b.mesh.L[a] = L[a];
        b.mesh.invL[a] = b.geo.invL[a];
        b.mesh.h[a] = p.meshSz[a] > 0 ? L[a] / p.meshSz[a] : 0.0;
        b.mesh.periodic[a] = p.periodic[a];
        b.tileCellW[a] = p.nc[a] > 0 ? L[a] / p.nc[a] : 0.0;
    }
    // band of the FP32 pre-test: 8 ulp of the largest coordinate per axis on
    // the distance and 16 ulp on r^2 (2^-23 = FP32 epsilon)
    const double r2 = p.rlist * p.rlist;
    const double eps = 1.1920928955078125e-07;
    const double band = 8.0 * 2.0 * p.rlist * 1.01 * lmax * eps * 1.7320508075688772 + 16.0 * r2 * eps;
    b.pf.lo2 = (float)(r2 - band);
    b.pf.hi2 = (float)(r2 + band);
    const double rc = p.rlist + 1e-5 * (lmax + p.rlist);
    b.cu.r2 = (float)(rc * rc);
    if (open) {
        // an open axis leaves positions unwrapped (no bound on |x - lo|): every
        // candidate takes the FP64 test, and no axis is culled
        b.pf.lo2 = -1.0f;
        b.pf.hi2 = __builtin_inff();
        b.cu.on[0] = b.cu.on[1] = b.cu.on[2] = 0;
    }
    b.vol = L[0] * L[1] * L[2];
    b.gatherScale = (float)(-sqrt(p.qqr) / b.vol);
    b.tileMargin = 0.5 * p.skin + 1e-6 * L[0];
}

struct PppmState {
    bool on = false;
    int gatherTeam = 8;           // lanes per charged atom in the field gather
    int spreadMode = 1;           // 2 owner-computes cell blocks, 1 LDS tiles with global atomics, 0 global atomics
    int gatherMode = 2;           // 2 LDS tiles of the field, 1 row teams from the mesh records, 0 from the positions
    int tileKB = 48;              // LDS budget of a gather tile's field box (KB)
    int spreadKB = 48;            // LDS budget of a spread tile's charge box (KB)
    int tileTeam = 8;             // lanes per charged site in the tile gather
    int tilePlanar = 1;           // 1: the tile gather loads the three planar components (no interleaving pass)
    int tileCols = 1;             // cell columns per tile side in the tiled spread and gather (1 or 2)
    int fftMode = 1;              // 1: the mesh solve in three LDS passes where the mesh allows, 0: hipFFT
    int lastSpread = -1, lastGather = -1, lastFft = -1;  // paths the last evaluation took (mdx_pppm_paths)
    int sz[3] = {64, 64, 64};
    int order = 5;
    double rcut = 9.0;
    double qqr = 332.06371;
    double alpha = 0.0;
    double alphaUser = -1.0;   // > 0: fixed alpha, else secant solve
    int alphaAtoms = 0;        // > 0: atom count of the secant solve (a ring's beads x atoms), else n
    double totalQ2 = 0.0;
    double err = 0.0;
    double boxAtGreen[3] = {0, 0, 0};
    long long *mesh = nullptr;       // sx*sy*sz fixed point
    float *real = nullptr;           // sx*sy*sz
    float2 *rhoHat = nullptr;        // sx*sy*(sz/2+1)
    float2 *eHat = nullptr;          // 3 * sx*sy*(sz/2+1)
    float *eReal = nullptr;          // 3 * sx*sy*sz
    float4 *eField4 = nullptr;       // the field per mesh point (Ex, Ey, Ez, 0)
    float *green = nullptr;          // sx*sy*(sz/2+1)
    hipfftHandle planR2C = 0;
    hipStream_t planStream = nullptr;  // stream the plans are bound to
    hipfftHandle planC2R = 0;
    bool plansMade = false;
    // FixChargeEwald::setError: the mesh is the smallest the doubling search
    // of setGridToErrorTolerance finds for this RMS force error (0: off)
    double errorTarget = 0.0;
    // FixChargeEwald::setLongRangeInterval: the mesh part runs every
    // `interval` steps counted from the run's first; the steps between add the
    // forces it stored by atom id (fixed point, 3 per id)
    int interval = 1;
    long long turnInit = 0;
    long long *stored = nullptr;
    bool storedValid = false;
};

// Flexible water: quartic O-H bonds, harmonic H-O-H angle and a massless M
// site at gamma r_O + (1 - gamma)/2 (r_H1 + r_H2) (q-TIP4P/F)
struct FlexState {
    bool on = false;
    int nmol = 0;
    int4 *ids = nullptr;    // by molecule: O, H1, H2, M ids
    int4 *atoms = nullptr;  // by molecule: current indices
    double k2 = 0, k3 = 0, k4 = 0, r0 = 0, ktheta = 0, theta0 = 0, gamma = 0;
};

struct E3BState {
    bool on = false;
    int nmol = 0;
    float E2 = 0, Ea = 0, Eb = 0, Ec = 0, k2 = 0, k3 = 0, rs = 0, rf = 0;
    float r2b = 6.2f;       // two-body cutoff (molecule-list radius of the reference)
    float rmax = 6.1572f;   // three-body pair test
    int4 *molIds = nullptr;   // by molecule: atom ids (O, H1, H2, M)
    int4 *molAtoms = nullptr; // by molecule: current atom indices
    int *atomMolO = nullptr;  // by current atom index: molecule index if the atom is an E3B oxygen, else -1
    int *molIdsOfO = nullptr; // by atom id: molecule index if oxygen, else -1
    int cap = 0;
    int *nl = nullptr;        // nmol * cap molecule indices
    int *nlCount = nullptr;
    float4 *tot = nullptr;    // per molecule totals (x accepted, y H1 donated, z H2 donated)
    int team = 32;            // lanes per molecule in the E3B passes
    int listTeam = 8;         // lanes per molecule building the molecule list
    bool oxUncharged = false;  // every E3B oxygen site has q = 0 (its neighbors need no charged scan)
    std::vector<int> oxIds;    // by molecule: oxygen atom id
};

// Nose-Hoover chain thermostat state, kept on the device so a captured step
// needs no host round trip: dev = [vel M | force M | mass M | pos M | ke | scale]
struct NoseHooverState {
    bool on = false, fresh = true;
    int chain = 10;
    double T = 0.0, freq = 0.0, boltz = 0.0, mvv = 0.0, ndf = 0.0;
    double *dev = nullptr;
};

// Berendsen barostat (FixPressureBerendsen)
struct BarostatState {
    bool on = false;
    double P0 = 1.0, period = 1000.0, bulk = 10.0, maxDil = 1e-5, nktv2p = 68568.415, mvv = 1.0;
    int every = 10;
    double lastP = 0.0;
};

// learned E3B three-body surrogate (surrogate.hip)
// surrogate sizes: radial basis functions, raw channels per molecule (acc,
// d1, d2 per basis + three pair products), network inputs, hidden width
constexpr int kSurNB = 2, kSurNR = 3 * kSurNB + 3, kSurNF = 3 * kSurNB + 6, kSurNH = 64;
constexpr int kSurNFP = (kSurNF + 15) / 16 * 16;  // network inputs padded to whole MFMA k-blocks
constexpr int kSurNW = (kSurNR + 3) / 4 * 4;  // row stride of the per-molecule weights (16-byte rows)
constexpr int kSurNParam = kSurNF * kSurNH + kSurNH + kSurNH * kSurNH + kSurNH + kSurNH + 1 + 2 * kSurNF + 2;

struct SurrogateState {
    bool on = false;
    float *params = nullptr;  // packed model (kSurNParam)
    float *feat = nullptr;    // per molecule raw channels (kSurNR)
    float *emol = nullptr;    // per molecule energy
    float *wgt = nullptr;     // per molecule dE / d(raw channel) (kSurNR, row stride kSurNW)
    unsigned short *frag = nullptr;  // BF16 MFMA operand fragments of W1, W2 (both orientations)
    int nmolAlloc = 0;
    float *trainX = nullptr;  // recorded samples: raw channels
    float *trainT = nullptr;  // recorded samples: E3B three-body energy of the molecule
    int nsamp = 0, capSamp = 0;
    float *part = nullptr, *m1 = nullptr, *m2 = nullptr, *loss = nullptr;
    long step = 0;
    int forceLanes = 16;      // lanes per molecule in the force pass (8, 16, 32, 64)
    // E3B as judge: every judgeEvery steps of a run with MDX_SURROGATE the
    // three-body forces of both forms are compared on the current frame; an
    // RMSE above judgeMax makes the rest of the run (and later runs) use E3B
    int judgeEvery = 0;
    double judgeMax = 0.0;
    long long judgeChecks = 0;
    double judgeLast = 0.0, judgeWorst = 0.0;
    bool fellBack = false;
    unsigned long long *fSave = nullptr, *fE3B = nullptr;  // 3 n each
    double *judgeSum = nullptr;                             // device scalar
    int judgeN = 0;                                         // atoms the buffers hold
};

// kernels timed with HIP events when profiling is on (mdx_profile_enable)
enum ProfKernel {
    PK_BUILD = 0, PK_PAIR, PK_SPREAD, PK_GATHER, PK_E3B1, PK_E3B2, PK_FFT, PK_EFIELD,
    PK_SUR_FEAT, PK_SUR_MLP, PK_SUR_FORCE, PK_N
};

struct ProfRec {
    int kernel;
    hipEvent_t a, b;
};

struct RigidState {
    bool on = false;
    bool fresh = false;       // velocities not yet constrained (start of the next run)
    int nmol = 0;
    int nsites = 0;           // sites in rigid molecules (3 or 4 per molecule)
    int4 *ids = nullptr;      // by molecule: O, H1, H2, M ids (M = -1: three sites)
    int4 *atoms = nullptr;    // by molecule: current indices
    int *member = nullptr;    // by current index: 1 for sites of a rigid molecule
    double mO = 0, mH = 0, dOH = 0, dHH = 0, gamma = 0;
};

// compute flag (internal): the virial's trace is enough (a barostat step);
// the pair and E3B kernels then add r . F into ACC_VTRACE only
constexpr int kFlagVirialTrace = 1 << 12;  // MDX_VIRIAL_TRACE of the C ABI

// Device memory and FFT plans that a merged batch graph uses, released with it.
struct BatchKeep {
    std::vector<void *> mem;
    std::vector<hipfftHandle> plans;
    bool empty() const { return mem.empty() && plans.empty(); }
    void release() {
        for (void *p : mem) hipFree(p);
        for (hipfftHandle h : plans) hipfftDestroy(h);
        mem.clear();
        plans.clear();
    }
};

struct Ctx {
    int device = 0;
    bool noAlloc = false;  // inside mdx_forces_compute_noalloc: any device allocation is an error
    bool inRun = false;    // inside mdx_run's step loop (the long-range interval applies)
    // Captured step graphs are kept across mdx_run calls while nothing they
    // bake in changed: serial identifies the context for its lifetime,
    // apiEpoch counts calls that may change the set-up, layoutEpoch counts
    // device (re)allocations and cell-grid set-ups (kernel arguments).
    uint64_t serial = 0;
    uint64_t apiEpoch = 0;
    uint64_t layoutEpoch = 0;
    struct GraphCache {
        hipGraphExec_t big = nullptr, one = nullptr;
        std::vector<uint64_t> key;
        BatchKeep keep;  // what merged batch launches use: tables, block maps, grouped FFTs
    } graphs;
    bool batchMerge = true;      // option batch_merge: one launch per kernel across a batch
    int batchMerged = 0;         // the last batch capture merged its replicas' launches (1) or not (0)
    bool vtrace = false;  // set by compute for the terms it launches
    bool fuseMolVirial = false;  // captured barostat steps: rigid second half sums ACC_MOL
    bool fuseThermo = false;     // captured thermostat steps: rigid second half sums m v^2
    bool fuseCheck = false;      // captured steps: rigid first half gives the skin check's maxima
    bool deferMSpread = false;   // compute leaves the M-site forces for a later flex_spread_m (ring contraction)
    hipStream_t stream = nullptr;
    // second stream of captured steps: the P3M chain runs beside the pair and
    // E3B kernels (fork after the position refresh, join before the kick)
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    bool meshStream = true;       // option "mesh_stream"
    std::string err;

    // system
    int n = 0;
    Geometry geo{};
    double4 *pos[2] = {nullptr, nullptr};
    double4 *vel[2] = {nullptr, nullptr};
    float *q[2] = {nullptr, nullptr};
    int *type[2] = {nullptr, nullptr};
    int *id[2] = {nullptr, nullptr};
    int cur = 0;
    int *idToIdx = nullptr;
    float4 *xq = nullptr;
    float4 *meshRec = nullptr;             // per charged site: mesh offsets and point (mesh_record)
    unsigned long long *force = nullptr;   // 3 per atom
    unsigned long long *acc = nullptr;     // kAccTotal: kAccRep replicas of ACC_N slots
    BoxDev *box = nullptr;                 // device copy of the box-dependent launch data
    BoxDev *boxLast = nullptr;             // the box of the last list build (copied by k_scan)
    int replayChunk = 0;                   // option replay_chunk: chunk of each run replayed once
    std::vector<LaunchRec> *rec = nullptr; // set while a batch records this context's step launches
    BoxDev boxHost;                        // host mirror (last upload or read-back)
    double4 *posLast = nullptr;
    std::vector<double> invMassHost;
    std::vector<float> qHost;             // charges by id, as last set
    std::vector<int> typeHost;            // types by id
    std::vector<double> ljEps;            // 24 eps per type pair (host copy)
    bool pairFlagsDirty = true;
    int *pairTypeFlags = nullptr;         // per type: bit 0 LJ with a charged type, bit 1 with an uncharged one       // by id
    bool hasCharges = false;
    int nCharged = 0;             // atoms with q != 0: indices [0, nCharged) once a list is built
    int maxType = 0;

    // neighbor list
    double rlist = 0.0, skin = 0.0;
    bool nlistConfigured = false;
    int nc[3] = {0, 0, 0};
    // cell stencil half-width per axis (cell edge >= rlist / reach); thinner z
    // cells cut the z-runs closer to the sphere
    int reach[3] = {2, 2, 4};
    int ncell = 0;
    int *colRank = nullptr;
    int *cellCount = nullptr;
    int *cellStart = nullptr;     // 2 ncell + 1: cells of the charged atoms, then of the charge-free ones
    int *cellFill = nullptr;
    int *cellOf = nullptr;
    int *perm = nullptr;
    int *perm2 = nullptr;
    int2 *colXY = nullptr;        // by column rank: cell (x, y)
    int cap = 0;
    uint32_t *nl = nullptr;
    int *nlCountA = nullptr;      // leading (charged-neighbor) part of each row; charge-free partners follow
    int *nlCount = nullptr;
    // [0] rebuild flag, [1] max count, [2] e3b max count; captured steps:
    // [3] rebuild gate, [4] rebuilds, [5] thermostat gate, [6] turn;
    // [7] bit 0: an atom's candidate stream overflowed the build table, bit 1:
    // a charged site's P3M stencil left the cells next to its own;
    // [8] barostat gate; [10] last-workgroup ticket of the first half kick
    int *dFlags = nullptr;
    int *waveMax = nullptr;       // per-wave longest rows: [0, n + 4) pair list, then the molecule list
    int waveMaxN = 0, waveMaxA = 0, waveMaxB = 0;
    bool capture = false;         // a step is being recorded into a graph
    bool rowMaxPending = false;   // a captured build's row maxima go with the next position refresh
    bool decided = false;         // the first half kick of this captured step took the step's decisions
    bool refreshed = false;       // ... and refreshed xq, the mesh records, forces and accumulators
    bool forked = false;          // the captured rebuild recorded the mesh stream's fork after its copy-back
    bool baroFused = false;       // the rigid second half of this captured step ran the barostat step
    bool fuseStep = true;         // option "fuse_step": captured steps fold small launches into others
    int graphSteps = 0;           // mdx_run: 1 = captured steps (option "graph")
    unsigned *dMax = nullptr;     // 2 slots x (max |x - x_last|^2, max predicted^2) as float bits
    float *checkPart = nullptr;   // per-workgroup maxima of the displacement check (2 x kCheckBlocks)
    int *pinned = nullptr;        // host-pinned: [0..1] last build's max counts, [2..5] dMax copies, [6] stream overflow
    hipEvent_t capEvent = nullptr;
    bool capPending = false;
    long long listViolations = 0; // steps whose list was older than the skin allows
    int buildTeam = 64;           // lanes per atom in the list build (64: one wave per atom)
    bool buildTeamUser = false;   // set by mdx_teams_set
    int buildMode = 0;            // 0 (default) and 2 per-atom teams
    long long nRebuilds = 0;
    bool listValid = false;
    // Ring polymers with one list on the centroids (GridGPU.cu:163-189,
    // 888-895): while a ring call runs, the build bins and pairs atoms by the
    // centroid of their ring (cenById, by atom id) instead of their own
    // position, in buildPos (the centroids in this bead's order); the skin
    // check's partial maxima of every bead go to one ring table (ringPart,
    // ringPartNb workgroups) so every bead takes the same rebuild decision.
    const double4 *cenById = nullptr;
    double4 *buildPos = nullptr;
    const float *ringPart = nullptr;
    int ringPartNb = 0;
    bool centroidList = false;    // the current list was built on ring centroids

    // exclusions (CSR by id, entries id | tag << 30)
    int *exclOff = nullptr;
    uint32_t *exclIds = nullptr;
    int nExclIds = 0;
    int maxExcl = 0;
    int exclSpan = 0;             // largest |id_i - id_j| over excluded pairs
    uint32_t *exclIdx = nullptr;  // per atom: maxExcl partners as current index | tag << 30
    size_t exclIdxN = 0;

    // LJ + Ewald real space
    bool ljOn = false;
    int ntypes = 1;
    float *ljParams = nullptr;    // 3 * T*T: rc2, eps24, sig6
    float special[3] = {0.0f, 0.0f, 0.5f};
    bool ewaldReal = false;
    float alpha = 0.0f, qqr = 332.06371f, rq2 = 0.0f;
    int pairTeam = 16;

    PppmState pppm;
    E3BState e3b;
    FlexState flex;
    SurrogateState sur;
    NoseHooverState nh;
    BarostatState baro;
    RigidState rigid;
    bool mdxDisableTiledSpread = false;   // test knob: force the plain global-atomic spread

    // thermostat (velocity rescaling every thermoEvery steps)
    bool thermoOn = false;
    double thermoT = 0.0, thermoTau = 0.0;
    int thermoEvery = 1;
    double tconv = 0.0;                      // temperature per unit of sum m v^2
    unsigned long long *thermo = nullptr;    // [0] sum m v^2 (fixed point)

    // integration
    double dt = 0.005;
    double ftmToV = 1.0;
    long long turn = 0;
    int forceFlags = 0;

    hipEvent_t stepEvent[2] = {nullptr, nullptr};
    struct Snapshot {
        double4 *pos = nullptr, *vel = nullptr;
        float *q = nullptr;
        int *type = nullptr, *id = nullptr, *idToIdx = nullptr;
        long long turn = 0, rebuilds = 0;
        Geometry geo;
        BoxDev *box = nullptr;
        // the list the chunk started with, when one carried into it: the
        // positions and box of its build (a replay rebuilds it from them)
        bool listValid = false;
        double4 *posLast = nullptr;
        BoxDev *boxLast = nullptr;
        double *nh = nullptr;                 // Nose-Hoover chain state (4 chain + 2 doubles)
        int nhLen = 0;
        unsigned long long *thermo = nullptr; // thermostat sums (2 words)
        long long *pppmStored = nullptr;      // stored mesh forces (long-range interval > 1)
        bool pppmStoredValid = false;
    } snap;

    // profiling: event pairs recorded on the context stream around kernels
    bool prof = false;
    std::vector<ProfRec> profRecs;
    std::vector<hipEvent_t> eventPool;

    // all device allocations, freed in mdx_destroy
    std::vector<void *> allocs;
};

// Records a HIP event pair around the launches in its scope when profiling is on.
struct ProfScope {
    Ctx &c;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(Ctx &c_, int k_);
    ~ProfScope();
};

// Launches a per-step kernel body for this context, or records it when a
// batch is being put together (see mdx_launch.h).  The arguments are
// converted to the body's parameter types first, so every context of a batch
// records the same kernel whatever the argument expressions' types.
template <auto BODY, int NT, class... P>
void launch_typed(Ctx &c, int blocks, size_t shm, P... a) {
    if (c.rec) {
        LaunchRec r;
        r.kid = reinterpret_cast<const void *>(&k_many<BODY, NT, P...>);
        r.blocks = blocks;
        r.shm = shm;
        ArgPack<P...> p;
        pack_fill(p, a...);
        r.args.resize(sizeof(p));
        std::memcpy(r.args.data(), &p, sizeof(p));
        r.align = alignof(ArgPack<P...>);
        r.many = &launch_many<BODY, NT, P...>;
        c.rec->push_back(std::move(r));
        return;
    }
    if (blocks > 0) hipLaunchKernelGGL((k_one<BODY, NT, P...>), dim3(blocks), dim3(NT), shm, c.stream, a...);
}

template <class F>
struct BodyParams;
template <class... P>
struct BodyParams<void (*)(int, int, P...)> {
    template <auto BODY, int NT, class... A>
    static void launch(Ctx &c, int blocks, size_t shm, A... a) {
        launch_typed<BODY, NT, P...>(c, blocks, shm, static_cast<P>(a)...);
    }
    template <auto BODY, int NT, class... A>
    static void launch_on(hipStream_t s, int blocks, size_t shm, A... a) {
        if (blocks > 0) hipLaunchKernelGGL((k_one<BODY, NT, P...>), dim3(blocks), dim3(NT), shm, s, static_cast<P>(a)...);
    }
};

template <auto BODY, int NT, class... A>
void launch(Ctx &c, int blocks, size_t shm, A... a) {
    BodyParams<decltype(BODY)>::template launch<BODY, NT>(c, blocks, shm, a...);
}

// A per-step kernel body on a given stream, never recorded (launches that
// serve several contexts at once, such as the ring polymer's).
template <auto BODY, int NT, class... A>
void launch_on(hipStream_t s, int blocks, size_t shm, A... a) {
    BodyParams<decltype(BODY)>::template launch_on<BODY, NT>(s, blocks, shm, a...);
}

// A stream operation that is not a kernel of the launch layer (an FFT):
// run now, or recorded in order with the kernels of a batch.  fft (1 forward
// R2C of the charge mesh, 2 inverse C2R of the three field components) lets a
// merged batch run the transforms of replicas with equal meshes as one.
inline void stream_op(Ctx &c, std::function<void(hipStream_t)> f, int fft = 0) {
    if (c.rec) {
        LaunchRec r;
        r.host = std::move(f);
        r.fft = fft;
        c.rec->push_back(std::move(r));
        return;
    }
    f(c.stream);
}

// error helpers -------------------------------------------------------------
struct Error {
    int code;
    std::string msg;
};

#define MDX_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::mdx::Error{2, std::string(#call) + ": " + hipGetErrorString(e_)};   \
    } while (0)

#define MDX_REQUIRE(cond, code, msg)                                                     \
    do {                                                                                 \
        if (!(cond)) throw ::mdx::Error{(code), (msg)};                                  \
    } while (0)

template <class T>
T *dalloc(Ctx &c, size_t count) {
    MDX_REQUIRE(!c.noAlloc, 3,
                "this call would allocate device memory (first list build, grown rows, new buffers): "
                "call mdx_forces_compute once to set the context up");
    void *p = nullptr;
    if (count == 0) count = 1;
    ++c.layoutEpoch;
    MDX_HIP(hipMalloc(&p, count * sizeof(T)));
    c.allocs.push_back(p);
    return static_cast<T *>(p);
}

void dfree(Ctx &c, void *p);

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// device helpers --------------------------------------------------------------
__device__ __forceinline__ long long to_fx(double v) { return __double2ll_rn(v * kFxScale); }

__device__ __forceinline__ void fx_add(unsigned long long *p, double v) {
    atomicAdd(p, (unsigned long long)to_fx(v));
}

// this workgroup's replica of the accumulator block
__device__ __forceinline__ unsigned long long *acc_rep(unsigned long long *acc) {
    return acc + (blockIdx.x & (kAccRep - 1)) * ACC_N;
}

// ---- Berendsen barostat step on the device (FixPressureBerendsen::stepFinal,
// src/Fixes/FixPressureBerendsen.cu:20-51; alpha as handleBoundsChange,
// src/Fixes/FixChargeEwald.cu:1200-1229), run by one wave: its own launch or
// the last workgroup of the rigid second half.
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

__device__ inline double alpha_solve_any(int order, int natoms, double q2, const double *L, const int *sz, double rcut) {
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
__device__ __forceinline__ void baro_wave(const int *__restrict__ gate, BoxDev *__restrict__ bx, const BoxParams &bp,
                                          const BaroArgs &ba, const unsigned long long *__restrict__ acc) {
    const int lane = threadIdx.x & 63;
    // every input is read up front, the gate among them, so the wave waits
    // for one memory round trip rather than one per dependent read
    constexpr int kSlots[6] = {ACC_VIR, ACC_VIR + 1, ACC_VIR + 2, ACC_MOL, ACC_MOL + 1, ACC_VTRACE};
    const int open = gate ? *gate : 1;
    long long t[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) t[k] = lane < kAccRep ? (long long)acc[lane * ACC_N + kSlots[k]] : 0ll;
    const int topen = ba.tgate ? *ba.tgate : 1;
    const long long th = ba.rescaled ? (long long)ba.thermo[0] : 0ll;
    const double vol = bx->vol;
    double glo[3], gL[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        glo[d] = bx->geo.lo[d];
        gL[d] = bx->geo.L[d];
    }
    if (open == 0) return;
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t[k] += __shfl_xor(t[k], o);
    if (lane != 0) return;
    // trace of the virial: the tensor's diagonal plus the trace-only terms
    const double W = (double)t[0] * kFxInv + (double)t[1] * kFxInv + (double)t[2] * kFxInv + (double)t[5] * kFxInv;
    const double corr = (double)t[3] * kFxInv;
    double mv2 = (double)t[4] * kFxInv;
    if (ba.rescaled && topen) {
        const double T = (double)th * kFxInv * ba.tconv;
        if (T > 0.0) {
            const double l2 = ba.tmode == 0 ? ba.T0 / T : 1.0 + ba.coupling * (ba.T0 / T - 1.0);
            mv2 *= fmax(l2, 0.0);
        }
    }
    const double P = (mv2 * ba.mvv + W - corr) / (3.0 * vol) * ba.nktv2p;
    double mu = cbrt(1.0 - ba.dt / ba.period * (ba.P0 - P) / ba.bulk);
    mu = fmin(fmax(mu, 1.0 - ba.maxDil), 1.0 + ba.maxDil);
    bx->pressure = P;
    bx->mu = mu;
    double lo[3], L[3], h[3];
    for (int d = 0; d < 3; ++d) {
        const double ctr = glo[d] + 0.5 * gL[d];
        L[d] = gL[d] * mu;
        lo[d] = ctr - 0.5 * L[d];
        h[d] = bp.meshSz[d] > 0 ? L[d] / bp.meshSz[d] : 0.0;  // box_derive's mesh spacing
        if (bp.periodic[d] && L[d] < ba.minL[d]) bx->regrid = 1;
    }
    // box_derive leaves alpha, alphaF and greenLim as they are
    box_derive(*bx, bp, lo, L);
    if (ba.alphaSolve) {
        const double a = alpha_solve_any(ba.order, ba.natoms, ba.q2, L, bp.meshSz, ba.rcut);
        bx->alpha = a;
        bx->alphaF = (float)a;
        const double hm = cbrt(h[0] * h[1] * h[2]);
        bx->greenLim = green_lim(a, hm);
    }
}


// sub-wave team sum by butterfly shuffles (T divides 64)
template <int T>
__device__ __forceinline__ float team_sum(float v) {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, T);
    return v;
}

// XCD-aware block order.  Workgroups are dispatched round-robin over the
// chip's 8 XCDs (workgroup b to XCD b mod 8), each with its own L2; kernels
// whose blocks walk spatially sorted data take logical block
// (b mod 8) * per + b / 8 so that each XCD works through one contiguous
// eighth (its L2 then holds that region's share of a shared table, not all
// of it).  The grid is rounded up to a multiple of 8; blocks whose logical
// index is past the end return.
constexpr int kXcds = 8;
inline int xcd_grid(long long nlogical) { return cdiv(nlogical, kXcds) * kXcds; }
__device__ __forceinline__ int xcd_block(int blk, int nlogical) {
    const int per = (nlogical + kXcds - 1) / kXcds;
    return (blk % kXcds) * per + blk / kXcds;
}

// Device-side step decisions of a captured step: rebuild the list iff the
// largest displacement since the last build exceeds skin/2 (the reference's
// setBuildFlag test, GridGPU.cu:755-774), count it, and set the thermostat
// and barostat gates for the steps where the turn reaches a multiple of their
// intervals; clears the displacement slots for the next step.  The maxima
// come from nb per-workgroup partials; the first wave of the calling
// workgroup reduces them and its thread 0 writes.
struct DecideArgs {
    int *ticket;                 // last-workgroup counter when fused into another launch
    const float *part;
    unsigned *dmax;
    int *flags;
    float half2;
    int every, baroEvery;
    unsigned long long *thermo;  // the rigid second half of this step sums m v^2 into it
};

__device__ __forceinline__ void part_max(int nb, const float *__restrict__ part, float &d2, float &p2) {
    d2 = 0.f;
    p2 = 0.f;
    for (int b = threadIdx.x; b < nb; b += 64) {
        d2 = fmaxf(d2, part[2 * b]);
        p2 = fmaxf(p2, part[2 * b + 1]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        d2 = fmaxf(d2, __shfl_xor(d2, o, 64));
        p2 = fmaxf(p2, __shfl_xor(p2, o, 64));
    }
}

__device__ __forceinline__ void decide_wg(int nb, const DecideArgs &d) {
    if (threadIdx.x >= 64) return;
    float d2, p2;
    part_max(nb, d.part, d2, p2);
    if (threadIdx.x != 0) return;
    int *flags = d.flags;
    const int build = d2 > d.half2 ? 1 : 0;
    flags[3] = build;
    flags[4] += build;
    const int turn = flags[6] + 1;
    flags[6] = turn;
    flags[5] = (d.every > 0 && turn % d.every == 0) ? 1 : 0;
    if (d.thermo && flags[5]) {
        d.thermo[0] = 0ull;
        d.thermo[1] = 0ull;
    }
    flags[8] = (d.baroEvery > 0 && turn % d.baroEvery == 0) ? 1 : 0;
    d.dmax[0] = 0u;
    d.dmax[1] = 0u;
}

// True in the last workgroup of a launch to get here (every thread of every
// workgroup calls it once).  Device-scope fences order each workgroup's
// earlier writes before its ticket and the last workgroup's later reads after
// it, so the last workgroup sees what every workgroup wrote; it returns the
// ticket to 0 for the next launch.
__device__ __forceinline__ bool last_block(int *ticket, int nblk) {
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(ticket, 1) == nblk - 1 ? 1 : 0;
        if (last) *ticket = 0;
    }
    __syncthreads();
    if (last) __threadfence();
    return last != 0;
}

// maximum over the 64 lanes of a wave (every lane gets it)
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Longest pair row and molecule row of a build from the per-wave slots a
// (na) and b (nb), folded into the running maxima flags[1], flags[2]: one
// workgroup of any size (at most 16 waves).
__device__ __forceinline__ void row_max_wg(int na, const int *__restrict__ a, int nb, const int *__restrict__ b,
                                           int *__restrict__ flags) {
    __shared__ int red[2][16];
    const int nt = blockDim.x;
    int ma = 0, mb = 0;
    // sixteen loads in flight per thread and pass
    constexpr int U = 16;
    for (int k0 = threadIdx.x; k0 < na; k0 += U * nt) {
        int v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = k0 + u * nt < na ? a[k0 + u * nt] : 0;
#pragma unroll
        for (int u = 0; u < U; ++u) ma = max(ma, v[u]);
    }
    for (int k0 = threadIdx.x; k0 < nb; k0 += U * nt) {
        int v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = k0 + u * nt < nb ? b[k0 + u * nt] : 0;
#pragma unroll
        for (int u = 0; u < U; ++u) mb = max(mb, v[u]);
    }
    ma = wave_max(ma);
    mb = wave_max(mb);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = ma;
        red[1][w] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < nt / 64; ++k) {
            ma = max(ma, red[0][k]);
            mb = max(mb, red[1][k]);
        }
        flags[1] = max(flags[1], ma);
        flags[2] = max(flags[2], mb);
    }
}

// the row maxima of a captured step's build, taken by one extra workgroup of
// a launch that follows the build (one launch fewer per step)
struct RowMaxArgs {
    int on;
    const int *gate;
    int na, nb;
    const int *a, *b;
    int *flags;
};


// orders LDS traffic between the lanes of one wave (no cross-wave barrier)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int T>
__device__ __forceinline__ double team_sum_d(double v) {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, T);
    return v;
}

// device words that gate kernels inside a captured step (nullptr: always run)
inline const int *rebuild_gate(const Ctx &c) { return c.capture ? c.dFlags + 3 : nullptr; }
inline const int *thermo_gate(const Ctx &c) { return c.capture ? c.dFlags + 5 : nullptr; }
inline const int *baro_gate(const Ctx &c) { return c.capture ? c.dFlags + 8 : nullptr; }
// cells are sized for a box 0.1 % smaller while a barostat runs, so that the
// grid outlives about a hundred maximal dilations
inline double grid_headroom(const Ctx &c) { return c.baro.on ? 0.999 : 1.0; }

// per-module entry points ------------------------------------------------------
// nlist.hip
void nlist_setup(Ctx &c);
void nlist_check(Ctx &c);             // sets dFlags[0] when a rebuild is due
DecideArgs decide_args(const Ctx &c);  // the step decisions of a captured step
RowMaxArgs row_max_args(Ctx &c);        // a captured build's pending row maxima (clears the request)
void row_max_flush(Ctx &c);             // launch them alone if still pending
void flag7_check(int v);              // raises the error dFlags[7] holds, if any
void nlist_rebuild(Ctx &c, bool forced);  // rebuild if dFlags[0] or forced (host reads flag)
void refresh_xq(Ctx &c);
// refresh_xq + zero_forces(c, true) in one launch
void prepare_forces(Ctx &c);
void fill_int(Ctx &c, int *p, int count, int v, const int *gate = nullptr);  // kernel fill (no memset nodes)
void nlist_rebuild_async(Ctx &c);     // rebuild now; capacity checked later (capacity_ok)
bool capacity_ok(Ctx &c);             // false: rows grew, results since the last check are invalid
bool grow_if_needed(Ctx &c);
constexpr int kCheckBlocks = 512;     // workgroups of the displacement check
void check_max(Ctx &c, int parity);   // max displacement / speed into dMax slot `parity`
void decide_step(Ctx &c);             // captured step: device rebuild / thermostat gates
void skin_check(Ctx &c);              // the displacement check's partial maxima only
void skin_decide(Ctx &c);             // the decisions from them (a ring's table when set)

// pair.hip
void pair_forces(Ctx &c, bool energy, bool virial);
// per-type bits of the pair kernel: 1 = LJ with charged atoms, 2 = LJ with
// charge-free atoms (host copy of what the kernel reads)
std::vector<int> pair_type_flag_table(const Ctx &c);

// pppm.hip
void pppm_prepare(Ctx &c);            // alpha + Green function for the current box
void pppm_forces(Ctx &c, bool energy, bool virial);
// the mesh part of one step: pppm_forces, or the stored forces between the
// steps of a long-range interval (energy and virial requests always run the mesh)
void pppm_step(Ctx &c, bool energy, bool virial);
// setGridToErrorTolerance: mesh for the error target at the current box
// (grows the mesh, reallocating when it changed)
void pppm_mesh_for_error(Ctx &c);
void pppm_release(Ctx &c);
// File-scope device scratch for the P3M mesh work arrays of
// mdx_forces_compute_noalloc: its address on `device` (looked up once per
// device), the largest mesh it holds, and a scope that points the context's
// mesh arrays at it (cleared charge mesh) for the duration of one call.
constexpr int kScratchMeshMax = 64 * 64 * 64;
void pppm_scratch_init(int device);
struct PppmScratchScope {
    Ctx &c;
    long long *mesh;
    float *real;
    float2 *rhoHat, *eHat;
    float *eReal;
    float4 *eField4;
    explicit PppmScratchScope(Ctx &c_);
    ~PppmScratchScope();
};


// e3b.hip
void e3b_forces(Ctx &c, bool energy, bool virial, bool three_body = true, float *emol = nullptr);

// surrogate.hip
void surrogate_features(Ctx &c, float *raw);
// the learned three-body term with the analytic E3B two-body term (one
// fused descriptor + network launch and the force pass); with_feat also
// writes the raw channels
void surrogate_forces(Ctx &c, bool energy, bool virial, bool with_feat = false);
void surrogate_normalize(Ctx &c);
// BF16 weight fragments of the inference kernel from the FP32 parameters
void surrogate_pack(Ctx &c);
double surrogate_fit(Ctx &c, int epochs, int batch, double lr, unsigned seed);
// summed gradient of the trained parameters over one mini-batch of recorded
// samples (rows (bstart + s) mul + add mod nsamp), loss after them, into the
// device array out (surrogate_trained_count() + 1 floats)
void surrogate_grad(Ctx &c, int bstart, int bsize, unsigned mul, unsigned add, float *out);
int surrogate_trained_count();
// RMSE over atoms and components between the E3B and the surrogate
// three-body forces at the current positions; the context's force buffer
// is left as it was (synchronises)
double surrogate_judge(Ctx &c);
void fill_float(Ctx &c, float *p, int count, float v);

// rigid.hip
void rigid_map(Ctx &c);
// acc[slot] = sum F.(r - R_com), acc[slot+1] = sum M V^2 (gate: nullptr = always)
void molecular_virial(Ctx &c, int slot, const int *gate);
// isotropic scaling by the device box's mu about its centre (molecules rigid)
void dilate(Ctx &c, const int *gate);

// Isotropic dilation about the box centre by bx->mu (the barostat kernel set
// the new box and mu): thread t < nmol moves rigid molecule t with its centre
// of mass (geometry kept), thread t < n scales site t unless it belongs to a
// rigid molecule (Mod::scaleSystem, src/Mod.cu:150-240).  The intramolecular
// minimum images (about 1 A) are the same in the old and the new box.
struct DilateArgs {
    int nmol, n;
    const int4 *atoms;
    const int *member;
    double mO, mH;
    double4 *pos;
    // the thermostat's velocity rescale of the same step, when it rides on
    // the same launch (vblocks > 0): its own gate, the summed m v^2, the rule
    int vblocks;
    const int *tgate;
    double4 *vel;
    const unsigned long long *thermo;
    double tconv, T0, coupling;
    int tmode;
};
DilateArgs dilate_args(const Ctx &c);
__device__ inline double dil_mi(double d, const Geometry &g, int k) {
    return g.periodic[k] ? d - g.L[k] * rint(d * g.invL[k]) : d;
}
// velocity rescale of site i (FixNVTRescale / Berendsen weak coupling), as vscale_body
__device__ inline void vscale_thread(int i, int n, double4 *vel, const unsigned long long *thermo, double tconv,
                                     double T0, int mode, double coupling) {
    if (i >= n) return;
    const double T = (double)(long long)thermo[0] * kFxInv * tconv;
    if (!(T > 0.0)) return;
    double l2 = mode == 0 ? T0 / T : 1.0 + coupling * (T0 / T - 1.0);
    if (l2 < 0.0) l2 = 0.0;
    const double lam = sqrt(l2);
    double4 v = vel[i];
    v.x *= lam;
    v.y *= lam;
    v.z *= lam;
    vel[i] = v;
}

__device__ inline void dilate_thread(int t, const DilateArgs &d, const BoxDev *bx) {
    const Geometry &g = bx->geo;
    const double mu = bx->mu;
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = g.lo[k] + 0.5 * g.L[k];
    if (t < d.nmol) {
        const int4 a = d.atoms[t];
        const double mT = d.mO + 2.0 * d.mH;
        const double4 o = d.pos[a.x], p1 = d.pos[a.y], p2 = d.pos[a.z];
        const double oc[3] = {o.x, o.y, o.z};
        const double h1[3] = {o.x + dil_mi(p1.x - o.x, g, 0), o.y + dil_mi(p1.y - o.y, g, 1),
                              o.z + dil_mi(p1.z - o.z, g, 2)};
        const double h2[3] = {o.x + dil_mi(p2.x - o.x, g, 0), o.y + dil_mi(p2.y - o.y, g, 1),
                              o.z + dil_mi(p2.z - o.z, g, 2)};
        double sh[3];
        for (int k = 0; k < 3; ++k) {
            const double R = (1.0 / mT) * (d.mO * oc[k] + d.mH * (h1[k] + h2[k]));
            sh[k] = (c[k] + mu * (R - c[k])) - R;
        }
        const int site[4] = {a.x, a.y, a.z, a.w};
        for (int s = 0; s < 4; ++s) {
            if (site[s] < 0) continue;
            double4 p = d.pos[site[s]];
            p.x += sh[0];
            p.y += sh[1];
            p.z += sh[2];
            d.pos[site[s]] = p;
        }
    }
    if (t < d.n && !(d.member && d.member[t])) {
        double4 p = d.pos[t];
        p.x = c[0] + mu * (p.x - c[0]);
        p.y = c[1] + mu * (p.y - c[1]);
        p.z = c[2] + mu * (p.z - c[2]);
        d.pos[t] = p;
    }
}
// device box: upload the host state, read it back, and the barostat step on
// the device (pressure, dilation, new box, alpha and Green function)
BoxParams box_params(const Ctx &c);
void box_upload(Ctx &c);
void box_pull(Ctx &c);
// rescale: the thermostat's velocity rescale of this step goes with the
// barostat's launches (its m v^2 summed by the second half kick)
void baro_step(Ctx &c, const int *gate, bool rescale = false);
BaroArgs baro_args(const Ctx &c);  // the barostat step's arguments for this context
// Green function from the device box on stream s (gate: nullptr = always)
// dl != nullptr: the same launch also dilates (extra workgroups after the
// Green function's); returns whether it did
bool pppm_green_device(Ctx &c, const int *gate, const DilateArgs *dl = nullptr);
void rigid_first_half(Ctx &c);
// ring.hip: ring-polymer beads, one context per bead, coupled by the
// normal-mode propagation of the free ring polymer
constexpr int kRingTab = 4;  // pointers per bead in RingSpec::tab
struct RingSpec {
    int P = 0;            // beads
    int natoms = 0;       // atoms per bead
    double *C = nullptr;  // P x P orthonormal normal-mode matrix, C[n * P + k]
    double *cw = nullptr; // per mode k: cos(w_k dt), sin(w_k dt) / w_k, w_k sin(w_k dt)
    void **tab = nullptr; // per bead: pos, vel, idToIdx, force (device pointers)
    // ring-polymer contraction of the mesh part (false: every bead its own P3M)
    bool contract = false;
    double4 *cen = nullptr;   // centroids by atom id
    // one neighbor list on the centroids, bead k paired with bead k (false:
    // every bead its own list on its own positions)
    bool centroidList = false;
    float *part = nullptr;    // every bead's skin-check partial maxima (P x 2 x kCheckBlocks)
    std::vector<double4 *> buildPos;  // per bead: centroids in the bead's order
    std::vector<float *> ownPart;     // per bead: its own check table, restored afterwards
    float *q = nullptr;       // charges by atom id
    long long *cf = nullptr;  // centroid forces by atom id, fixed point (kept zero between uses)
};
// ring-level launches, issued (or recorded) on the first bead's context
void ring_drift(const RingSpec &r, Ctx &c0);
void ring_centroids(const RingSpec &r, Ctx &c0, const int *gate = nullptr);
void ring_apply(const RingSpec &r, Ctx &c0);
// pppm.hip: the mesh part of every bead from the ring's centroids, on the
// first bead's mesh (centroid force into r.cf, then onto every bead)
void ring_pppm(Ctx &c0, const RingSpec &r);
// the mesh solve of these P3M settings runs as the three LDS passes (no hipFFT)
bool pppm_lds_fft(const PppmState &P);

// flex.hip: molecule indices after a re-sort, M placement, intramolecular
// forces and the M force spread onto O, H1, H2
void flex_map(Ctx &c);
void flex_place_m(Ctx &c);
void flex_forces(Ctx &c, bool energy, bool virial);
void flex_spread_m(Ctx &c);
void rigid_second_half(Ctx &c, bool kick);
int fused_check_blocks(const Ctx &c);  // workgroups of k_rigid_vv1 when it can give the skin check, else 0

// integrate.hip
void integrate_first_half(Ctx &c);
// thermostat half step and half kick without the drift (ring polymers)
void integrate_kick_first(Ctx &c);
void integrate_second_half(Ctx &c);
// accumulator slots summed over the replicas (synchronises the stream)
void acc_read(Ctx &c, double out[ACC_N]);
// clear slots [slot, slot + count) in every replica
void acc_clear(Ctx &c, int slot, int count);
void zero_forces(Ctx &c, bool clear_acc);
void nose_hoover_start(Ctx &c);       // chain masses and forces from the current kinetic energy
void thermo_measure(Ctx &c);          // sum m v^2 into thermo[0]
void thermostat_apply(Ctx &c);

}  // namespace mdx
