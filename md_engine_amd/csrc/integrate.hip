// FP64 velocity-Verlet integration on the cell-ordered arrays.  Replaces the
// reference's preForce_cu (half kick + drift + force zero,
// src/Integrators/IntegratorVerlet.cu:307-338) and postForce_cu (half kick,
// 610-627).  Forces are read from the 64-bit fixed-point buffer; the inverse
// mass rides in vel.w (GPUData.h:18-28 keeps it in vs.w), massless sites
// (inverse mass 0) are left to their constraint fix.
#include "mdx_internal.h"

namespace mdx {

namespace {

// clears slots [slot, slot + count) of every accumulator replica
__device__ __forceinline__ void acc_clear_body(int blk, int nblk, unsigned long long *__restrict__ acc, int slot, int count) {
    const int r = threadIdx.x / ACC_N, k = threadIdx.x % ACC_N;
    if (r < kAccRep && k >= slot && k < slot + count) acc[r * ACC_N + k] = 0ull;
}

__device__ __forceinline__ void zero_u64_body(int blk, int nblk, const int *__restrict__ gate, int n, unsigned long long *__restrict__ p) {
    if (gate && *gate == 0) return;
    int i = blk * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__device__ __forceinline__ double fx_to_d(unsigned long long v) { return (double)(long long)v * kFxInv; }

__device__ __forceinline__ void vv1_body(int blk, int nblk, int n, double4 *__restrict__ pos, double4 *__restrict__ vel,
                      const unsigned long long *__restrict__ f, double hdtf, double dt,
                      const int *__restrict__ member) {
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (member && member[i]) return;  // moved by its rigid molecule
    double4 v = vel[i];
    if (v.w == 0.0) return;
    double4 x = pos[i];
    const double s = hdtf * v.w;
    v.x += s * fx_to_d(f[3 * i + 0]);
    v.y += s * fx_to_d(f[3 * i + 1]);
    v.z += s * fx_to_d(f[3 * i + 2]);
    x.x += dt * v.x;
    x.y += dt * v.y;
    x.z += dt * v.z;
    vel[i] = v;
    pos[i] = x;
}

// half kick only (the drift of a ring polymer is its normal-mode propagation)
__device__ __forceinline__ void kick_body(int blk, int nblk, int n, double4 *__restrict__ vel, const unsigned long long *__restrict__ f, double hdtf) {
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 v = vel[i];
    if (v.w == 0.0) return;
    const double s = hdtf * v.w;
    v.x += s * fx_to_d(f[3 * i + 0]);
    v.y += s * fx_to_d(f[3 * i + 1]);
    v.z += s * fx_to_d(f[3 * i + 2]);
    vel[i] = v;
}

__device__ __forceinline__ void vv2_body(int blk, int nblk, int n, double4 *__restrict__ vel, const unsigned long long *__restrict__ f,
                      double hdtf, const int *__restrict__ member) {
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (member && member[i]) return;
    double4 v = vel[i];
    if (v.w == 0.0) return;
    const double s = hdtf * v.w;
    v.x += s * fx_to_d(f[3 * i + 0]);
    v.y += s * fx_to_d(f[3 * i + 1]);
    v.z += s * fx_to_d(f[3 * i + 2]);
    vel[i] = v;
}

// Sum of m v^2 over massive sites (FP64 per block, one fixed-point atomic per
// block) -> thermo[0]; the temperature follows from the degrees of freedom.
__device__ __forceinline__ void mv2_body(int blk, int nblk, const int *__restrict__ gate, int n, const double4 *__restrict__ vel,
                                             unsigned long long *__restrict__ thermo) {
    if (gate && *gate == 0) return;
    __shared__ double wsum[4];
    int i = blk * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        double4 v = vel[i];
        if (v.w > 0.0) s = (v.x * v.x + v.y * v.y + v.z * v.z) / v.w;
    }
    s = team_sum_d<64>(s);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += wsum[w];
        atomicAdd(&thermo[0], (unsigned long long)__double2ll_rn(t * kFxScale));
    }
}

// Velocity rescaling toward T0: mode 0 sets T to T0 (FixNVTRescale,
// src/Fixes/FixNVTRescale.cu), mode 1 is Berendsen weak coupling
// lambda^2 = 1 + (interval dt / tau)(T0/T - 1).
__device__ __forceinline__ void vscale_body(int blk, int nblk, const int *__restrict__ gate, int n, double4 *__restrict__ vel, const unsigned long long *__restrict__ thermo,
                         double tconv, double T0, int mode, double coupling) {
    if (gate && *gate == 0) return;
    vscale_thread(blk * blockDim.x + threadIdx.x, n, vel, thermo, tconv, T0, mode, coupling);
}

__device__ __forceinline__ void fill_float_body(int blk, int nblk, int n, float *__restrict__ p, float v) {
    int i = blk * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Nose-Hoover chain (FixNoseHoover::thermostatIntegrate,
// src/Fixes/FixNoseHoover.cu:795-866, after Martyna, Tuckerman, Tobias and
// Klein 1996): one thread advances the chain by half a step and leaves the
// velocity scale factor in st[4M + 1]; `measure` refreshes twice the kinetic
// energy (sum m v^2, which equipartition pairs with ndf kT) from thermo[0]
// first (the stepFinal half).  Chain positions are integrated
// as well so that the conserved extended energy can be reported.
__device__ __forceinline__ void nh_half_body(int blk, int nblk, const int *__restrict__ gate, const unsigned long long *__restrict__ thermo,
                          double *__restrict__ st, int M, double ndf, double kT, double dt, double mvv,
                          int measure) {
    if (gate && *gate == 0) return;
    double *vel = st, *frc = st + M, *mass = st + 2 * M, *pos = st + 3 * M;
    double &ke = st[4 * M];
    const double nkt = ndf * kT;
    if (measure) {
        ke = (double)(long long)thermo[0] * kFxInv * mvv;  // sum m v^2 = 2 KE, paired with ndf kT
        frc[0] = (ke - nkt) / mass[0];
    }
    const double t2 = 0.5 * dt, t4 = 0.25 * dt, t8 = 0.125 * dt;
    vel[M - 1] += t4 * frc[M - 1];
    for (int k = M - 2; k > 0; --k) {
        const double pf = exp(-t8 * vel[k + 1]);
        vel[k] = (vel[k] * pf + t4 * frc[k]) * pf;
    }
    double pf = exp(-t8 * vel[1]);
    vel[0] = (vel[0] * pf + t4 * frc[0]) * pf;
    const double s = exp(-t2 * vel[0]);
    ke *= s * s;
    for (int k = 0; k < M; ++k) pos[k] += t2 * vel[k];
    vel[0] *= pf;
    frc[0] = (ke - nkt) / mass[0];
    vel[0] = (vel[0] + t4 * frc[0]) * pf;
    for (int k = 1; k < M - 1; ++k) {
        pf = exp(-t8 * vel[k + 1]);
        vel[k] *= pf;
        frc[k] = (mass[k - 1] * vel[k - 1] * vel[k - 1] - kT) / mass[k];
        vel[k] = (vel[k] + t4 * frc[k]) * pf;
    }
    frc[M - 1] = (mass[M - 2] * vel[M - 2] * vel[M - 2] - kT) / mass[M - 1];
    vel[M - 1] += t4 * frc[M - 1];
    st[4 * M + 1] = s;
}

// chain start (FixNoseHoover::prepareFinal, FixNoseHoover.cu:240-300 and
// updateThermalMasses, 1034-1042): Q_0 = ndf kT / w^2, Q_k = kT / w^2
__device__ __forceinline__ void nh_start_body(int blk, int nblk, const unsigned long long *__restrict__ thermo, double *__restrict__ st, int M,
                           double ndf, double kT, double freq, double mvv) {
    double *vel = st, *frc = st + M, *mass = st + 2 * M, *pos = st + 3 * M;
    const double ke = (double)(long long)thermo[0] * kFxInv * mvv;  // 2 KE
    st[4 * M] = ke;
    st[4 * M + 1] = 1.0;
    for (int k = 0; k < M; ++k) {
        vel[k] = 0.0;
        pos[k] = 0.0;
        mass[k] = (k == 0 ? ndf : 1.0) * kT / (freq * freq);
    }
    frc[0] = (ke - ndf * kT) / mass[0];
    for (int k = 1; k < M; ++k) frc[k] = (mass[k - 1] * vel[k - 1] * vel[k - 1] - kT) / mass[k];
}

__device__ __forceinline__ void vscale_dev_body(int blk, int nblk, int n, double4 *__restrict__ vel, const double *__restrict__ scale) {
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = *scale;
    double4 v = vel[i];
    v.x *= s;
    v.y *= s;
    v.z *= s;
    vel[i] = v;
}

}  // namespace

static void measure_mv2(Ctx &c, const int *gate) {
    launch<zero_u64_body, 64>(c, 1, 0, gate, 2, c.thermo);
    launch<mv2_body, 256>(c, cdiv(c.n, 256), 0, gate, c.n, c.vel[c.cur], c.thermo);
}

static void nh_half(Ctx &c, bool final_half) {
    NoseHooverState &H = c.nh;
    if (final_half) measure_mv2(c, nullptr);  // every step, also inside a captured step
    launch<nh_half_body, 1>(c, 1, 0, nullptr, c.thermo, H.dev, H.chain, H.ndf, H.boltz * H.T, c.dt, H.mvv, final_half ? 1 : 0);
    launch<vscale_dev_body, 256>(c, cdiv(c.n, 256), 0, c.n, c.vel[c.cur], H.dev + 4 * H.chain + 1);
}

void nose_hoover_start(Ctx &c) {
    NoseHooverState &H = c.nh;
    measure_mv2(c, nullptr);
    launch<nh_start_body, 1>(c, 1, 0, c.thermo, H.dev, H.chain, H.ndf, H.boltz * H.T, H.freq, H.mvv);
    H.fresh = false;
}

void fill_float(Ctx &c, float *p, int count, float v) {
    launch<fill_float_body, 256>(c, cdiv(count, 256), 0, count, p, v);
}

void thermo_measure(Ctx &c) { measure_mv2(c, thermo_gate(c)); }

void thermostat_apply(Ctx &c) {
    if (!c.fuseThermo) thermo_measure(c);  // else summed by the rigid second half
    const double coupling = c.thermoTau > 0 ? c.thermoEvery * c.dt / c.thermoTau : 1.0;
    launch<vscale_body, 256>(c, cdiv(c.n, 256), 0, thermo_gate(c), c.n, c.vel[c.cur], c.thermo, c.tconv, c.thermoT, c.thermoTau > 0 ? 1 : 0, coupling);
}

void acc_read(Ctx &c, double out[ACC_N]) {
    unsigned long long h[kAccTotal];
    MDX_HIP(hipMemcpyAsync(h, c.acc, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    MDX_HIP(hipStreamSynchronize(c.stream));
    for (int s = 0; s < ACC_N; ++s) {
        unsigned long long t = 0;
        for (int r = 0; r < kAccRep; ++r) t += h[r * ACC_N + s];
        out[s] = (double)(long long)t * kFxInv;
    }
}

void acc_clear(Ctx &c, int slot, int count) {
    launch<acc_clear_body, kAccRep * ACC_N>(c, 1, 0, c.acc, slot, count);
}

void zero_forces(Ctx &c, bool clear_acc) {
    const int threads = 256;
    launch<zero_u64_body, 256>(c, cdiv(3LL * c.n, threads), 0, nullptr, 3 * c.n, c.force);
    if (clear_acc)
        launch<zero_u64_body, 256>(c, cdiv(kAccTotal, 256), 0, nullptr, kAccTotal, c.acc);
}

void integrate_first_half(Ctx &c) {
    const int threads = 256;
    if (c.nh.on) nh_half(c, false);
    const int *member = c.rigid.on ? c.rigid.member : nullptr;
    if (!c.rigid.on || c.rigid.nsites < c.n)
        launch<vv1_body, 256>(c, cdiv(c.n, threads), 0, c.n, c.pos[c.cur], c.vel[c.cur], c.force, 0.5 * c.dt * c.ftmToV, c.dt, member);
    if (c.rigid.on) rigid_first_half(c);
}

void integrate_kick_first(Ctx &c) {
    if (c.nh.on) nh_half(c, false);
    launch<kick_body, 256>(c, cdiv(c.n, 256), 0, c.n, c.vel[c.cur], c.force, 0.5 * c.dt * c.ftmToV);
}

void integrate_second_half(Ctx &c) {
    const int threads = 256;
    const int *member = c.rigid.on ? c.rigid.member : nullptr;
    if (!c.rigid.on || c.rigid.nsites < c.n)
        launch<vv2_body, 256>(c, cdiv(c.n, threads), 0, c.n, c.vel[c.cur], c.force, 0.5 * c.dt * c.ftmToV, member);
    if (c.rigid.on) rigid_second_half(c, true);
    if (c.nh.on) nh_half(c, true);
}

}  // namespace mdx
