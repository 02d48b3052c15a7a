// Learned E3B three-body surrogate (no counterpart in the reference, which
// only has the analytic E3B of src/Evaluators/ThreeBodyE3B.h): per-molecule
// descriptors from the E3B molecule list, an NF -> 64 -> 64 -> 1 SiLU MLP on
// BF16 MFMA tiles (FP32 accumulation) with its hand-written backward pass for
// forces, and on-device training (per-molecule energy regression, Adam).
//
// Raw channels of molecule i for NB radial functions
// g_k(r) = exp(-kappa_k r) fc(r) (fc: the E3B switching polynomial RS..RF),
// kappa_k = k3 (1 - 0.1 (KE - k)) for k <= KE = NB - 2, so basis KE is the
// E3B bond function f, and kappa = 2 k3 for basis K2 = NB - 1 (f^2 up to the
// switch):
//   acc_k = sum_{j != i, H of j} g_k(|H - O_i|),
//   d1_k  = sum_{j != i} g_k(|H1_i - O_j|), d2_k likewise for H2_i,
//   pAA = sum_j f(H1_i - O_j) f(H2_i - O_j), pBB = sum_j f(H1_j - O_i) f(H2_j - O_i),
//   pAB = sum_j D_ij D_ji  (D_ij = f(H1_i - O_j) + f(H2_i - O_j));
// network input F = [acc, d1 + d2, d1 * d2, pAA, pBB, pAB, acc_KE s_KE,
// acc_KE^2, s_KE^2] (s = d1 + d2), standardised by (mu, sd).
// Forces: dE/dF per molecule (MLP backward) -> per-raw-channel weights ->
// every owner recomputes its own pairs (deterministic, no cross-molecule
// atomics beyond the fixed-point force buffer).
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int NB = kSurNB, NR = kSurNR, NF = kSurNF, NH = kSurNH, NW = kSurNW;
// network inputs padded with zero columns to KB whole k-blocks of 16 (the
// padding has zero weights and zero inputs, so it changes no result)
constexpr int NFP = kSurNFP, KB = NFP / 16;
static_assert(NW >= NR && NW % 4 == 0, "weight rows hold the raw channels in float4 units");
constexpr int KE = NB - 2;  // basis with kappa = k3: the E3B hydrogen-bond function
constexpr int K2 = NB - 1;  // basis with kappa = 2 k3: f^2 up to the switch
static_assert(K2 == KE + 1, "the last channel squares the E3B one");
// raw channel offsets: acc, d1, d2 (NB each), then pAA, pBB, pAB
constexpr int RD1 = NB, RD2 = 2 * NB, RAA = 3 * NB, RBB = 3 * NB + 1, RAB = 3 * NB + 2;
// radial basis g_k(r) = exp(-kappa_k r) fc(r), kappa_k = k3 (1 - 0.1 (KE - k))
// for k <= KE and kappa = 2 k3 for the last channel (k3 = 1.907 / A,
// switching 5.0-5.2 A)
constexpr float RS_C = 5.0f, RF_C = 5.2f, KAPPA0 = (1.0f - 0.1f * KE) * 1.907f, DKAPPA = 0.1f * 1.907f;

// parameter offsets in the packed model (oracle/surrogate.py layout)
constexpr int OW1 = 0, OB1 = OW1 + NF * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH, OMU = OB3 + 1, OSD = OMU + NF, OES = OSD + NF, OEH = OES + 1;
constexpr int NTRAIN = OMU;  // W1 b1 W2 b2 w3 b3 are trained; mu sd escale eshift are fixed

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float kappa_k(int k) {
    return k < K2 ? KAPPA0 + DKAPPA * (float)k : 2.0f * (KAPPA0 + DKAPPA * (float)KE);
}

// exp(-kappa_k r) of every channel: one exponential per factor, the rest by
// recurrence and the last channel as the square of the E3B one
__device__ __forceinline__ void basis_exp(float r, float (&e)[NB]) {
    const float step = __expf(-DKAPPA * r);
    float v = __expf(-KAPPA0 * r);
#pragma unroll
    for (int k = 0; k < K2; ++k) {
        e[k] = v;
        v *= step;
    }
    e[K2] = e[KE] * e[KE];
}

__device__ __forceinline__ short to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (short)(u >> 16);
}

__device__ __forceinline__ float bf16_round(float f) {
    return __uint_as_float(((unsigned)(unsigned short)to_bf16(f)) << 16);
}

__device__ __forceinline__ void fcut(float r, float &s, float &ds) {
    const float inv = 1.0f / ((RF_C - RS_C) * (RF_C - RS_C) * (RF_C - RS_C));
    if (r <= RS_C) {
        s = 1.0f;
        ds = 0.0f;
    } else if (r >= RF_C) {
        s = 0.0f;
        ds = 0.0f;
    } else {
        s = (RF_C - r) * (RF_C - r) * (RF_C + 2.0f * r - 3.0f * RS_C) * inv;
        ds = 6.0f * (RF_C - r) * (RS_C - r) * inv;
    }
}

// logistic by the hardware reciprocal (about 1 ulp; the network runs on BF16
// operands)
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

struct Geo3 {
    float L[3], invL[3];
};

__device__ __forceinline__ Geo3 geo3(const BoxDev *bx) {
    Geo3 g;
    for (int d = 0; d < 3; ++d) {
        g.L[d] = bx->Lf[d];
        g.invL[d] = bx->invLf[d];
    }
    return g;
}

__device__ __forceinline__ float3 mi(float4 a, float4 b, const Geo3 &g) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    dx -= g.L[0] * rintf(dx * g.invL[0]);
    dy -= g.L[1] * rintf(dy * g.invL[1]);
    dz -= g.L[2] * rintf(dz * g.invL[2]);
    return make_float3(dx, dy, dz);
}

// ---------------------------------------------------------------------------
// descriptors: 16 lanes per molecule split its neighbours, two neighbours
// per lane and round with all their loads in flight together; each lane sums
// all 45 raw channels in registers, the 16 partial rows go through LDS and
// lanes add up the channel columns.  A neighbour whose O-O distance exceeds
// RF_C + r_OH cannot reach the cutoff.
constexpr float kSkipOO2 = (RF_C + 1.0f) * (RF_C + 1.0f);
constexpr int kFeatLanes = 16;

struct FeatAcc {
    float acc[NB], d1[NB], d2[NB];
    float pAA, pBB, pAB;
};

// channels of one neighbour j of molecule i
__device__ __forceinline__ void feat_pair(FeatAcc &A, float4 Oi, float4 H1i, float4 H2i, float4 Oj, float4 H1j,
                                          float4 H2j, const Geo3 &g) {
    const float3 oo = mi(Oj, Oi, g);
    if (oo.x * oo.x + oo.y * oo.y + oo.z * oo.z > kSkipOO2) return;
    // bonds H1_j, H2_j -> O_i (acceptor) and H1_i, H2_i -> O_j (donor)
    const float3 r4[4] = {mi(H1j, Oi, g), mi(H2j, Oi, g), mi(H1i, Oj, g), mi(H2i, Oj, g)};
    float fb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        fb[q] = 0.f;
        const float r = sqrtf(r4[q].x * r4[q].x + r4[q].y * r4[q].y + r4[q].z * r4[q].z);
        float sw, ds;
        fcut(r, sw, ds);
        if (sw <= 0.f) continue;
        float e[NB];
        basis_exp(r, e);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const float v = e[k] * sw;
            if (q < 2) A.acc[k] += v;
            else if (q == 2) A.d1[k] += v;
            else A.d2[k] += v;
            if (k == KE) fb[q] = v;
        }
    }
    A.pAA += fb[2] * fb[3];
    A.pBB += fb[0] * fb[1];
    A.pAB += (fb[2] + fb[3]) * (fb[0] + fb[1]);
}

__global__ void __launch_bounds__(256) k_sur_feat(int nmol, const int4 *__restrict__ molAtoms,
                                                  const float4 *__restrict__ xq, const BoxDev *__restrict__ bx,
                                                  int cap, const int *__restrict__ nl,
                                                  const int *__restrict__ nlCount, float *__restrict__ feat) {
    const Geo3 g = geo3(bx);
    __shared__ float part[256 / kFeatLanes][kFeatLanes][NR + 1];
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gt / kFeatLanes, tl = threadIdx.x % kFeatLanes, team = threadIdx.x / kFeatLanes;
    FeatAcc A;
#pragma unroll
    for (int k = 0; k < NB; ++k) A.acc[k] = A.d1[k] = A.d2[k] = 0.f;
    A.pAA = A.pBB = A.pAB = 0.f;
    if (i < nmol) {
        const int4 ai = molAtoms[i];
        const float4 Oi = xq[ai.x], H1i = xq[ai.y], H2i = xq[ai.z];
        const int cnt = min(nlCount[i], cap);
        const int *row = nl + (size_t)i * cap;
        for (int t = tl; t < cnt; t += 2 * kFeatLanes) {
            const bool two = t + kFeatLanes < cnt;
            const int mA = row[t], mB = two ? row[t + kFeatLanes] : mA;
            const int4 aA = molAtoms[mA], aB = molAtoms[mB];
            const float4 OA = xq[aA.x], H1A = xq[aA.y], H2A = xq[aA.z];
            const float4 OB = xq[aB.x], H1B = xq[aB.y], H2B = xq[aB.z];
            feat_pair(A, Oi, H1i, H2i, OA, H1A, H2A, g);
            if (two) feat_pair(A, Oi, H1i, H2i, OB, H1B, H2B, g);
        }
    }
    float *pr = part[team][tl];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        pr[k] = A.acc[k];
        pr[RD1 + k] = A.d1[k];
        pr[RD2 + k] = A.d2[k];
    }
    pr[RAA] = A.pAA;
    pr[RBB] = A.pBB;
    pr[RAB] = A.pAB;
    __syncthreads();
    if (i < nmol) {
        float *f = feat + (size_t)i * NR;
        for (int col = tl; col < NR; col += kFeatLanes) {
            float sum = 0.f;
            for (int l = 0; l < kFeatLanes; ++l) sum += part[team][l][col];
            f[col] = sum;
        }
    }
}

// network input column `col` of one molecule from its raw channels
__device__ __forceinline__ float feature_of(const float *__restrict__ f, int col) {
    if (col >= NF) return 0.f;  // padding
    if (col < NB) return f[col];
    if (col < 2 * NB) return f[RD1 + col - NB] + f[RD2 + col - NB];
    if (col < 3 * NB) return f[RD1 + col - 2 * NB] * f[RD2 + col - 2 * NB];
    if (col < 3 * NB + 3) return f[col];  // pAA, pBB, pAB
    const float x = f[KE], sk = f[RD1 + KE] + f[RD2 + KE];
    if (col == 3 * NB + 3) return x * sk;
    if (col == 3 * NB + 4) return x * x;
    return sk * sk;
}

// ---------------------------------------------------------------------------
// wave-level 16x16 tiles: C += A (16 x K) B (K x 16), operands FP32 in LDS
// converted to BF16 on load; A(m,k) = A[m*am + k*ak], B(k,n) = B[k*bk + n*bn].
// Lane l: A row l%16, k 4*(l/16)+j; B k 4*(l/16)+j, column l%16; C rows
// 4*(l/16)+j, column l%16 (v_mfma_f32_16x16x16bf16_1k layout).
__device__ __forceinline__ f32x4 mm16(const float *A, int am, int ak, const float *B, int bk, int bn, int K,
                                      f32x4 c) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int kk = k0 + 4 * q + j;
            a[j] = to_bf16(A[r * am + kk * ak]);
            b[j] = to_bf16(B[kk * bk + r * bn]);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    return c;
}

__device__ __forceinline__ void st16(float *D, int ld, int n0, f32x4 c) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) D[(4 * q + j) * ld + n0 + r] = c[j];
}

__device__ __forceinline__ f32x4 zero4() {
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
    return z;
}

// per-wave LDS scratch of the fused MLP (16 molecules)
struct WaveBuf {
    float X[16 * NFP];
    float Z1[16 * NH];
    float H1[16 * NH];
    float Z2[16 * NH];
    float T[16 * NH];  // dZ2, then dH1 / dZ1
};

// Standardised input rows of 16 samples from their raw channels.
__device__ __forceinline__ void load_x(WaveBuf &w, const float *__restrict__ raw, const int *rows, int nrow,
                                       const float *mu, const float *sd) {
    const int l = threadIdx.x & 63;
    for (int e = l; e < 16 * NFP; e += 64) {
        const int s = e / NFP, col = e % NFP;
        const float v = s < nrow ? feature_of(raw + (size_t)rows[s] * NR, col) : 0.f;
        w.X[s * NFP + col] = col < NF ? (v - mu[col]) / sd[col] : 0.f;
    }
}

// forward through both hidden layers: Z1, H1, Z2 (and H2 into T)
__device__ __forceinline__ void forward(WaveBuf &w, const float *W1, const float *b1, const float *W2,
                                        const float *b2) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int n0 = 0; n0 < NH; n0 += 16) st16(w.Z1, NH, n0, mm16(w.X, NFP, 1, W1 + n0, NH, 1, NFP, zero4()));
    wave_sync();
    for (int e = l; e < 16 * NH; e += 64) {
        const float z = w.Z1[e] + b1[e % NH];
        w.Z1[e] = z;
        w.H1[e] = silu(z);
    }
    wave_sync();
#pragma unroll
    for (int n0 = 0; n0 < NH; n0 += 16) st16(w.Z2, NH, n0, mm16(w.H1, NH, 1, W2 + n0, NH, 1, NH, zero4()));
    wave_sync();
    for (int e = l; e < 16 * NH; e += 64) {
        const float z = w.Z2[e] + b2[e % NH];
        w.Z2[e] = z;
        w.T[e] = silu(z);
    }
    wave_sync();
}

// network outputs of the 16 rows: H2 (16 x 64) . w3 as one MFMA column tile
// (B column 0 = w3, other columns 0); lanes with column 0 hold rows 4q + j
__device__ __forceinline__ void out_rows(const WaveBuf &w, const float *w3, float b3, float *e16) {
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    f32x4 c = zero4();
    for (int k0 = 0; k0 < NH; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int kk = k0 + 4 * q + j;
            a[j] = to_bf16(w.T[r * NH + kk]);
            b[j] = r == 0 ? to_bf16(w3[kk]) : (short)0;
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    if (r == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) e16[4 * q + j] = c[j] + b3;
    }
}

// given dE/de per row in g[16] (LDS), back to dZ2 (T), dZ1 (T after) ; X grads optional
__device__ __forceinline__ void backward_hidden(WaveBuf &w, const float *gscale, const float *w3, const float *W2) {
    const int l = threadIdx.x & 63;
    for (int e = l; e < 16 * NH; e += 64) w.T[e] = gscale[e / NH] * w3[e % NH] * dsilu(w.Z2[e]);  // dZ2
    wave_sync();
    // dH1 = dZ2 W2^T  -> into H1's place is needed later for dW2; use Z1 buffer after dsilu
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mm16(w.T, NH, 1, W2 + t * 16 * NH, 1, NH, NH, zero4());
    wave_sync();
    // dZ1 = dH1 * silu'(Z1), kept in Z1
    {
        const int r = l & 15, q = l >> 4;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int e = (4 * q + j) * NH + t * 16 + r;
                w.Z1[e] = acc[t][j] * dsilu(w.Z1[e]);
            }
    }
    wave_sync();
}

constexpr int kWaves = 2;  // waves (16-molecule tiles) per workgroup of the training kernel

// ---------------------------------------------------------------------------
// inference with the activations kept in registers.  Every product is taken
// transposed (hidden x molecules): the C tile of one layer, rows 4 (l / 16) + j
// and column l % 16 of lane l, is exactly the B operand layout of the next
// layer's k-block, so a 16-molecule tile runs forward and backward through
// both layers without touching LDS.  The weights come as BF16 A-operand
// fragments packed once per model (k_sur_pack): lane l of fragment (t, kb)
// holds 4 consecutive k of row l % 16.
//   Z1^T = W1^T X^T, Z2^T = W2^T H1^T (forward), dH1^T = W2 dZ2^T, dX^T = W1 dZ1^T
constexpr int FR_A1 = 0;                    // W1^T: 4 hidden tiles x KB feature blocks
constexpr int FR_A2 = FR_A1 + 4 * KB;       // W2^T: 4 x 4
constexpr int FR_A3 = FR_A2 + 16;           // W2:   4 x 4
constexpr int FR_A4 = FR_A3 + 16;           // W1:   KB feature tiles x 4 hidden blocks
constexpr int FR_N = FR_A4 + 4 * KB;

// one thread per (fragment, lane): 4 BF16 values
__global__ void k_sur_pack(const float *__restrict__ P, unsigned short *__restrict__ frag) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= FR_N * 64) return;
    const int fr = e / 64, l = e % 64, r = l & 15, q = l >> 4;
    const float *W1 = P + OW1, *W2 = P + OW2;
    for (int j = 0; j < 4; ++j) {
        float v;
        if (fr < FR_A2) {  // A = W1^T: row hidden 16 t + r, k feature 16 kb + 4 q + j
            const int t = (fr - FR_A1) / KB, kb = (fr - FR_A1) % KB, f = 16 * kb + 4 * q + j;
            v = f < NF ? W1[f * NH + 16 * t + r] : 0.f;
        } else if (fr < FR_A3) {  // A = W2^T: row hidden2 16 t + r, k hidden1 16 kb + 4 q + j
            const int t = (fr - FR_A2) / 4, kb = (fr - FR_A2) % 4;
            v = W2[(16 * kb + 4 * q + j) * NH + 16 * t + r];
        } else if (fr < FR_A4) {  // A = W2: row hidden1 16 t + r, k hidden2 16 kb + 4 q + j
            const int t = (fr - FR_A3) / 4, kb = (fr - FR_A3) % 4;
            v = W2[(16 * t + r) * NH + 16 * kb + 4 * q + j];
        } else {  // A = W1: row feature 16 t + r, k hidden 16 kb + 4 q + j
            const int t = (fr - FR_A4) / 4, kb = (fr - FR_A4) % 4, f = 16 * t + r;
            v = f < NF ? W1[f * NH + 16 * kb + 4 * q + j] : 0.f;
        }
        frag[(size_t)e * 4 + j] = (unsigned short)to_bf16(v);
    }
}

__device__ __forceinline__ bf16x4 ld_frag(const uint2 *sfr, int fr, int l) {
    const uint2 u = sfr[fr * 64 + l];
    bf16x4 a;
    a[0] = (short)(u.x & 0xFFFFu);
    a[1] = (short)(u.x >> 16);
    a[2] = (short)(u.y & 0xFFFFu);
    a[3] = (short)(u.y >> 16);
    return a;
}

__device__ __forceinline__ bf16x4 pack4(f32x4 v) {
    bf16x4 b;
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = to_bf16(v[j]);
    return b;
}

constexpr int kMlpWaves = 2;

__global__ void __launch_bounds__(64 * kMlpWaves) k_sur_infer(int nmol, const float *__restrict__ P,
                                                              const unsigned short *__restrict__ frag,
                                                              const float *__restrict__ raw, float *__restrict__ emol,
                                                              float *__restrict__ wgt,
                                                              unsigned long long *__restrict__ acc, int energy) {
    // the weight fragments and vectors are staged in LDS once per workgroup
    __shared__ uint2 sfr[FR_N * 64];
    __shared__ float sb1[NH], sb2[NH], sw3[NH], sw3r[NH], smu[NFP], sisd[NFP];
    __shared__ float sraw[kMlpWaves][16][NR + 1];
    __shared__ float sdx[kMlpWaves][16][NFP + 1];
    static_assert(FR_N * 64 % (64 * kMlpWaves) == 0, "fragment staging");
    {
        // all loads first, then the LDS stores
        uint2 tmp[FR_N * 64 / (64 * kMlpWaves)];
#pragma unroll
        for (int u = 0; u < FR_N * 64 / (64 * kMlpWaves); ++u)
            tmp[u] = ((const uint2 *)frag)[threadIdx.x + u * 64 * kMlpWaves];
#pragma unroll
        for (int u = 0; u < FR_N * 64 / (64 * kMlpWaves); ++u) sfr[threadIdx.x + u * 64 * kMlpWaves] = tmp[u];
    }
    for (int e = threadIdx.x; e < NH; e += blockDim.x) {
        sb1[e] = P[OB1 + e];
        sb2[e] = P[OB2 + e];
        sw3[e] = P[OW3 + e];
        sw3r[e] = bf16_round(P[OW3 + e]);
    }
    for (int e = threadIdx.x; e < NFP; e += blockDim.x) {
        smu[e] = e < NF ? P[OMU + e] : 0.f;
        sisd[e] = e < NF ? 1.0f / P[OSD + e] : 0.f;
    }
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    const int m0 = (blockIdx.x * kMlpWaves + wv) * 16;
    const int nrow = max(0, min(16, nmol - m0));
    {
        float tmp[(16 * NR + 63) / 64];
#pragma unroll
        for (int u = 0; u < (16 * NR + 63) / 64; ++u) {
            const int e = l + 64 * u, s2 = e / NR;
            tmp[u] = e < 16 * NR && s2 < nrow ? raw[(size_t)m0 * NR + e] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < (16 * NR + 63) / 64; ++u) {
            const int e = l + 64 * u;
            if (e < 16 * NR) sraw[wv][e / NR][e % NR] = tmp[u];
        }
    }
    __syncthreads();
    if (nrow == 0) return;  // whole wave; no barrier follows
    const float *fm = sraw[wv][min(r, nrow - 1)];
    // X^T k-blocks: lane holds features 16 kb + 4 q + j of molecule r
    bf16x4 xb[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        f32x4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = 16 * kb + 4 * q + j;
            v[j] = (feature_of(fm, col) - smu[col]) * sisd[col];
        }
        xb[kb] = pack4(v);
    }
    // layer 1
    f32x4 z1[4];
    bf16x4 h1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ld_frag(sfr, FR_A1 + KB * t + kb, l), xb[kb], c, 0, 0, 0);
        f32x4 h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] += sb1[16 * t + 4 * q + j];
            h[j] = silu(c[j]);
        }
        z1[t] = c;
        h1[t] = pack4(h);
    }
    // layer 2 and the output row
    f32x4 z2[4];
    float part = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ld_frag(sfr, FR_A2 + 4 * t + kb, l), h1[kb], c, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] += sb2[16 * t + 4 * q + j];
            // the output layer as in the training kernel: BF16 H2 and w3
            part += bf16_round(silu(c[j])) * sw3r[16 * t + 4 * q + j];
        }
        z2[t] = c;
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    const float es = P[OES];
    const float e = es * (part + P[OB3]) + P[OEH];
    double esum = 0.0;
    if (q == 0 && r < nrow) {
        emol[m0 + r] = e;
        esum = (double)e;
    }
    // backward: dZ2^T = es w3 silu'(Z2^T)
    bf16x4 dz2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 d;
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = es * sw3[16 * t + 4 * q + j] * dsilu(z2[t][j]);
        dz2[t] = pack4(d);
    }
    // dZ1^T = (W2 dZ2^T) silu'(Z1^T)
    bf16x4 dz1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ld_frag(sfr, FR_A3 + 4 * t + kb, l), dz2[kb], c, 0, 0, 0);
        f32x4 d;
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = c[j] * dsilu(z1[t][j]);
        dz1[t] = pack4(d);
    }
    // dX^T = W1 dZ1^T, then dE/dF = dX / sd into LDS (row = molecule)
#pragma unroll
    for (int t = 0; t < KB; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ld_frag(sfr, FR_A4 + 4 * t + kb, l), dz1[kb], c, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = 16 * t + 4 * q + j;
            sdx[wv][r][col] = c[j] * sisd[col];
        }
    }
    wave_sync();
    // dE/d(raw channel) per molecule
    for (int e2 = l; e2 < 16 * NB; e2 += 64) {
        const int s = e2 / NB, k = e2 % NB;
        if (s >= nrow) continue;
        const float *f = sraw[wv][s];
        const float *gx = sdx[wv][s];
        const float d1 = f[RD1 + k], d2 = f[RD2 + k];
        float ga = gx[k], gs = gx[NB + k];
        const float gp = gx[2 * NB + k];
        if (k == KE) {
            const float x = f[KE], sk = f[RD1 + KE] + f[RD2 + KE];
            ga += gx[3 * NB + 3] * sk + 2.0f * gx[3 * NB + 4] * x;
            gs += gx[3 * NB + 3] * x + 2.0f * gx[3 * NB + 5] * sk;
        }
        float *o = wgt + (size_t)(m0 + s) * NW;
        o[k] = ga;
        o[RD1 + k] = gs + gp * d2;
        o[RD2 + k] = gs + gp * d1;
    }
    if (l < 48) {
        const int s = l / 3, c3 = l % 3;
        if (s < nrow) wgt[(size_t)(m0 + s) * NW + RAA + c3] = sdx[wv][s][3 * NB + c3];
    }
    if (energy) {
        esum = team_sum_d<64>(esum);
        if (l == 0 && esum != 0.0) fx_add(&acc_rep(acc)[ACC_E3B3], esum);
    }
}

// ---------------------------------------------------------------------------
// Descriptors, the analytic E3B two-body term and the network in one launch.
// A workgroup of 256 lanes takes 16 molecules, 16 lanes each: every lane
// walks a share of its molecule's list, sums the 45 raw channels and the
// two-body force on the oxygen (F = d k2 E2 exp(-k2 r) / r for r_OO < r2b,
// the same term as the E3B pass, EvaluatorE3B.h:67-87); the lanes' rows are
// reduced through LDS, and the first wave runs the 16 molecules through the
// network as one MFMA column tile (forward, then the backward pass to
// dE / d(raw channel)), with the BF16 weight fragments read straight from the
// packed table.  Writes the per-molecule energies and channel weights for the
// force pass (and the raw channels when `feat` is given).
struct TwoBody {
    float E2, k2, r2b2;
};

template <bool EV, bool VIR>
__device__ __forceinline__ void sur_fused_body(int blk, int nblk, int nmol, const int4 *__restrict__ molAtoms,
                                                   const float4 *__restrict__ xq, const BoxDev *__restrict__ bx,
                                                   int cap, const int *__restrict__ nl, const int *__restrict__ nlCount,
                                                   TwoBody tb, const float *__restrict__ P,
                                                   const unsigned short *__restrict__ frag,
                                                   float *__restrict__ feat, float *__restrict__ emol,
                                                   float *__restrict__ wgt, unsigned long long *__restrict__ force,
                                                   unsigned long long *__restrict__ acc) {
    const Geo3 g = geo3(bx);
    __shared__ float part[256 / kFeatLanes][kFeatLanes][NR + 1];
    __shared__ float sdx[16][NFP + 1];
    __shared__ double red[4][7];
    const int gt = blk * blockDim.x + threadIdx.x;
    const int i = gt / kFeatLanes, tl = threadIdx.x % kFeatLanes, team = threadIdx.x / kFeatLanes;
    const int m0 = blk * (256 / kFeatLanes);
    FeatAcc A;
#pragma unroll
    for (int k = 0; k < NB; ++k) A.acc[k] = A.d1[k] = A.d2[k] = 0.f;
    A.pAA = A.pBB = A.pAB = 0.f;
    float fx = 0.f, fy = 0.f, fz = 0.f, e2 = 0.f;
    float w0 = 0.f, w1 = 0.f, w2 = 0.f, w3 = 0.f, w4 = 0.f, w5 = 0.f;
    int4 ai = make_int4(0, 0, 0, 0);
    if (i < nmol) {
        ai = molAtoms[i];
        const float4 Oi = xq[ai.x], H1i = xq[ai.y], H2i = xq[ai.z];
        const int cnt = min(nlCount[i], cap);
        const int *row = nl + (size_t)i * cap;
        for (int t = tl; t < cnt; t += 2 * kFeatLanes) {
            const bool two = t + kFeatLanes < cnt;
            const int mA = row[t], mB = two ? row[t + kFeatLanes] : mA;
            const int4 aA = molAtoms[mA], aB = molAtoms[mB];
            const float4 OA = xq[aA.x], H1A = xq[aA.y], H2A = xq[aA.z];
            const float4 OB = xq[aB.x], H1B = xq[aB.y], H2B = xq[aB.z];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (u == 1 && !two) break;
                const float4 Oj = u ? OB : OA;
                const float3 oo = mi(Oj, Oi, g);
                const float r2 = oo.x * oo.x + oo.y * oo.y + oo.z * oo.z;
                if (r2 < tb.r2b2) {
                    const float r = sqrtf(r2);
                    const float ex = __expf(-tb.k2 * r);
                    const float sc = tb.k2 * tb.E2 * ex / r;
                    // d = r_Oi - r_Oj = -oo: a repulsion along d
                    fx -= oo.x * sc;
                    fy -= oo.y * sc;
                    fz -= oo.z * sc;
                    if (EV) e2 += 0.5f * tb.E2 * ex;
                    if (VIR) {
                        w0 += oo.x * sc * oo.x;
                        w1 += oo.y * sc * oo.y;
                        w2 += oo.z * sc * oo.z;
                        w3 += oo.x * sc * oo.y;
                        w4 += oo.x * sc * oo.z;
                        w5 += oo.y * sc * oo.z;
                    }
                }
                if (u == 0) feat_pair(A, Oi, H1i, H2i, OA, H1A, H2A, g);
                else feat_pair(A, Oi, H1i, H2i, OB, H1B, H2B, g);
            }
        }
    }
    fx = team_sum<kFeatLanes>(fx);
    fy = team_sum<kFeatLanes>(fy);
    fz = team_sum<kFeatLanes>(fz);
    if (tl == 0 && i < nmol) {
        fx_add(&force[3 * ai.x + 0], fx);
        fx_add(&force[3 * ai.x + 1], fy);
        fx_add(&force[3 * ai.x + 2], fz);
    }
    float *pr = part[team][tl];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        pr[k] = A.acc[k];
        pr[RD1 + k] = A.d1[k];
        pr[RD2 + k] = A.d2[k];
    }
    pr[RAA] = A.pAA;
    pr[RBB] = A.pBB;
    pr[RAB] = A.pAB;
    if (EV || VIR) {
        // two-body energy and virial: wave sums, one atomic per workgroup and value
        const double v[7] = {e2, 0.5 * w0, 0.5 * w1, 0.5 * w2, 0.5 * w3, 0.5 * w4, 0.5 * w5};
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int s2 = 0; s2 < 7; ++s2) {
            const double x = team_sum_d<64>(v[s2]);
            if (lane == 0) red[w][s2] = x;
        }
    }
    __syncthreads();
    if (EV || VIR) {
        if (threadIdx.x < 7) {
            double x = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
            if (x != 0.0) fx_add(&acc_rep(acc)[threadIdx.x == 0 ? ACC_E3B2 : ACC_VIR + threadIdx.x - 1], x);
        }
    }
    // the 16 molecules' channel sums: row 0 of each team's slab
    for (int e = threadIdx.x; e < 16 * NR; e += 256) {
        const int s = e / NR, col = e % NR;
        float sum = 0.f;
        for (int l = 1; l < kFeatLanes; ++l) sum += part[s][l][col];
        // lane 0's own partial stays in place; add the others after reading all
        part[s][0][col] += sum;
        if (feat && m0 + s < nmol) feat[(size_t)(m0 + s) * NR + col] = part[s][0][col];
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;  // the network runs in the first wave; no barrier follows
    const int l = threadIdx.x, r = l & 15, q = l >> 4;
    const int nrow = max(0, min(16, nmol - m0));
    if (nrow == 0) return;
    const float *fm = part[min(r, nrow - 1)][0];
    const uint2 *fr2 = (const uint2 *)frag;
    auto ldf = [&](int f) {
        const uint2 u = fr2[f * 64 + l];
        bf16x4 a;
        a[0] = (short)(u.x & 0xFFFFu);
        a[1] = (short)(u.x >> 16);
        a[2] = (short)(u.y & 0xFFFFu);
        a[3] = (short)(u.y >> 16);
        return a;
    };
    bf16x4 xb[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        f32x4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = 16 * kb + 4 * q + j;
            v[j] = col < NF ? (feature_of(fm, col) - P[OMU + col]) / P[OSD + col] : 0.f;
        }
        xb[kb] = pack4(v);
    }
    f32x4 z1[4];
    bf16x4 h1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ldf(FR_A1 + KB * t + kb), xb[kb], c, 0, 0, 0);
        f32x4 h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] += P[OB1 + 16 * t + 4 * q + j];
            h[j] = silu(c[j]);
        }
        z1[t] = c;
        h1[t] = pack4(h);
    }
    f32x4 z2[4];
    float outp = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ldf(FR_A2 + 4 * t + kb), h1[kb], c, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] += P[OB2 + 16 * t + 4 * q + j];
            outp += bf16_round(silu(c[j])) * bf16_round(P[OW3 + 16 * t + 4 * q + j]);
        }
        z2[t] = c;
    }
    outp += __shfl_xor(outp, 16);
    outp += __shfl_xor(outp, 32);
    const float es = P[OES];
    const float eo = es * (outp + P[OB3]) + P[OEH];
    double esum = 0.0;
    if (q == 0 && r < nrow) {
        emol[m0 + r] = eo;
        esum = (double)eo;
    }
    bf16x4 dz2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 d;
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = es * P[OW3 + 16 * t + 4 * q + j] * dsilu(z2[t][j]);
        dz2[t] = pack4(d);
    }
    bf16x4 dz1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ldf(FR_A3 + 4 * t + kb), dz2[kb], c, 0, 0, 0);
        f32x4 d;
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = c[j] * dsilu(z1[t][j]);
        dz1[t] = pack4(d);
    }
#pragma unroll
    for (int t = 0; t < KB; ++t) {
        f32x4 c = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ldf(FR_A4 + 4 * t + kb), dz1[kb], c, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = 16 * t + 4 * q + j;
            sdx[r][col] = col < NF ? c[j] / P[OSD + col] : 0.f;
        }
    }
    wave_sync();
    for (int e = l; e < 16 * NB; e += 64) {
        const int s = e / NB, k = e % NB;
        if (s >= nrow) continue;
        const float *f = part[s][0];
        const float *gx = sdx[s];
        const float d1 = f[RD1 + k], d2 = f[RD2 + k];
        float ga = gx[k], gs = gx[NB + k];
        const float gp = gx[2 * NB + k];
        if (k == KE) {
            const float x = f[KE], sk = f[RD1 + KE] + f[RD2 + KE];
            ga += gx[3 * NB + 3] * sk + 2.0f * gx[3 * NB + 4] * x;
            gs += gx[3 * NB + 3] * x + 2.0f * gx[3 * NB + 5] * sk;
        }
        float *o = wgt + (size_t)(m0 + s) * NW;
        o[k] = ga;
        o[RD1 + k] = gs + gp * d2;
        o[RD2 + k] = gs + gp * d1;
    }
    if (l < 48) {
        const int s = l / 3, c3 = l % 3;
        if (s < nrow) wgt[(size_t)(m0 + s) * NW + RAA + c3] = sdx[s][3 * NB + c3];
    }
    if (EV) {
        esum = team_sum_d<64>(esum);
        if (l == 0 && esum != 0.0) fx_add(&acc_rep(acc)[ACC_E3B3], esum);
    }
}

// ---------------------------------------------------------------------------
// forces: 16 lanes per molecule split its neighbours (owner-computes); each
// lane holds the molecule's 45 raw-channel weights, gathers the neighbour's,
// and the lanes' partial forces are summed by shuffles.  A bond H -> O of the
// pair (i, j) with value a = f feeds: the donor channel of H's molecule and
// the acceptor channel of O's molecule (radial weights), the product of the
// two bonds of H's molecule to that O (pAA of H's molecule, pBB of O's), and
// the mutual product D_ij D_ji (pAB of both).  VIR: the virial of every
// bond, -(dE/dr / r) d (x) d, half from each of its two owners.
// forces on molecule i's own atoms from its bonds with one neighbour j
template <bool VIR>
__device__ __forceinline__ void force_pair(float4 Oi, float4 H1i, float4 H2i, float4 Oj, float4 H1j, float4 H2j,
                                           const float (&wi)[NW], const float (&wj)[NW], const Geo3 &g,
                                           float3 &FO, float3 &F1, float3 &F2, float &v0, float &v1, float &v2,
                                           float &v3, float &v4, float &v5) {
    const float3 oo = mi(Oj, Oi, g);
    if (oo.x * oo.x + oo.y * oo.y + oo.z * oo.z > kSkipOO2) return;
    // q = 0, 1: H1_i, H2_i -> O_j (i's donor channels, j's acceptor);
    // q = 2, 3: H1_j, H2_j -> O_i (j's donor channels, i's acceptor)
    const float3 d[4] = {mi(H1i, Oj, g), mi(H2i, Oj, g), mi(H1j, Oi, g), mi(H2j, Oi, g)};
    float r[4], sw[4], ds[4], fb[4], dfb[4], eb[4][NB];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        r[q] = sqrtf(d[q].x * d[q].x + d[q].y * d[q].y + d[q].z * d[q].z);
        fcut(r[q], sw[q], ds[q]);
        basis_exp(r[q], eb[q]);
        const float ek = eb[q][KE];
        fb[q] = ek * sw[q];
        dfb[q] = ek * (ds[q] - kappa_k(KE) * sw[q]);
    }
    const float wAB = wi[RAB] + wj[RAB];
    const float Dij = fb[0] + fb[1], Dji = fb[2] + fb[3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (sw[q] <= 0.f && ds[q] == 0.f) continue;
        // d/dr [exp(-kappa r) fc] = exp(-kappa r) (fc' - kappa fc)
        float dEdr = 0.f;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const float w = q == 0 ? wi[RD1 + k] + wj[k]
                          : q == 1 ? wi[RD2 + k] + wj[k]
                          : q == 2 ? wj[RD1 + k] + wi[k] : wj[RD2 + k] + wi[k];
            dEdr += w * eb[q][k] * (ds[q] - kappa_k(k) * sw[q]);
        }
        // pair products: the other bond of the same hydrogen pair, and D of the reverse direction
        const float other = fb[q ^ 1];
        const float wpair = q < 2 ? wi[RAA] + wj[RBB] : wj[RAA] + wi[RBB];
        dEdr += (wpair * other + wAB * (q < 2 ? Dji : Dij)) * dfb[q];
        const float c = dEdr / r[q];
        if (VIR) {
            const float h = -0.5f * c;
            v0 += h * d[q].x * d[q].x;
            v1 += h * d[q].y * d[q].y;
            v2 += h * d[q].z * d[q].z;
            v3 += h * d[q].x * d[q].y;
            v4 += h * d[q].x * d[q].z;
            v5 += h * d[q].y * d[q].z;
        }
        // F_H = -c d, F_O = +c d  (d = r_H - r_O)
        if (q == 0) {
            F1.x -= c * d[q].x;
            F1.y -= c * d[q].y;
            F1.z -= c * d[q].z;
        } else if (q == 1) {
            F2.x -= c * d[q].x;
            F2.y -= c * d[q].y;
            F2.z -= c * d[q].z;
        } else {
            FO.x += c * d[q].x;
            FO.y += c * d[q].y;
            FO.z += c * d[q].z;
        }
    }
}

template <bool VIR, int kForceLanes>
__device__ __forceinline__ void sur_force_body(int blk, int nblk, int nmol, const int4 *__restrict__ molAtoms,
                                                   const float4 *__restrict__ xq, const BoxDev *__restrict__ bx,
                                                   int cap, const int *__restrict__ nl,
                                                   const int *__restrict__ nlCount, const float *__restrict__ wgt,
                                                   unsigned long long *__restrict__ force,
                                                   unsigned long long *__restrict__ acc) {
    // the molecule's own channel weights are shared by its lanes through LDS;
    // a neighbour's are fetched only once its O-O distance passed the test
    constexpr int M = 256 / kForceLanes;
    __shared__ float4 swi[M][NW / 4];
    const Geo3 g = geo3(bx);
    const int gt = blk * blockDim.x + threadIdx.x;
    const int i = gt / kForceLanes, tl = threadIdx.x % kForceLanes, team = threadIdx.x / kForceLanes;
    for (int v = tl; v < NW / 4; v += kForceLanes)
        swi[team][v] = i < nmol ? ((const float4 *)(wgt + (size_t)i * NW))[v] : make_float4(0, 0, 0, 0);
    __syncthreads();
    float wi[NW];
#pragma unroll
    for (int v = 0; v < NW / 4; ++v) {
        const float4 t4 = swi[team][v];
        wi[4 * v] = t4.x;
        wi[4 * v + 1] = t4.y;
        wi[4 * v + 2] = t4.z;
        wi[4 * v + 3] = t4.w;
    }
    float3 FO = make_float3(0.f, 0.f, 0.f), F1 = FO, F2 = FO;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
    int4 ai = make_int4(0, 0, 0, 0);
    if (i < nmol) {
        ai = molAtoms[i];
        const float4 Oi = xq[ai.x], H1i = xq[ai.y], H2i = xq[ai.z];
        const int cnt = min(nlCount[i], cap);
        const int *row = nl + (size_t)i * cap;
        for (int t = tl; t < cnt; t += kForceLanes) {
            const int mj = row[t];
            const int4 aj = molAtoms[mj];
            const float4 Oj = xq[aj.x];
            const float3 oo = mi(Oj, Oi, g);
            if (oo.x * oo.x + oo.y * oo.y + oo.z * oo.z > kSkipOO2) continue;
            const float4 H1j = xq[aj.y], H2j = xq[aj.z];
            float wj[NW];
            const float4 *w4 = (const float4 *)(wgt + (size_t)mj * NW);
#pragma unroll
            for (int v = 0; v < NW / 4; ++v) {
                const float4 t4 = w4[v];
                wj[4 * v] = t4.x;
                wj[4 * v + 1] = t4.y;
                wj[4 * v + 2] = t4.z;
                wj[4 * v + 3] = t4.w;
            }
            force_pair<VIR>(Oi, H1i, H2i, Oj, H1j, H2j, wi, wj, g, FO, F1, F2, v0, v1, v2, v3, v4, v5);
        }
    }
    FO.x = team_sum<kForceLanes>(FO.x);
    FO.y = team_sum<kForceLanes>(FO.y);
    FO.z = team_sum<kForceLanes>(FO.z);
    F1.x = team_sum<kForceLanes>(F1.x);
    F1.y = team_sum<kForceLanes>(F1.y);
    F1.z = team_sum<kForceLanes>(F1.z);
    F2.x = team_sum<kForceLanes>(F2.x);
    F2.y = team_sum<kForceLanes>(F2.y);
    F2.z = team_sum<kForceLanes>(F2.z);
    if (tl == 0 && i < nmol) {
        fx_add(&force[3 * ai.x + 0], FO.x);
        fx_add(&force[3 * ai.x + 1], FO.y);
        fx_add(&force[3 * ai.x + 2], FO.z);
        fx_add(&force[3 * ai.y + 0], F1.x);
        fx_add(&force[3 * ai.y + 1], F1.y);
        fx_add(&force[3 * ai.y + 2], F1.z);
        fx_add(&force[3 * ai.z + 0], F2.x);
        fx_add(&force[3 * ai.z + 1], F2.y);
        fx_add(&force[3 * ai.z + 2], F2.z);
    }
    if (VIR) {
        // wave sums, then one atomic per workgroup and component
        __shared__ double red[4][6];
        const double vv[6] = {v0, v1, v2, v3, v4, v5};
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double x = team_sum_d<64>(vv[k]);
            if (lane == 0) red[w][k] = x;
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const double x = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
            if (x != 0.0) fx_add(&acc_rep(acc)[ACC_VIR + threadIdx.x], x);
        }
    }
}

// ---------------------------------------------------------------------------
// training: per-wave partial gradients of the mean squared error of the
// standardised per-molecule energy over one mini-batch
constexpr int kTrainWaves = 256;  // 128 blocks of 2 waves
constexpr int kGradStride = NTRAIN + 1;  // trained parameters + loss

__global__ void __launch_bounds__(64 * kWaves) k_sur_grad(int nsamp, int bstart, int bsize, unsigned mulA, unsigned addB,
                                                  const float *__restrict__ P, const float *__restrict__ raw,
                                                  const float *__restrict__ target, float *__restrict__ part) {
    __shared__ float sW1[NFP * NH], sW2[NH * NH], sb1[NH], sb2[NH], sw3[NH], smu[NFP], ssd[NFP];
    __shared__ WaveBuf wb[kWaves];
    __shared__ float sg[kWaves][16], se[kWaves][16];
    __shared__ int srow[kWaves][16];
    for (int e = threadIdx.x; e < NFP * NH; e += blockDim.x) sW1[e] = e < NF * NH ? P[OW1 + e] : 0.f;
    for (int e = threadIdx.x; e < NH * NH; e += blockDim.x) sW2[e] = P[OW2 + e];
    for (int e = threadIdx.x; e < NH; e += blockDim.x) {
        sb1[e] = P[OB1 + e];
        sb2[e] = P[OB2 + e];
        sw3[e] = P[OW3 + e];
    }
    for (int e = threadIdx.x; e < NFP; e += blockDim.x) {
        smu[e] = e < NF ? P[OMU + e] : 0.f;
        ssd[e] = e < NF ? P[OSD + e] : 1.f;
    }
    __syncthreads();
    const float b3 = P[OB3], es = P[OES], eh = P[OEH];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, q = l >> 4;
    const int wave = blockIdx.x * kWaves + wv;
    WaveBuf &w = wb[wv];
    f32x4 gW1[KB][4], gW2[4][4];
#pragma unroll
    for (int a = 0; a < KB; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) gW1[a][b] = zero4();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) gW2[a][b] = zero4();
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
    const int ntile = (bsize + 15) / 16;
    const float inv = 1.0f / (float)bsize;
    for (int t = wave; t < ntile; t += gridDim.x * kWaves) {
        const int s0 = t * 16;
        const int nrow = min(16, bsize - s0);
        if (l < 16) {
            const int s = min(l, nrow - 1);
            // pseudo-shuffled sample order: (bstart + s0 + s) * a + b mod nsamp
            srow[wv][l] = (int)(((unsigned long long)(unsigned)(bstart + s0 + s) * mulA + addB) % (unsigned)nsamp);
        }
        wave_sync();
        load_x(w, raw, srow[wv], nrow, smu, ssd);
        wave_sync();
        forward(w, sW1, sb1, sW2, sb2);
        out_rows(w, sw3, b3, se[wv]);
        wave_sync();
        if (l < 16) {
            float gr = 0.f;
            if (l < nrow) {
                const float tt = (target[srow[wv][l]] - eh) / es;
                const float diff = se[wv][l] - tt;
                gr = 2.0f * diff * inv;
                se[wv][l] = diff * diff * inv;
            } else {
                se[wv][l] = 0.f;
            }
            sg[wv][l] = gr;
        }
        wave_sync();
        if (l == 0)
            for (int s = 0; s < 16; ++s) loss += se[wv][s];
        // dw3 += H2^T g, db3 += sum g (lane l owns hidden unit l)
        for (int s = 0; s < 16; ++s) gw3 += sg[wv][s] * w.T[s * NH + l];
        if (l == 0)
            for (int s = 0; s < 16; ++s) gb3 += sg[wv][s];
        wave_sync();
        backward_hidden(w, sg[wv], sw3, sW2);  // T = dZ2, Z1 = dZ1
        for (int s = 0; s < 16; ++s) {
            gb2 += w.T[s * NH + l];
            gb1 += w.Z1[s * NH + l];
        }
        // dW2 += H1^T dZ2: A(m=a, k=s) = H1[s][a], B(k=s, n=b) = dZ2[s][b]
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) gW2[a][b] = mm16(w.H1 + a * 16, 1, NH, w.T + b * 16, NH, 1, 16, gW2[a][b]);
        // dW1 += X^T dZ1
#pragma unroll
        for (int a = 0; a < KB; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) gW1[a][b] = mm16(w.X + a * 16, 1, NFP, w.Z1 + b * 16, NH, 1, 16, gW1[a][b]);
        wave_sync();
    }
    float *o = part + (size_t)wave * kGradStride;
#pragma unroll
    for (int a = 0; a < KB; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (a * 16 + 4 * q + j < NF) o[OW1 + (a * 16 + 4 * q + j) * NH + b * 16 + r] = gW1[a][b][j];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) o[OW2 + (a * 16 + 4 * q + j) * NH + b * 16 + r] = gW2[a][b][j];
    o[OB1 + l] = gb1;
    o[OB2 + l] = gb2;
    o[OW3 + l] = gw3;
    if (l == 0) {
        o[OB3] = gb3;
        o[NTRAIN] = loss;
    }
}

// sum of the per-wave partials in wave order (deterministic), then Adam
__global__ void k_sur_adam(const float *__restrict__ part, int nwaves, float *__restrict__ P,
                           float *__restrict__ m1, float *__restrict__ m2, float lr, float c1, float c2,
                           float *__restrict__ lossOut) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > NTRAIN) return;
    float gsum = 0.f;
    for (int w = 0; w < nwaves; ++w) gsum += part[(size_t)w * kGradStride + e];
    if (e == NTRAIN) {
        *lossOut = gsum;
        return;
    }
    const float b1 = 0.9f, b2 = 0.999f;
    const float a = b1 * m1[e] + (1.f - b1) * gsum;
    const float v = b2 * m2[e] + (1.f - b2) * gsum * gsum;
    m1[e] = a;
    m2[e] = v;
    P[e] -= lr * (a * c1) / (sqrtf(v * c2) + 1e-8f);
}

// the summed gradient (and loss) of one mini-batch, without an update
__global__ void k_sur_gsum(const float *__restrict__ part, int nwaves, float *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > NTRAIN) return;
    float gsum = 0.f;
    for (int w = 0; w < nwaves; ++w) gsum += part[(size_t)w * kGradStride + e];
    out[e] = gsum;
}

// sum over 3n force components of (a - b)^2 (fixed point to FP64), one
// workgroup: the lanes' partial sums are combined in a fixed order
__global__ void __launch_bounds__(256) k_force_sqdiff(int m, const unsigned long long *__restrict__ a,
                                                      const unsigned long long *__restrict__ b,
                                                      double *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const double d = (double)(long long)(a[i] - b[i]) * kFxInv;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// per-feature mean and standard deviation of the network inputs over the
// recorded samples (one block per input column, fixed reduction order)
__global__ void __launch_bounds__(256) k_sur_norm(int nsamp, const float *__restrict__ raw,
                                                  const float *__restrict__ target, float *__restrict__ P) {
    __shared__ double s1[256], s2[256];
    const int col = blockIdx.x;  // 0..NF-1 inputs, NF = target
    double a = 0.0, b = 0.0;
    for (int s = threadIdx.x; s < nsamp; s += blockDim.x) {
        float v;
        if (col == NF) {
            v = target[s];
        } else {
            v = feature_of(raw + (size_t)s * NR, col);
        }
        a += v;
        b += (double)v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = s1[0] / nsamp;
        const double var = fmax(s2[0] / nsamp - m * m, 0.0);
        const float sd = (float)fmax(sqrt(var), 1e-6);
        if (col == NF) {
            P[OEH] = (float)m;
            P[OES] = sd;
        } else {
            P[OMU + col] = (float)m;
            P[OSD + col] = sd;
        }
    }
}


}  // namespace

static_assert(OEH + 1 == kSurNParam, "packed surrogate layout");

void surrogate_features(Ctx &c, float *raw) {
    E3BState &E = c.e3b;
    hipLaunchKernelGGL(k_sur_feat, dim3(cdiv((long long)E.nmol * kFeatLanes, 256)), dim3(256), 0, c.stream, E.nmol,
                       E.molAtoms, c.xq, c.box, E.cap, E.nl, E.nlCount, raw);
}

void surrogate_forces(Ctx &c, bool energy, bool virial, bool with_feat) {
    SurrogateState &S = c.sur;
    E3BState &E = c.e3b;
    {
        ProfScope ps(c, PK_SUR_FEAT);
        const TwoBody tb{E.E2, E.k2, E.r2b * E.r2b};
        const dim3 g1(cdiv((long long)E.nmol * kFeatLanes, 256));
        float *feat = with_feat ? S.feat : nullptr;
#define MDX_FUSED(EV, VI)                                                                                \
    launch<sur_fused_body<EV, VI>, 256>(c, g1.x, 0, E.nmol, E.molAtoms, c.xq, c.box, E.cap, E.nl, E.nlCount, tb, \
                                        S.params, S.frag, feat, S.emol, S.wgt, c.force, c.acc)
        if (energy && virial) MDX_FUSED(true, true);
        else if (energy) MDX_FUSED(true, false);
        else if (virial) MDX_FUSED(false, true);
        else MDX_FUSED(false, false);
#undef MDX_FUSED
    }
    ProfScope ps(c, PK_SUR_FORCE);
    const int lanes = S.forceLanes;
    const dim3 grid(cdiv((long long)E.nmol * lanes, 256));
#define MDX_SFORCE(V, T)                                                                          \
    launch<sur_force_body<V, T>, 256>(c, grid.x, 0, E.nmol, E.molAtoms, c.xq, c.box, E.cap, E.nl, E.nlCount, \
                                      S.wgt, c.force, c.acc)
    switch (lanes) {
        case 8: if (virial) MDX_SFORCE(true, 8); else MDX_SFORCE(false, 8); break;
        case 16: if (virial) MDX_SFORCE(true, 16); else MDX_SFORCE(false, 16); break;
        case 64: if (virial) MDX_SFORCE(true, 64); else MDX_SFORCE(false, 64); break;
        default: if (virial) MDX_SFORCE(true, 32); else MDX_SFORCE(false, 32); break;
    }
#undef MDX_SFORCE
}

void surrogate_pack(Ctx &c) {
    SurrogateState &S = c.sur;
    if (!S.frag) S.frag = dalloc<unsigned short>(c, (size_t)FR_N * 64 * 4);
    hipLaunchKernelGGL(k_sur_pack, dim3(cdiv(FR_N * 64, 256)), dim3(256), 0, c.stream, S.params, S.frag);
}

void surrogate_normalize(Ctx &c) {
    SurrogateState &S = c.sur;
    hipLaunchKernelGGL(k_sur_norm, dim3(NF + 1), dim3(256), 0, c.stream, S.nsamp, S.trainX, S.trainT, S.params);
}

static void surrogate_train_alloc(Ctx &c) {
    SurrogateState &S = c.sur;
    if (!S.part) {
        S.part = dalloc<float>(c, (size_t)kTrainWaves * kGradStride);
        S.m1 = dalloc<float>(c, NTRAIN);
        S.m2 = dalloc<float>(c, NTRAIN);
        S.loss = dalloc<float>(c, 1);
        fill_float(c, S.m1, NTRAIN, 0.f);
        fill_float(c, S.m2, NTRAIN, 0.f);
    }
}

int surrogate_trained_count() { return NTRAIN; }

double surrogate_judge(Ctx &c) {
    SurrogateState &S = c.sur;
    MDX_REQUIRE(S.fSave && S.judgeN == c.n, 3, "judge buffers not set up");
    const size_t bytes = sizeof(unsigned long long) * 3 * (size_t)c.n;
    MDX_HIP(hipMemcpyAsync(S.fSave, c.force, bytes, hipMemcpyDeviceToDevice, c.stream));
    // both forms share the analytic two-body term: the difference of the
    // totals is the difference of the three-body forces
    prepare_forces(c);
    e3b_forces(c, false, false, true);
    MDX_HIP(hipMemcpyAsync(S.fE3B, c.force, bytes, hipMemcpyDeviceToDevice, c.stream));
    prepare_forces(c);
    surrogate_forces(c, false, false);
    hipLaunchKernelGGL(k_force_sqdiff, dim3(1), dim3(256), 0, c.stream, 3 * c.n, S.fE3B, c.force, S.judgeSum);
    MDX_HIP(hipMemcpyAsync(c.force, S.fSave, bytes, hipMemcpyDeviceToDevice, c.stream));
    double h = 0.0;
    MDX_HIP(hipMemcpyAsync(&h, S.judgeSum, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    MDX_HIP(hipStreamSynchronize(c.stream));
    return std::sqrt(h / (3.0 * c.n));
}

void surrogate_grad(Ctx &c, int bstart, int bsize, unsigned mul, unsigned add, float *out) {
    SurrogateState &S = c.sur;
    surrogate_train_alloc(c);
    hipLaunchKernelGGL(k_sur_grad, dim3(kTrainWaves / kWaves), dim3(64 * kWaves), 0, c.stream, S.nsamp, bstart, bsize,
                       mul, add, S.params, S.trainX, S.trainT, S.part);
    hipLaunchKernelGGL(k_sur_gsum, dim3(cdiv(NTRAIN + 1, 256)), dim3(256), 0, c.stream, S.part, kTrainWaves, out);
}

double surrogate_fit(Ctx &c, int epochs, int batch, double lr, unsigned seed) {
    SurrogateState &S = c.sur;
    surrogate_train_alloc(c);
    const int n = S.nsamp;
    batch = std::min(batch, n);
    double last = 0.0;
    float hl = 0.f;
    unsigned state = seed * 2654435761u + 12345u;
    for (int ep = 0; ep < epochs; ++ep) {
        // affine pseudo-permutation of the samples for this epoch (a odd and
        // coprime with n is enough for a bijection when n is a power of two;
        // otherwise some samples repeat within an epoch, which is harmless)
        state = state * 1664525u + 1013904223u;
        unsigned a = (state | 1u) % (unsigned)n;
        if (a == 0) a = 1;
        while (std::gcd(a, (unsigned)n) != 1) ++a;
        state = state * 1664525u + 1013904223u;
        const unsigned b = state % (unsigned)n;
        double esum = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 + batch <= n; b0 += batch) {
            ++S.step;
            const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)S.step));
            const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)S.step));
            hipLaunchKernelGGL(k_sur_grad, dim3(kTrainWaves / kWaves), dim3(64 * kWaves), 0, c.stream, n, b0, batch, a, b,
                               S.params, S.trainX, S.trainT, S.part);
            hipLaunchKernelGGL(k_sur_adam, dim3(cdiv(NTRAIN + 1, 256)), dim3(256), 0, c.stream, S.part,
                               kTrainWaves, S.params, S.m1, S.m2, (float)lr, c1, c2, S.loss);
            if (ep == epochs - 1) {
                MDX_HIP(hipMemcpyAsync(&hl, S.loss, sizeof(float), hipMemcpyDeviceToHost, c.stream));
                MDX_HIP(hipStreamSynchronize(c.stream));
                esum += hl;
                ++nb;
            }
        }
        if (nb) last = esum / nb;
    }
    surrogate_pack(c);
    return last;
}

}  // namespace mdx
