/*
 * mdx_oracle.c -- FP64 CPU restatement of DASH's per-step force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * md_engine_amd/: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product path never does.
 *
 * Every function cites the reference file:line whose arithmetic it restates
 * (paths relative to the dreid1991/md_engine source tree).  All arithmetic is
 * IEEE double.  The translation unit is compiled with -ffp-contract=off so the
 * canonical neighbor distance formula (orc_dist2) rounds exactly like the
 * device version in md_engine_amd/csrc/nlist.hip.
 *
 * Pinning (what fixes each function independently of this restatement) is
 * listed in the oracle's Python wrapper (oracle/__init__.py) and in tests/.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Box geometry                                                              */
/* ------------------------------------------------------------------------ */

/* BoundsGPU::minImage (src/BoundsGPU.h:91-99): v -= L * rint(v * invL) on
 * periodic axes.  FP64 overload used by SETTLE at BoundsGPU.h:103-107. */
static inline double orc_mi(double d, double L, double invL, int periodic) {
    if (periodic) {
        double img = rint(d * invL);
        d = d - img * L;
    }
    return d;
}

/* Canonical squared distance used for neighbor-list membership.  The device
 * builder evaluates the same expression in the same order with contraction
 * disabled, so membership decisions are bit-identical.  The reference tests
 * dot(x_j + shift - x_i) < r_list^2 (GridGPU.cu:326-340, 514-520); for
 * r_list < L/2 the shift it uses is the minimum image. */
static inline double orc_dist2(const double *xi, const double *xj,
                               const double *L, const double *invL,
                               const int *periodic) {
    double dx = orc_mi(xj[0] - xi[0], L[0], invL[0], periodic[0]);
    double dy = orc_mi(xj[1] - xi[1], L[1], invL[1], periodic[1]);
    double dz = orc_mi(xj[2] - xi[2], L[2], invL[2], periodic[2]);
    return dx * dx + dy * dy + dz * dz;
}

/* periodicWrap (GridGPU.cu:138-161) / BoundsGPU::wrapCoords (BoundsGPU.h:128-143):
 * x -= L * floor((x - lo) / L).  The division is replaced by the product with
 * the host-computed 1/L, identically on the device. */
void orc_wrap(int n, double *x, const double *lo, const double *L,
              const double *invL, const int *periodic) {
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            if (!periodic[a]) continue;
            double img = floor((x[3 * i + a] - lo[a]) * invL[a]);
            x[3 * i + a] = x[3 * i + a] - img * L[a];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Neighbor lists                                                            */
/* ------------------------------------------------------------------------ */

/* Brute-force full neighbor list: the reference's own CPU verifier
 * (GridGPU::verifyNeighborlists, GridGPU.cu:1109-1200) compares per-atom id
 * sets built by an O(N^2) scan.  Row i of nbr (capacity cap) holds j ascending;
 * j != i and orc_dist2 < rlist2.  Returns the maximum count (rows longer than
 * cap are truncated, counts[] still holds the full count). */
int orc_nlist_brute(int n, const double *x, const double *L, const double *invL,
                    const int *periodic, double rlist2, int cap, int *counts,
                    int *nbr) {
    int maxc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxc)
    for (int i = 0; i < n; ++i) {
        int c = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            if (orc_dist2(x + 3 * i, x + 3 * j, L, invL, periodic) < rlist2) {
                if (c < cap) nbr[(size_t)i * cap + c] = j;
                ++c;
            }
        }
        counts[i] = c;
        if (c > maxc) maxc = c;
    }
    return maxc;
}

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Cell-list full neighbor list with the same membership test as
 * orc_nlist_brute (used for large systems and the CPU baseline).  Cells of
 * edge >= r_list, 27-cell stencil with periodic wrap, as in countNumNeighbors /
 * assignNeighbors (GridGPU.cu:326-463, 466-734); positions must already be
 * wrapped.  Rows are sorted ascending. */
int orc_nlist_cell(int n, const double *x, const double *lo, const double *L,
                   const double *invL, const int *periodic, double rlist,
                   int cap, int *counts, int *nbr) {
    int nc[3];
    for (int a = 0; a < 3; ++a) {
        nc[a] = (int)floor(L[a] / (rlist * (1.0 + 1e-9)));
        if (nc[a] < 1) nc[a] = 1;
    }
    int ncell = nc[0] * nc[1] * nc[2];
    int *cell_of = (int *)malloc(sizeof(int) * (size_t)n);
    int *start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *order = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        int c[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = (int)floor((x[3 * i + a] - lo[a]) * invL[a] * nc[a]);
            if (c[a] < 0) c[a] = 0;
            if (c[a] >= nc[a]) c[a] = nc[a] - 1;
        }
        cell_of[i] = (c[0] * nc[1] + c[1]) * nc[2] + c[2];
        start[cell_of[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) start[c + 1] += start[c];
    {
        int *fill = (int *)malloc(sizeof(int) * (size_t)ncell);
        memcpy(fill, start, sizeof(int) * (size_t)ncell);
        for (int i = 0; i < n; ++i) order[fill[cell_of[i]]++] = i;
        free(fill);
    }
    double rlist2 = rlist * rlist;
    int maxc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxc)
    for (int i = 0; i < n; ++i) {
        int ci = cell_of[i];
        int c0[3] = {ci / (nc[1] * nc[2]), (ci / nc[2]) % nc[1], ci % nc[2]};
        /* unique neighbor cells per axis */
        int lst[3][3], nl[3];
        for (int a = 0; a < 3; ++a) {
            if (nc[a] < 3) {
                nl[a] = nc[a];
                for (int k = 0; k < nc[a]; ++k) lst[a][k] = k;
            } else {
                nl[a] = 3;
                for (int k = 0; k < 3; ++k)
                    lst[a][k] = (c0[a] + k - 1 + nc[a]) % nc[a];
            }
        }
        int c = 0;
        int *row = nbr + (size_t)i * cap;
        for (int ax = 0; ax < nl[0]; ++ax)
            for (int ay = 0; ay < nl[1]; ++ay)
                for (int az = 0; az < nl[2]; ++az) {
                    int cc = (lst[0][ax] * nc[1] + lst[1][ay]) * nc[2] + lst[2][az];
                    for (int s = start[cc]; s < start[cc + 1]; ++s) {
                        int j = order[s];
                        if (j == i) continue;
                        if (orc_dist2(x + 3 * i, x + 3 * j, L, invL, periodic) < rlist2) {
                            if (c < cap) row[c] = j;
                            ++c;
                        }
                    }
                }
        counts[i] = c;
        if (c > maxc) maxc = c;
        qsort(row, (size_t)(c < cap ? c : cap), sizeof(int), cmp_int);
    }
    free(cell_of);
    free(start);
    free(order);
    return maxc;
}

/* ------------------------------------------------------------------------ */
/* LJ + Ewald real-space pair forces                                         */
/* ------------------------------------------------------------------------ */

/* Fused pair loop of compute_force_iso / compute_energy_iso
 * (src/Evaluators/PairEvaluateIso.h:8-207, 212-358) with EvaluatorLJ
 * (PairEvaluatorLJ.h:17-60) and ChargeEvaluatorEwald (ChargeEvaluatorEwald.h:11-63).
 *
 * nbr rows hold neighbor index | tag << 30 (tag 0 none, 1 = 1-2, 2 = 1-3,
 * 3 = 1-4; globalDefs.h:40).  Multipliers {1, c12, c13, c14}
 * (PairEvaluateIso.h:31).  Per type pair (T x T, row-major): rc2, eps24 = 24 eps,
 * sig6 = sigma^6 (FixLJCut.cu:83-122).  dr = minImage(x_i - x_j).
 * Outputs: f (n x 3) forces, e (n) per-atom energies with the reference's 0.5
 * weights, vir[6] = 0.5 * sum F (x) dr ordered xx yy zz xy xz yz (helpers.h:34-41).
 * Any output pointer may be NULL. */
void orc_pair(int n, const double *x, const double *q, const int *type,
              int ntypes, const double *rc2, const double *eps24,
              const double *sig6, const double *special, int with_lj,
              int with_charge, double alpha, double qqr, double rq2,
              const double *L, const double *invL, const int *periodic,
              const int *counts, const uint32_t *nbr, int cap, double *f,
              double *e, double *vir) {
    double mult[4] = {1.0, special[0], special[1], special[2]};
    double vtot[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel
    {
        double vloc[6] = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            double fi[3] = {0, 0, 0}, ei = 0.0;
            int ti = type[i];
            const uint32_t *row = nbr + (size_t)i * cap;
            for (int k = 0; k < counts[i]; ++k) {
                uint32_t raw = row[k];
                int j = (int)(raw & 0x3FFFFFFFu);
                double m = mult[raw >> 30];
                double dr[3];
                for (int a = 0; a < 3; ++a)
                    dr[a] = orc_mi(x[3 * i + a] - x[3 * j + a], L[a], invL[a], periodic[a]);
                double lenSqr = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                int tp = ti * ntypes + type[j];
                double fs = 0.0;
                int computed = 0;
                if (with_lj && lenSqr < rc2[tp]) {
                    computed = 1;
                    if (m != 0.0) {
                        double p1 = eps24[tp] * 2.0 * sig6[tp] * sig6[tp];
                        double p2 = eps24[tp] * sig6[tp];
                        double r2inv = 1.0 / lenSqr;
                        double r6inv = r2inv * r2inv * r2inv;
                        fs += r6inv * r2inv * (p1 * r6inv - p2) * m;
                        double eps = eps24[tp] / 24.0;
                        double s6r6 = sig6[tp] * r6inv;
                        double rc6 = rc2[tp] * rc2[tp] * rc2[tp];
                        double s6rc6 = sig6[tp] / rc6;
                        double off = s6rc6 * (s6rc6 - 1.0);
                        ei += 0.5 * 4.0 * eps * (s6r6 * (s6r6 - 1.0) - off) * m;
                    }
                }
                if (with_charge && lenSqr < rq2) {
                    computed = 1;
                    double ls = lenSqr < 1e-10 ? 1e-10 : lenSqr;
                    double r2inv = 1.0 / ls;
                    double rinv = sqrt(r2inv);
                    double len = sqrt(ls);
                    double qq = qqr * q[i] * q[j];
                    double erfcv = erfc(alpha * len);
                    double s = qq * (erfcv * rinv + (2.0 * 0.5641895835477563 * alpha) *
                                                         exp(-alpha * alpha * ls));
                    if (m < 1.0) s -= (1.0 - m) * qq * rinv;
                    fs += s * r2inv;
                    double pref = qq / len;
                    double eng = pref * erfcv;
                    if (m < 1.0) eng -= (1.0 - m) * pref;
                    ei += 0.5 * eng;
                }
                if (computed) {
                    double F[3] = {dr[0] * fs, dr[1] * fs, dr[2] * fs};
                    fi[0] += F[0];
                    fi[1] += F[1];
                    fi[2] += F[2];
                    vloc[0] += 0.5 * F[0] * dr[0];
                    vloc[1] += 0.5 * F[1] * dr[1];
                    vloc[2] += 0.5 * F[2] * dr[2];
                    vloc[3] += 0.5 * F[0] * dr[1];
                    vloc[4] += 0.5 * F[0] * dr[2];
                    vloc[5] += 0.5 * F[1] * dr[2];
                }
            }
            if (f) {
                f[3 * i] = fi[0];
                f[3 * i + 1] = fi[1];
                f[3 * i + 2] = fi[2];
            }
            if (e) e[i] = ei;
        }
#pragma omp critical
        for (int a = 0; a < 6; ++a) vtot[a] += vloc[a];
    }
    if (vir)
        for (int a = 0; a < 6; ++a) vir[a] = vtot[a];
}

/* ------------------------------------------------------------------------ */
/* P3M (Deserno-Holm particle-particle particle-mesh)                        */
/* ------------------------------------------------------------------------ */

/* Charge-assignment weights of order p (Deserno & Holm 1998, App. E), as in
 * W_p_3 / W_p_5 / W_p_7 (src/Fixes/FixChargeEwald.cu:111-115, 163-172,
 * 220-234).  i runs over [-(p-1)/2, (p-1)/2], x in [-1/2, 1/2]. */
double orc_wp(int order, int i, double x) {
    double x2 = x * x;
    if (order == 1) return 1.0;
    if (order == 3) {
        if (i == -1) return 0.125 - 0.5 * x + 0.5 * x2;
        if (i == 0) return 0.75 - x2;
        return 0.125 + 0.5 * x + 0.5 * x2;
    }
    if (order == 5) {
        double x4 = x2 * x2;
        if (i == -2) return (1.0 / 384.0) * (1.0 - 8.0 * x + 24.0 * x2 - 32.0 * x * x2 + 16.0 * x4);
        if (i == -1) return (1.0 / 96.0) * (19.0 - 44.0 * x + 24.0 * x2 + 16.0 * x * x2 - 16.0 * x4);
        if (i == 0) return (1.0 / 192.0) * (115.0 - 120.0 * x2 + 48.0 * x4);
        if (i == 1) return (1.0 / 96.0) * (19.0 + 44.0 * x + 24.0 * x2 - 16.0 * x * x2 - 16.0 * x4);
        return (1.0 / 384.0) * (1.0 + 8.0 * x + 24.0 * x2 + 32.0 * x * x2 + 16.0 * x4);
    }
    /* order 7 */
    {
        double x4 = x2 * x2, x6 = x2 * x4, x3 = x2 * x, x5 = x3 * x2;
        if (i == 0) return (1.0 / 11520.0) * (5887.0 - 4620.0 * x2 + 1680.0 * x4 - 320.0 * x6);
        if (i == -3) return (1.0 / 46080.0) * (1.0 - 12.0 * x + 60.0 * x2 - 160.0 * x3 + 240.0 * x4 - 192.0 * x5 + 64.0 * x6);
        if (i == -2) return (1.0 / 23040.0) * (361.0 - 1416.0 * x + 2220.0 * x2 - 1600.0 * x3 + 240.0 * x4 + 384.0 * x5 - 192.0 * x6);
        if (i == -1) return (1.0 / 46080.0) * (10543.0 - 17340.0 * x + 4740.0 * x2 + 6880.0 * x3 - 4080.0 * x4 - 960.0 * x5 + 960.0 * x6);
        if (i == 1) return (1.0 / 46080.0) * (10543.0 + 17340.0 * x + 4740.0 * x2 - 6880.0 * x3 - 4080.0 * x4 + 960.0 * x5 + 960.0 * x6);
        if (i == 2) return (1.0 / 23040.0) * (361.0 + 1416.0 * x + 2220.0 * x2 + 1600.0 * x3 + 240.0 * x4 - 384.0 * x5 - 192.0 * x6);
        return (1.0 / 46080.0) * (1.0 + 12.0 * x + 60.0 * x2 + 160.0 * x3 + 240.0 * x4 + 192.0 * x5 + 64.0 * x6);
    }
}

/* amp_table (FixChargeEwald.cu:869-877): Deserno-Holm RMS error coefficients. */
static const double orc_amp_table[7][7] = {
    {2.0 / 3.0, 0, 0, 0, 0, 0, 0},
    {1.0 / 50.0, 5.0 / 294.0, 0, 0, 0, 0, 0},
    {1.0 / 588.0, 7.0 / 1440.0, 21.0 / 3872.0, 0, 0, 0, 0},
    {1.0 / 4320.0, 3.0 / 1936.0, 7601.0 / 2271360.0, 143.0 / 28800.0, 0, 0, 0},
    {1.0 / 23232.0, 7601.0 / 12628160.0, 143.0 / 69120.0, 517231.0 / 106536960.0,
     106640677.0 / 11737571328.0, 0, 0},
    {691.0 / 68140800.0, 13.0 / 57600.0, 47021.0 / 35512320.0, 9694607.0 / 2095994880.0,
     733191589.0 / 59609088000.0, 326190917.0 / 11700633600.0, 0},
    {1.0 / 345600.0, 3617.0 / 35512320.0, 745739.0 / 838397952.0, 56399353.0 / 12773376000.0,
     25091609.0 / 1560084480.0, 1755948832039.0 / 36229939200000.0,
     48887769399.0 / 37838389248.0}};

/* One entry of the RMS-error coefficient table (FixChargeEwald.cu:854-862),
 * for tests that perturb the table. */
double orc_pppm_amp(int order, int m) { return orc_amp_table[order - 1][m]; }

/* DeltaF_k (FixChargeEwald.cu:880-895). */
double orc_deltaf_k(double a, int natoms, double total_q2, const double *L,
                    const int *sz, int order) {
    double h[3] = {L[0] / sz[0], L[1] / sz[1], L[2] / sz[2]};
    double sum[3] = {0, 0, 0};
    for (int m = 0; m < order; ++m) {
        double amp = orc_amp_table[order - 1][m];
        for (int d = 0; d < 3; ++d) sum[d] += amp * pow(h[d] * a, 2 * m);
    }
    double acc = 0.0;
    for (int d = 0; d < 3; ++d)
        acc += 1.0 / (L[d] * L[d]) * pow(a * h[d], order) *
               sqrt(a * L[d] / natoms * sqrt(2.0 * M_PI) * sum[d]);
    return total_q2 / 3.0 * acc;
}

/* DeltaF_real (FixChargeEwald.cu:897-900). */
double orc_deltaf_real(double a, int natoms, double total_q2, const double *L,
                       double rcut) {
    return 2 * total_q2 / sqrt(natoms * rcut * L[0] * L[1] * L[2]) * exp(-a * a * rcut * rcut);
}

/* find_optimal_parameters (FixChargeEwald.cu:935-967): secant solve of
 * DeltaF_k(alpha) = DeltaF_real(alpha) from x_a = 0, x_b = 4.79853 / rcut,
 * relative tolerance 1e-5, at most 100 iterations.  total_q2 = qqr * sum q^2
 * (setTotalQ2, FixChargeEwald.cu:903-933); natoms counts every site, charged or
 * not.  Returns alpha; *err receives DeltaF_k + DeltaF_real. */
double orc_pppm_alpha(int natoms, double total_q2, const double *L, const int *sz,
                      int order, double rcut, double *err) {
    double xa = 0.0, xb = 4.79853 / rcut;
    double ya = orc_deltaf_k(xa, natoms, total_q2, L, sz, order) -
                orc_deltaf_real(xa, natoms, total_q2, L, rcut);
    double yb = orc_deltaf_k(xb, natoms, total_q2, L, sz, order) -
                orc_deltaf_real(xb, natoms, total_q2, L, rcut);
    double tol = 1e-5;
    int it = 0, maxit = 100;
    while ((fabs(yb) / orc_deltaf_real(xb, natoms, total_q2, L, rcut) > tol) && it < maxit) {
        double kinv = (xb - xa) / (yb - ya);
        ya = yb;
        xa = xb;
        xb = xa - ya * kinv;
        yb = orc_deltaf_k(xb, natoms, total_q2, L, sz, order) -
             orc_deltaf_real(xb, natoms, total_q2, L, rcut);
        ++it;
    }
    if (err)
        *err = orc_deltaf_k(xb, natoms, total_q2, L, sz, order) +
               orc_deltaf_real(xb, natoms, total_q2, L, rcut);
    return xb;
}

/* sinc with the reference's small-argument series (FixChargeEwald.cu:302-312). */
static double orc_sinc(double x) {
    if (x < 0.1 && x > -0.1) {
        double x2 = x * x;
        return 1.0 - x2 * 0.16666666667 + x2 * x2 * 0.008333333333333333 -
               x2 * x2 * x2 * 0.00019841269841269841;
    }
    return sin(x) / x;
}

static double orc_kvec(int id, int n, double L) {
    int m = id > n / 2 ? id - n : id;
    return 2.0 * ORC_PI * m / L;
}

/* Optimal influence function, Green_function_cu (FixChargeEwald.cu:314-366)
 * with sum_limits from calc_Green_function (FixChargeEwald.cu:1097-1103).
 * G is written for the full mesh, index (x * sy + y) * sz + z. */
void orc_pppm_green(const double *L, const int *sz, double alpha, int order,
                    double *G) {
    double h[3] = {L[0] / sz[0], L[1] / sz[1], L[2] / sz[2]};
    int lim = (int)(alpha * pow(h[0] * h[1] * h[2], 1.0 / 3.0) / 3.14159 *
                    (sqrt(-log(10E-7)))) + 1;
    double f4a2 = 0.25 / alpha / alpha;
#pragma omp parallel for collapse(2)
    for (int ix = 0; ix < sz[0]; ++ix)
        for (int iy = 0; iy < sz[1]; ++iy)
            for (int iz = 0; iz < sz[2]; ++iz) {
                double k[3] = {orc_kvec(ix, sz[0], L[0]), orc_kvec(iy, sz[1], L[1]),
                               orc_kvec(iz, sz[2], L[2])};
                double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                size_t idx = ((size_t)ix * sz[1] + iy) * sz[2] + iz;
                if (k2 == 0.0) {
                    G[idx] = 0.0;
                    continue;
                }
                double s1 = 0.0, s2 = 0.0;
                for (int mx = -lim; mx <= lim; ++mx)
                    for (int my = -lim; my <= lim; ++my)
                        for (int mz = -lim; mz <= lim; ++mz) {
                            double kp[3] = {k[0] + 2.0 * ORC_PI * mx / h[0],
                                            k[1] + 2.0 * ORC_PI * my / h[1],
                                            k[2] + 2.0 * ORC_PI * mz / h[2]};
                            double kl = kp[0] * kp[0] + kp[1] * kp[1] + kp[2] * kp[2];
                            double W = orc_sinc(kp[0] * h[0] * 0.5) *
                                       orc_sinc(kp[1] * h[1] * 0.5) *
                                       orc_sinc(kp[2] * h[2] * 0.5);
                            double W2 = pow(W, order * 2);
                            s1 += 4.0 * ORC_PI * exp(-kl * f4a2) *
                                  (k[0] * kp[0] + k[1] * kp[1] + k[2] * kp[2]) / kl * W2;
                            s2 += W2;
                        }
                G[idx] = s1 / (s2 * s2) / k2;
            }
}

/* In-place radix-2 complex FFT of length n (power of two) over data with
 * element stride `stride` (complex elements, interleaved re/im).  sign = -1 is
 * the forward transform exp(-2 pi i jk/n), +1 the unnormalised inverse, the
 * conventions of the 3-D C2C transforms at FixChargeEwald.cu:1309, 1344-1346. */
static void orc_fft1(double *data, int n, size_t stride, int sign, double *tmp) {
    for (int i = 0; i < n; ++i) {
        tmp[2 * i] = data[2 * i * stride];
        tmp[2 * i + 1] = data[2 * i * stride + 1];
    }
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = tmp[2 * i], ti = tmp[2 * i + 1];
            tmp[2 * i] = tmp[2 * j];
            tmp[2 * i + 1] = tmp[2 * j + 1];
            tmp[2 * j] = tr;
            tmp[2 * j + 1] = ti;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        double ang = sign * 2.0 * ORC_PI / len;
        for (int k = 0; k < len / 2; ++k) {
            double wr = cos(ang * k), wi = sin(ang * k);
            for (int i = 0; i < n; i += len) {
                int a = i + k, b = i + k + len / 2;
                double br = tmp[2 * b] * wr - tmp[2 * b + 1] * wi;
                double bi = tmp[2 * b] * wi + tmp[2 * b + 1] * wr;
                tmp[2 * b] = tmp[2 * a] - br;
                tmp[2 * b + 1] = tmp[2 * a + 1] - bi;
                tmp[2 * a] += br;
                tmp[2 * a + 1] += bi;
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        data[2 * i * stride] = tmp[2 * i];
        data[2 * i * stride + 1] = tmp[2 * i + 1];
    }
}

/* 3-D complex FFT of a (sx, sy, sz) row-major mesh (z fastest). */
void orc_fft3(double *c, const int *sz, int sign) {
    int nx = sz[0], ny = sz[1], nz = sz[2];
#pragma omp parallel
    {
        int nmax = nx > ny ? nx : ny;
        if (nz > nmax) nmax = nz;
        double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)nmax);
#pragma omp for collapse(2)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                orc_fft1(c + 2 * (((size_t)x * ny + y) * nz), nz, 1, sign, tmp);
#pragma omp for collapse(2)
        for (int x = 0; x < nx; ++x)
            for (int z = 0; z < nz; ++z)
                orc_fft1(c + 2 * ((size_t)x * ny * nz + z), ny, (size_t)nz, sign, tmp);
#pragma omp for collapse(2)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z)
                orc_fft1(c + 2 * ((size_t)y * nz + z), nx, (size_t)ny * nz, sign, tmp);
        free(tmp);
    }
}

static inline int orc_wrapi(int p, int n) {
    p %= n;
    return p < 0 ? p + n : p;
}

/* Mesh position of a (wrapped) coordinate: nearest point int((pos + h/2)/h) and
 * offset d = pos/h - ngp (map_charge_to_grid_order_*_cu,
 * FixChargeEwald.cu:174-190).  The reference subtracts box lo from the
 * unwrapped coordinate; this restatement wraps first (see orc_pppm). */
static void orc_mesh_point(double pos, double h, int *ngp, double *d) {
    *ngp = (int)((pos + 0.5 * h) / h);
    *d = pos / h - *ngp;
}

/* Full P3M long-range step in FP64, in the order of FixChargeEwald::compute
 * (FixChargeEwald.cu:1231-1478): spread (174-217), forward FFT, E_field_cu
 * (387-434), three inverse FFTs, gather (561-627).  Optional energy
 * (singlePointEng, 1481-1611: 0.5 sum |rho_k|^2 G / V - alpha/sqrt(pi) Q2) and
 * virial (virials_cu 731-802 with mapVirialToSingleAtom 815-818: 0.5 W / V).
 *
 * x is wrapped into [lo, lo+L) before assignment (the reference truncates
 * (pos - lo + h/2)/h toward zero, which misplaces coordinates that drifted more
 * than h/2 below lo between rebuilds).  G must be the full-mesh influence
 * function.  f may be NULL; energy/virial outputs may be NULL.  rho_out, when
 * non-NULL, receives the real spread mesh (sx*sy*sz). */
void orc_pppm(int n, const double *x, const double *q, const double *lo,
              const double *L, const int *sz, int order, double qqr,
              double alpha, double total_q2, const double *G, double *f,
              double *energy, double *vir, double *rho_out) {
    size_t ng = (size_t)sz[0] * sz[1] * sz[2];
    double h[3] = {L[0] / sz[0], L[1] / sz[1], L[2] / sz[2]};
    double invL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
    double V = L[0] * L[1] * L[2];
    double Qunit = sqrt(qqr);
    int half = (order - 1) / 2;
    double *rho = (double *)calloc(2 * ng, sizeof(double));
    for (int i = 0; i < n; ++i) {
        if (q[i] == 0.0) continue;
        double pos[3];
        int ngp[3];
        double d[3];
        for (int a = 0; a < 3; ++a) {
            double xa = x[3 * i + a];
            double img = floor((xa - lo[a]) * invL[a]);
            xa = xa - img * L[a];
            pos[a] = xa - lo[a];
            orc_mesh_point(pos[a], h[a], &ngp[a], &d[a]);
        }
        double qi = Qunit * q[i];
        for (int ix = -half; ix <= half; ++ix) {
            int px = orc_wrapi(ngp[0] + ix, sz[0]);
            double wx = qi * orc_wp(order, ix, d[0]);
            for (int iy = -half; iy <= half; ++iy) {
                int py = orc_wrapi(ngp[1] + iy, sz[1]);
                double wy = wx * orc_wp(order, iy, d[1]);
                for (int iz = -half; iz <= half; ++iz) {
                    int pz = orc_wrapi(ngp[2] + iz, sz[2]);
                    rho[2 * (((size_t)px * sz[1] + py) * sz[2] + pz)] += wy * orc_wp(order, iz, d[2]);
                }
            }
        }
    }
    if (rho_out)
        for (size_t g = 0; g < ng; ++g) rho_out[g] = rho[2 * g];
    orc_fft3(rho, sz, -1);
    if (energy || vir) {
        double es = 0.0, vs[6] = {0, 0, 0, 0, 0, 0};
        for (int ix = 0; ix < sz[0]; ++ix)
            for (int iy = 0; iy < sz[1]; ++iy)
                for (int iz = 0; iz < sz[2]; ++iz) {
                    size_t g = ((size_t)ix * sz[1] + iy) * sz[2] + iz;
                    double E = (rho[2 * g] * rho[2 * g] + rho[2 * g + 1] * rho[2 * g + 1]) * G[g];
                    es += E;
                    double k[3] = {orc_kvec(ix, sz[0], L[0]), orc_kvec(iy, sz[1], L[1]),
                                   orc_kvec(iz, sz[2], L[2])};
                    double kl = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                    if (kl == 0.0) continue;
                    double diff = -2.0 * (1.0 / kl + 0.25 / (alpha * alpha));
                    vs[0] += (1.0 + diff * k[0] * k[0]) * E;
                    vs[1] += (1.0 + diff * k[1] * k[1]) * E;
                    vs[2] += (1.0 + diff * k[2] * k[2]) * E;
                    vs[3] += diff * k[0] * k[1] * E;
                    vs[4] += diff * k[0] * k[2] * E;
                    vs[5] += diff * k[1] * k[2] * E;
                }
        if (energy) *energy = 0.5 * es / V - alpha / sqrt(M_PI) * total_q2;
        if (vir)
            for (int a = 0; a < 6; ++a) vir[a] = 0.5 * vs[a] / V;
    }
    if (f) {
        double *E3 = (double *)malloc(sizeof(double) * 2 * ng * 3);
        for (int ix = 0; ix < sz[0]; ++ix)
            for (int iy = 0; iy < sz[1]; ++iy)
                for (int iz = 0; iz < sz[2]; ++iz) {
                    size_t g = ((size_t)ix * sz[1] + iy) * sz[2] + iz;
                    double k[3] = {orc_kvec(ix, sz[0], L[0]), orc_kvec(iy, sz[1], L[1]),
                                   orc_kvec(iz, sz[2], L[2])};
                    double gr = rho[2 * g] * G[g], gi = rho[2 * g + 1] * G[g];
                    for (int a = 0; a < 3; ++a) {
                        /* E_hat = i k G rho_hat  (E_field_cu, FixChargeEwald.cu:426-431) */
                        E3[2 * (a * ng + g)] = -k[a] * gi;
                        E3[2 * (a * ng + g) + 1] = k[a] * gr;
                    }
                }
        for (int a = 0; a < 3; ++a) orc_fft3(E3 + 2 * a * ng, sz, +1);
        for (int i = 0; i < n; ++i) {
            f[3 * i] = f[3 * i + 1] = f[3 * i + 2] = 0.0;
            if (q[i] == 0.0) continue;
            double pos[3];
            int ngp[3];
            double d[3];
            for (int a = 0; a < 3; ++a) {
                double xa = x[3 * i + a];
                double img = floor((xa - lo[a]) * invL[a]);
                xa = xa - img * L[a];
                pos[a] = xa - lo[a];
                orc_mesh_point(pos[a], h[a], &ngp[a], &d[a]);
            }
            double Ev[3] = {0, 0, 0};
            for (int ix = -half; ix <= half; ++ix) {
                int px = orc_wrapi(ngp[0] + ix, sz[0]);
                double wx = orc_wp(order, ix, d[0]);
                for (int iy = -half; iy <= half; ++iy) {
                    int py = orc_wrapi(ngp[1] + iy, sz[1]);
                    double wy = wx * orc_wp(order, iy, d[1]);
                    for (int iz = -half; iz <= half; ++iz) {
                        int pz = orc_wrapi(ngp[2] + iz, sz[2]);
                        double w = wy * orc_wp(order, iz, d[2]);
                        size_t g = ((size_t)px * sz[1] + py) * sz[2] + pz;
                        for (int a = 0; a < 3; ++a) Ev[a] += w * (-E3[2 * (a * ng + g)] / V);
                    }
                }
            }
            double qi = Qunit * q[i];
            for (int a = 0; a < 3; ++a) f[3 * i + a] = qi * Ev[a];
        }
        free(E3);
    }
    free(rho);
}

/* Classical Ewald sum (real space over images within rcut_real, reciprocal
 * space over |m_i| <= kmax, self term), FP64.  Not a reference function: it is
 * the converged ground truth the P3M path (real + reciprocal) is pinned to.
 * Forces in f (n x 3), total energy returned.  qqr multiplies all terms. */
double orc_ewald_direct(int n, const double *x, const double *q, const double *L,
                        double alpha, double qqr, double rcut_real, int kmax,
                        double *f) {
    double V = L[0] * L[1] * L[2];
    double e = 0.0;
    for (int i = 0; i < 3 * n; ++i) f[i] = 0.0;
    int nim[3];
    for (int a = 0; a < 3; ++a) nim[a] = (int)ceil(rcut_real / L[a]) + 1;
    double rc2 = rcut_real * rcut_real;
    double ereal = 0.0;
#pragma omp parallel for reduction(+ : ereal) schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        double fi[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j) {
            for (int ix = -nim[0]; ix <= nim[0]; ++ix)
                for (int iy = -nim[1]; iy <= nim[1]; ++iy)
                    for (int iz = -nim[2]; iz <= nim[2]; ++iz) {
                        if (i == j && ix == 0 && iy == 0 && iz == 0) continue;
                        double dr[3] = {x[3 * i] - x[3 * j] + ix * L[0],
                                        x[3 * i + 1] - x[3 * j + 1] + iy * L[1],
                                        x[3 * i + 2] - x[3 * j + 2] + iz * L[2]};
                        double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                        if (r2 > rc2) continue;
                        double r = sqrt(r2);
                        double qq = qqr * q[i] * q[j];
                        ereal += 0.5 * qq * erfc(alpha * r) / r;
                        double s = qq * (erfc(alpha * r) / r + 2.0 * alpha / sqrt(ORC_PI) *
                                                                   exp(-alpha * alpha * r2)) / r2;
                        for (int a = 0; a < 3; ++a) fi[a] += s * dr[a];
                    }
        }
        for (int a = 0; a < 3; ++a) f[3 * i + a] += fi[a];
    }
    e += ereal;
    double erecip = 0.0;
    for (int mx = -kmax; mx <= kmax; ++mx)
        for (int my = -kmax; my <= kmax; ++my)
            for (int mz = -kmax; mz <= kmax; ++mz) {
                if (mx == 0 && my == 0 && mz == 0) continue;
                double k[3] = {2 * ORC_PI * mx / L[0], 2 * ORC_PI * my / L[1], 2 * ORC_PI * mz / L[2]};
                double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                double pre = 4.0 * ORC_PI / (V * k2) * exp(-k2 / (4 * alpha * alpha));
                double sr = 0.0, si = 0.0;
                for (int j = 0; j < n; ++j) {
                    double ph = k[0] * x[3 * j] + k[1] * x[3 * j + 1] + k[2] * x[3 * j + 2];
                    sr += q[j] * cos(ph);
                    si += q[j] * sin(ph);
                }
                erecip += 0.5 * qqr * pre * (sr * sr + si * si);
                for (int i = 0; i < n; ++i) {
                    double ph = k[0] * x[3 * i] + k[1] * x[3 * i + 1] + k[2] * x[3 * i + 2];
                    double s = qqr * pre * q[i] * (sin(ph) * sr - cos(ph) * si);
                    for (int a = 0; a < 3; ++a) f[3 * i + a] += s * k[a];
                }
            }
    e += erecip;
    double q2 = 0.0;
    for (int i = 0; i < n; ++i) q2 += q[i] * q[i];
    e -= qqr * alpha / sqrt(ORC_PI) * q2;
    return e;
}

/* ------------------------------------------------------------------------ */
/* E3B three-body water                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    double E2, Ea, Eb, Ec, k2, k3, rs, rf;
    double r2b;   /* two-body list cutoff: rf + padding (FixE3B.cu:29, 656-664) */
    double rmax;  /* three-body pair test RMAX_E3B (ThreeBodyE3B.h:15) */
} orc_e3b_params;

/* switching (EvaluatorE3B.h:91-104) */
static double orc_e3b_s(const orc_e3b_params *p, double r) {
    if (r <= p->rs) return 1.0;
    if (r > p->rf) return 0.0;
    double v = (p->rf - r);
    v *= v;
    return v * ((p->rf + 2.0 * r - 3.0 * p->rs) / ((p->rf - p->rs) * (p->rf - p->rs) * (p->rf - p->rs)));
}

/* dswitchdr (EvaluatorE3B.h:107-117) */
static double orc_e3b_ds(const orc_e3b_params *p, double r) {
    if (r <= p->rs || r > p->rf) return 0.0;
    return 6.0 * (p->rf - r) * (p->rs - r) / ((p->rf - p->rs) * (p->rf - p->rs) * (p->rf - p->rs));
}

/* f(r) = s(r) exp(-k3 r) (threeBodyPairEnergy, EvaluatorE3B.h:119-125) and its
 * force vector g = dr * (k3 e s - e s') / r (threeBodyForce, 127-141), with
 * dr = r_H - r_O so that g = -grad_H f. */
static double orc_e3b_f(const orc_e3b_params *p, const double *dr, double *g) {
    double r = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    if (r > p->rf) {
        if (g) g[0] = g[1] = g[2] = 0.0;
        return 0.0;
    }
    double es = exp(-p->k3 * r);
    double s = orc_e3b_s(p, r);
    if (g) {
        double fs = p->k3 * es / r;
        if (r > p->rs) {
            fs *= s;
            fs -= es * orc_e3b_ds(p, r) / r;
        }
        g[0] = dr[0] * fs;
        g[1] = dr[1] * fs;
        g[2] = dr[2] * fs;
    }
    return s * es;
}

static void orc_e3b_dr(const double *x, int a, int b, const double *L,
                       const double *invL, const int *per, double *dr) {
    for (int d = 0; d < 3; ++d)
        dr[d] = orc_mi(x[3 * a + d] - x[3 * b + d], L[d], invL[d], per[d]);
}

/* Molecule list of E3B: for each molecule the molecules whose oxygen lies
 * within r2b of its own, ascending (FixE3B's grid over the oxygen positions,
 * FixE3B.cu:67-83 and 656-680, with cutoff rf + padding).  A cell grid of
 * edge >= r2b over the oxygens when every axis is periodic and holds at
 * least 3 cells; otherwise every pair is scanned.  Rows are sorted so that
 * the sums over a row run in the same order as an all-pairs scan.  Returns
 * the CSR offsets (nmol + 1) and fills *nbr (caller frees both). */
static int orc_cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int *orc_e3b_list(int nmol, const int *mol, const double *x, const double *L,
                         const double *invL, const int *per, double r2b, int **nbr) {
    const double r2 = r2b * r2b;
    int nc[3];
    int cells = 1;
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(L[d] / r2b);
        cells = cells && per[d] && nc[d] >= 3;
    }
    int cap = 64;
    int *cnt = (int *)calloc((size_t)nmol, sizeof(int));
    int *tmp = NULL;
    int *head = NULL, *next = NULL, *cellOf = NULL;
    if (cells) {
        const int ncell = nc[0] * nc[1] * nc[2];
        head = (int *)malloc(sizeof(int) * (size_t)ncell);
        next = (int *)malloc(sizeof(int) * (size_t)nmol);
        cellOf = (int *)malloc(sizeof(int) * (size_t)nmol);
        for (int c = 0; c < ncell; ++c) head[c] = -1;
        for (int i = nmol - 1; i >= 0; --i) {
            int cc[3];
            for (int d = 0; d < 3; ++d) {
                double u = x[3 * mol[4 * i] + d] * invL[d];
                u -= floor(u);
                int k = (int)(u * nc[d]);
                cc[d] = k >= nc[d] ? nc[d] - 1 : k;
            }
            const int c = (cc[0] * nc[1] + cc[1]) * nc[2] + cc[2];
            cellOf[i] = c;
            next[i] = head[c];
            head[c] = i;
        }
    }
    for (;;) {
        int overflow = 0;
        tmp = (int *)malloc(sizeof(int) * (size_t)nmol * cap);
#pragma omp parallel for schedule(dynamic, 32) reduction(| : overflow)
        for (int i = 0; i < nmol; ++i) {
            int *row = tmp + (size_t)i * cap;
            int k = 0;
            if (cells) {
                const int c = cellOf[i];
                const int cx = c / (nc[1] * nc[2]), cy = (c / nc[2]) % nc[1], cz = c % nc[2];
                for (int ax = -1; ax <= 1; ++ax)
                    for (int ay = -1; ay <= 1; ++ay)
                        for (int az = -1; az <= 1; ++az) {
                            const int nx = (cx + ax + nc[0]) % nc[0], ny = (cy + ay + nc[1]) % nc[1],
                                      nz = (cz + az + nc[2]) % nc[2];
                            for (int j = head[(nx * nc[1] + ny) * nc[2] + nz]; j >= 0; j = next[j]) {
                                if (j == i) continue;
                                double dr[3];
                                orc_e3b_dr(x, mol[4 * i], mol[4 * j], L, invL, per, dr);
                                if (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] < r2) {
                                    if (k < cap) row[k] = j;
                                    ++k;
                                }
                            }
                        }
            } else {
                for (int j = 0; j < nmol; ++j) {
                    if (j == i) continue;
                    double dr[3];
                    orc_e3b_dr(x, mol[4 * i], mol[4 * j], L, invL, per, dr);
                    if (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] < r2) {
                        if (k < cap) row[k] = j;
                        ++k;
                    }
                }
            }
            cnt[i] = k;
            if (k > cap) overflow = 1;
            else qsort(row, (size_t)k, sizeof(int), orc_cmp_int);
        }
        if (!overflow) break;
        free(tmp);
        cap *= 4;
    }
    int *off = (int *)malloc(sizeof(int) * ((size_t)nmol + 1));
    off[0] = 0;
    for (int i = 0; i < nmol; ++i) off[i + 1] = off[i] + cnt[i];
    *nbr = (int *)malloc(sizeof(int) * ((size_t)off[nmol] + 1));
    for (int i = 0; i < nmol; ++i) memcpy(*nbr + off[i], tmp + (size_t)i * cap, sizeof(int) * (size_t)cnt[i]);
    free(tmp);
    free(cnt);
    free(head);
    free(next);
    free(cellOf);
    return off;
}

/* E3B forces and energies (FixE3B::compute, FixE3B.cu:85-347, kernels
 * compute_E3B_force_twobody / _threebody, ThreeBodyE3B.h:18-549, force assembly
 * threeBodyPairForce, EvaluatorE3B.h:145-249).  mol holds (O, H1, H2, M) atom
 * indices per molecule.  Pairs come from the molecule list (orc_e3b_list);
 * orc_e3b_allpairs runs the same sums over an O(M^2) scan.
 *
 * Two-body: for every ordered molecule pair with |r_OO| < r2b (the molecule
 * list membership, rebuilt every step in the reference), F_O += dr k2 E2
 * exp(-k2 r)/r and E += 0.5 E2 exp(-k2 r) (EvaluatorE3B.h:67-87).
 * Three-body: for pairs with |r_OO| <= rmax, per-molecule totals
 * x = sum f(H_j -> O_i) (accepted), y/z = sum f(H1_i / H2_i -> O_j) (donated);
 * energy E_A + E_B + E_C with E_A = Ea sum_d (y_d z_d - sum_a f1_da f2_da),
 * E_B = Eb sum_m (x_m (y_m + z_m) - sum_a D(m->a) D(a->m)),
 * E_C = Ec/2 sum_a (x_a^2 - sum_d D(d->a)^2), D(d->a) = f1_da + f2_da; this is
 * the sum of the reference's per-pair decomposition (EvaluatorE3B.h:251-345).
 * Forces: each f(H_d^h -> O_a) carries the coefficient c = dE/df
 * (the fscale values of threeBodyPairForce) and adds c g to H and -c g to O.
 * f (natoms x 3) is ADDED to.  e2, e3 receive the energies (may be NULL). */
static void orc_e3b_core(int nmol, const int *mol, const double *x, const double *L,
                         const int *per, const orc_e3b_params *p, double *f, double *e2,
                         double *e3, const int *off, const int *nbr) {
    double invL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
    double *tot = (double *)calloc((size_t)nmol * 3, sizeof(double));
    double r2b2 = p->r2b * p->r2b;
    double etwo = 0.0;
    /* pass 1: two-body force + totals (full scan per molecule, owner computes) */
#pragma omp parallel for reduction(+ : etwo) schedule(dynamic, 16)
    for (int i = 0; i < nmol; ++i) {
        int Oi = mol[4 * i], H1i = mol[4 * i + 1], H2i = mol[4 * i + 2];
        double fO[3] = {0, 0, 0};
        double t[3] = {0, 0, 0};
        const int jn = off ? off[i + 1] - off[i] : nmol;
        for (int jj = 0; jj < jn; ++jj) {
            const int j = off ? nbr[off[i] + jj] : jj;
            if (j == i) continue;
            int Oj = mol[4 * j], H1j = mol[4 * j + 1], H2j = mol[4 * j + 2];
            double dr[3];
            orc_e3b_dr(x, Oi, Oj, L, invL, per, dr);
            double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            if (r2 < r2b2) {
                double r = sqrt(r2);
                double ex = exp(-p->k2 * r);
                double s = p->k2 * p->E2 * ex / r;
                for (int d = 0; d < 3; ++d) fO[d] += dr[d] * s;
                etwo += 0.5 * p->E2 * ex;
            }
            double r = sqrt(r2);
            if (r <= p->rmax) {
                double d1[3], d2[3];
                orc_e3b_dr(x, H1j, Oi, L, invL, per, d1);
                orc_e3b_dr(x, H2j, Oi, L, invL, per, d2);
                t[0] += orc_e3b_f(p, d1, NULL) + orc_e3b_f(p, d2, NULL);
                orc_e3b_dr(x, H1i, Oj, L, invL, per, d1);
                orc_e3b_dr(x, H2i, Oj, L, invL, per, d2);
                t[1] += orc_e3b_f(p, d1, NULL);
                t[2] += orc_e3b_f(p, d2, NULL);
            }
        }
        for (int d = 0; d < 3; ++d) f[3 * Oi + d] += fO[d];
        tot[3 * i] = t[0];
        tot[3 * i + 1] = t[1];
        tot[3 * i + 2] = t[2];
    }
    /* pass 2: three-body forces on my own O, H1, H2 (owner computes) */
    double ea = 0.0, eb = 0.0, ec = 0.0;
#pragma omp parallel for reduction(+ : ea, eb, ec) schedule(dynamic, 16)
    for (int i = 0; i < nmol; ++i) {
        int Oi = mol[4 * i], H1i = mol[4 * i + 1], H2i = mol[4 * i + 2];
        double xi = tot[3 * i], yi = tot[3 * i + 1], zi = tot[3 * i + 2];
        double FO[3] = {0, 0, 0}, FH1[3] = {0, 0, 0}, FH2[3] = {0, 0, 0};
        double sum_f1f2 = 0.0, sum_DD = 0.0, sum_D2in = 0.0;
        const int jn = off ? off[i + 1] - off[i] : nmol;
        for (int jj = 0; jj < jn; ++jj) {
            const int j = off ? nbr[off[i] + jj] : jj;
            if (j == i) continue;
            int Oj = mol[4 * j], H1j = mol[4 * j + 1], H2j = mol[4 * j + 2];
            double dr[3];
            orc_e3b_dr(x, Oi, Oj, L, invL, per, dr);
            double r = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
            if (!(r <= p->rmax)) continue;
            double xj = tot[3 * j], yj = tot[3 * j + 1], zj = tot[3 * j + 2];
            /* i donates: f(H1_i -> O_j), f(H2_i -> O_j) */
            double dA1[3], dA2[3], gA1[3], gA2[3];
            orc_e3b_dr(x, H1i, Oj, L, invL, per, dA1);
            orc_e3b_dr(x, H2i, Oj, L, invL, per, dA2);
            double fA1 = orc_e3b_f(p, dA1, gA1), fA2 = orc_e3b_f(p, dA2, gA2);
            /* j donates to i: f(H1_j -> O_i), f(H2_j -> O_i) */
            double dB1[3], dB2[3], gB1[3], gB2[3];
            orc_e3b_dr(x, H1j, Oi, L, invL, per, dB1);
            orc_e3b_dr(x, H2j, Oi, L, invL, per, dB2);
            double fB1 = orc_e3b_f(p, dB1, gB1), fB2 = orc_e3b_f(p, dB2, gB2);
            double Dij = fA1 + fA2, Dji = fB1 + fB2;
            /* coefficients dE/df for the four bonds (d, h, a) */
            double cA1 = p->Ea * (zi - fA2) + p->Eb * (xi + yj + zj - 2.0 * Dji) + p->Ec * (xj - Dij);
            double cA2 = p->Ea * (yi - fA1) + p->Eb * (xi + yj + zj - 2.0 * Dji) + p->Ec * (xj - Dij);
            double cB1 = p->Ea * (zj - fB2) + p->Eb * (xj + yi + zi - 2.0 * Dij) + p->Ec * (xi - Dji);
            double cB2 = p->Ea * (yj - fB1) + p->Eb * (xj + yi + zi - 2.0 * Dij) + p->Ec * (xi - Dji);
            for (int d = 0; d < 3; ++d) {
                FH1[d] += cA1 * gA1[d];
                FH2[d] += cA2 * gA2[d];
                FO[d] -= cB1 * gB1[d] + cB2 * gB2[d];
            }
            sum_f1f2 += fA1 * fA2;
            sum_DD += Dij * Dji;
            sum_D2in += Dji * Dji;
        }
        for (int d = 0; d < 3; ++d) {
            f[3 * Oi + d] += FO[d];
            f[3 * H1i + d] += FH1[d];
            f[3 * H2i + d] += FH2[d];
        }
        ea += p->Ea * (yi * zi - sum_f1f2);
        eb += p->Eb * (xi * (yi + zi) - sum_DD);
        ec += 0.5 * p->Ec * (xi * xi - sum_D2in);
    }
    free(tot);
    if (e2) *e2 = etwo;
    if (e3) *e3 = ea + eb + ec;
}

void orc_e3b(int nmol, const int *mol, const double *x, const double *L,
             const int *per, const orc_e3b_params *p, double *f, double *e2,
             double *e3) {
    double invL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
    int *nbr = NULL;
    int *off = orc_e3b_list(nmol, mol, x, L, invL, per, p->r2b > p->rmax ? p->r2b : p->rmax, &nbr);
    orc_e3b_core(nmol, mol, x, L, per, p, f, e2, e3, off, nbr);
    free(off);
    free(nbr);
}

void orc_e3b_allpairs(int nmol, const int *mol, const double *x, const double *L,
                      const int *per, const orc_e3b_params *p, double *f, double *e2,
                      double *e3) {
    orc_e3b_core(nmol, mol, x, L, per, p, f, e2, e3, NULL, NULL);
}

/* Brute-force E3B three-body energy straight from the bond-pair definitions
 * (Tainter, Shi & Skinner, JCTC 11, 2268 (2015)): every intermolecular
 * hydrogen bond b = (donor d, hydrogen h, acceptor a) with weight f(r_{H O});
 * every unordered pair of distinct bonds is type A (same donor, different
 * hydrogens, different acceptors), type B (the acceptor of one is the donor of
 * the other and the outer molecules differ) or type C (same acceptor,
 * different donors).  Independent of the totals decomposition in orc_e3b; used
 * to pin it.  O(B^2) in the number of bonds: small systems only. */
double orc_e3b_brute_energy(int nmol, const int *mol, const double *x,
                            const double *L, const int *per,
                            const orc_e3b_params *p) {
    double invL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
    int cap = 2 * nmol * nmol;
    int *bd = (int *)malloc(sizeof(int) * (size_t)cap * 3);
    double *bf = (double *)malloc(sizeof(double) * (size_t)cap);
    int nb = 0;
    for (int d = 0; d < nmol; ++d)
        for (int h = 1; h <= 2; ++h)
            for (int a = 0; a < nmol; ++a) {
                if (a == d) continue;
                double dro[3];
                orc_e3b_dr(x, mol[4 * d], mol[4 * a], L, invL, per, dro);
                double roo = sqrt(dro[0] * dro[0] + dro[1] * dro[1] + dro[2] * dro[2]);
                if (!(roo <= p->rmax)) continue;
                double dr[3];
                orc_e3b_dr(x, mol[4 * d + h], mol[4 * a], L, invL, per, dr);
                double fv = orc_e3b_f(p, dr, NULL);
                if (fv == 0.0) continue;
                bd[3 * nb] = d;
                bd[3 * nb + 1] = h;
                bd[3 * nb + 2] = a;
                bf[nb] = fv;
                ++nb;
            }
    double e = 0.0;
    for (int b1 = 0; b1 < nb; ++b1)
        for (int b2 = b1 + 1; b2 < nb; ++b2) {
            int d1 = bd[3 * b1], h1 = bd[3 * b1 + 1], a1 = bd[3 * b1 + 2];
            int d2 = bd[3 * b2], h2 = bd[3 * b2 + 1], a2 = bd[3 * b2 + 2];
            double ff = bf[b1] * bf[b2];
            if (d1 == d2 && h1 != h2 && a1 != a2) e += p->Ea * ff;
            else if ((a1 == d2 && d1 != a2) || (a2 == d1 && d2 != a1)) e += p->Eb * ff;
            else if (a1 == a2 && d1 != d2) e += p->Ec * ff;
        }
    free(bd);
    free(bf);
    return e;
}

/* ------------------------------------------------------------------------ */
/* Velocity Verlet (LJ fluid)                                                */
/* ------------------------------------------------------------------------ */

/* NVE velocity Verlet for a single-type LJ fluid, the step of
 * IntegratorVerlet::run (IntegratorVerlet.cu:649-770) with preForce_cu
 * (307-338: v += dt/2 F/m, x += dt v) and postForce_cu (610-627:
 * v += dt/2 F/m), ftm_to_v = 1 (LJ units).  Neighbor list rebuilt when any
 * atom moved more than skin/2 since the last build (setBuildFlag,
 * GridGPU.cu:755-774), checked every step.  Cell-list neighbor search with
 * r_list = rc + skin.  Writes the per-step total energy (kinetic + potential)
 * into etot[nsteps+1] when non-NULL; returns the number of rebuilds. */
int orc_lj_verlet(int n, double *x, double *v, double mass, const double *lo,
                  const double *L, double eps, double sigma, double rc,
                  double skin, double dt, int nsteps, double *etot) {
    int per[3] = {1, 1, 1};
    double invL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
    double rlist = rc + skin;
    int cap = 512;
    int *counts = (int *)malloc(sizeof(int) * (size_t)n);
    int *nbr = (int *)malloc(sizeof(int) * (size_t)n * cap);
    double *xlast = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *e = (double *)malloc(sizeof(double) * (size_t)n);
    double rc2 = rc * rc, eps24 = 24.0 * eps, sig6 = pow(sigma, 6.0);
    double special[3] = {0, 0, 0.5};
    double q0 = 0.0;
    (void)q0;
    int type0 = 0;
    (void)type0;
    int *types = (int *)calloc((size_t)n, sizeof(int));
    double *qs = (double *)calloc((size_t)n, sizeof(double));
    int rebuilds = 0;
    double thr = 0.25 * skin * skin;
#define ORC_BUILD()                                                              \
    do {                                                                         \
        orc_wrap(n, x, lo, L, invL, per);                                        \
        orc_nlist_cell(n, x, lo, L, invL, per, rlist, cap, counts, nbr);         \
        memcpy(xlast, x, sizeof(double) * 3 * (size_t)n);                        \
        ++rebuilds;                                                              \
    } while (0)
#define ORC_FORCE()                                                              \
    orc_pair(n, x, qs, types, 1, &rc2, &eps24, &sig6, special, 1, 0, 0.0, 1.0,   \
             0.0, L, invL, per, counts, (const uint32_t *)nbr, cap, f, e, NULL)
    ORC_BUILD();
    ORC_FORCE();
    if (etot) {
        double ke = 0.0, pe = 0.0;
        for (int i = 0; i < n; ++i) {
            ke += 0.5 * mass * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
            pe += e[i];
        }
        etot[0] = ke + pe;
    }
    for (int s = 0; s < nsteps; ++s) {
        int need = 0;
        for (int i = 0; i < 3 * n; ++i) {
            v[i] += 0.5 * dt * f[i] / mass;
            x[i] += dt * v[i];
        }
        for (int i = 0; i < n && !need; ++i)
            if (orc_dist2(x + 3 * i, xlast + 3 * i, L, invL, per) > thr) need = 1;
        if (need) ORC_BUILD();
        ORC_FORCE();
        for (int i = 0; i < 3 * n; ++i) v[i] += 0.5 * dt * f[i] / mass;
        if (etot) {
            double ke = 0.0, pe = 0.0;
            for (int i = 0; i < n; ++i) {
                ke += 0.5 * mass * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
                pe += e[i];
            }
            etot[s + 1] = ke + pe;
        }
    }
#undef ORC_BUILD
#undef ORC_FORCE
    free(counts);
    free(nbr);
    free(xlast);
    free(f);
    free(e);
    free(types);
    free(qs);
    return rebuilds;
}

/* OpenMP threads of the following calls (the CPU baseline's 1-thread leg). */
void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Rigid water: SETTLE, RATTLE velocities, M site                            */
/* ------------------------------------------------------------------------ */

static inline void v3sub(const double *a, const double *b, double *o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static inline double v3dot(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void v3cross(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void v3norm(double *a) {
    double s = 1.0 / sqrt(v3dot(a, a));
    a[0] *= s;
    a[1] *= s;
    a[2] *= s;
}

/* SETTLE for one water (Miyamoto & Kollman, J. Comput. Chem. 13, 952 (1992)),
 * the algorithm of settlePositions (src/Fixes/FixRigid.cu:850-1136).  Inputs
 * relative to the old oxygen: old positions A0 = 0, B0, C0 (constrained) and
 * unconstrained new positions A1, B1, C1; outputs the constrained new
 * positions in A1, B1, C1.  mO, mH masses, dOH, dHH distances. */
void orc_settle_one(const double *B0, const double *C0, double *A1, double *B1, double *C1,
                    double mO, double mH, double dOH, double dHH) {
    double mT = mO + 2.0 * mH, wO = mO / mT, wH = mH / mT;
    double rc = 0.5 * dHH;
    double dmid = sqrt(dOH * dOH - rc * rc);
    double ra = 2.0 * wH * dmid, rb = dmid - ra;
    double com[3], a1[3], b1[3], c1[3];
    for (int k = 0; k < 3; ++k) com[k] = wO * A1[k] + wH * (B1[k] + C1[k]);
    v3sub(A1, com, a1);
    v3sub(B1, com, b1);
    v3sub(C1, com, c1);
    double ez[3], ex[3], ey[3];
    v3cross(B0, C0, ez);
    v3cross(a1, ez, ex);
    v3cross(ez, ex, ey);
    v3norm(ez);
    v3norm(ex);
    v3norm(ey);
    double xb0 = v3dot(B0, ex), yb0 = v3dot(B0, ey), xc0 = v3dot(C0, ex), yc0 = v3dot(C0, ey);
    double za1 = v3dot(a1, ez);
    double xb1 = v3dot(b1, ex), yb1 = v3dot(b1, ey), zb1 = v3dot(b1, ez);
    double xc1 = v3dot(c1, ex), yc1 = v3dot(c1, ey), zc1 = v3dot(c1, ez);
    double sinphi = za1 / ra, cosphi = sqrt(1.0 - sinphi * sinphi);
    double sinpsi = (zb1 - zc1) / (2.0 * rc * cosphi), cospsi = sqrt(1.0 - sinpsi * sinpsi);
    double ya2 = ra * cosphi, xb2 = -rc * cospsi;
    double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
    double yb2 = t1 - t2, yc2 = t1 + t2;
    double alpha = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
    double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
    double gamma = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
    double ab2 = alpha * alpha + beta * beta;
    double sinth = (alpha * gamma - beta * sqrt(ab2 - gamma * gamma)) / ab2;
    double costh = sqrt(1.0 - sinth * sinth);
    double a3[3] = {-ya2 * sinth, ya2 * costh, za1};
    double b3[3] = {xb2 * costh - yb2 * sinth, xb2 * sinth + yb2 * costh, zb1};
    double c3[3] = {-xb2 * costh - yc2 * sinth, -xb2 * sinth + yc2 * costh, zc1};
    for (int k = 0; k < 3; ++k) {
        A1[k] = com[k] + a3[0] * ex[k] + a3[1] * ey[k] + a3[2] * ez[k];
        B1[k] = com[k] + b3[0] * ex[k] + b3[1] * ey[k] + b3[2] * ez[k];
        C1[k] = com[k] + c3[0] * ex[k] + c3[1] * ey[k] + c3[2] * ez[k];
    }
}

/* Iterative SHAKE (Ryckaert, Ciccotti & Berendsen 1977) on the same three
 * distance constraints, corrections along the old bond vectors; converged to
 * tol.  Independent of SETTLE; used to pin it. */
void orc_shake_one(const double *B0, const double *C0, double *A1, double *B1, double *C1,
                   double mO, double mH, double dOH, double dHH, double tol, int maxit) {
    double *p[3] = {A1, B1, C1};
    double o[3][3] = {{0, 0, 0}, {B0[0], B0[1], B0[2]}, {C0[0], C0[1], C0[2]}};
    double m[3] = {mO, mH, mH};
    int bi[3] = {0, 0, 1}, bj[3] = {1, 2, 2};
    double d[3] = {dOH, dOH, dHH};
    for (int it = 0; it < maxit; ++it) {
        double worst = 0.0;
        for (int k = 0; k < 3; ++k) {
            int i = bi[k], j = bj[k];
            double r[3], r0[3];
            v3sub(p[i], p[j], r);
            v3sub(o[i], o[j], r0);
            double diff = d[k] * d[k] - v3dot(r, r);
            if (fabs(diff) / (d[k] * d[k]) > worst) worst = fabs(diff) / (d[k] * d[k]);
            double g = diff / (2.0 * v3dot(r, r0) * (1.0 / m[i] + 1.0 / m[j]));
            for (int a = 0; a < 3; ++a) {
                p[i][a] += g * r0[a] / m[i];
                p[j][a] -= g * r0[a] / m[j];
            }
        }
        if (worst < tol) break;
    }
}

/* Velocity constraint for a rigid water (settleVelocities,
 * src/Fixes/FixRigid.cu:658-849): the three bond-relative velocities
 * (v_i - v_j).(r_i - r_j) are made zero by momentum-conserving impulses along
 * the bonds, solved exactly (3x3, Cramer).  r relative positions O, H1, H2
 * (any origin), v in/out. */
void orc_rattle_one(const double *rO, const double *rH1, const double *rH2, double *vO,
                    double *vH1, double *vH2, double mO, double mH) {
    const double *r[3] = {rO, rH1, rH2};
    double *v[3] = {vO, vH1, vH2};
    double im[3] = {1.0 / mO, 1.0 / mH, 1.0 / mH};
    int bi[3] = {0, 0, 1}, bj[3] = {1, 2, 2};
    double e[3][3], M[3][3], rhs[3];
    for (int k = 0; k < 3; ++k) v3sub(r[bi[k]], r[bj[k]], e[k]);
    for (int k = 0; k < 3; ++k) {
        double dv[3];
        v3sub(v[bi[k]], v[bj[k]], dv);
        rhs[k] = -v3dot(dv, e[k]);
        for (int l = 0; l < 3; ++l) {
            double si = (bi[l] == bi[k]) ? 1.0 : (bj[l] == bi[k] ? -1.0 : 0.0);
            double sj = (bi[l] == bj[k]) ? 1.0 : (bj[l] == bj[k] ? -1.0 : 0.0);
            M[k][l] = v3dot(e[k], e[l]) * (si * im[bi[k]] - sj * im[bj[k]]);
        }
    }
    double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                 M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                 M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    double lam[3];
    for (int c = 0; c < 3; ++c) {
        double A[3][3];
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) A[k][l] = (l == c) ? rhs[k] : M[k][l];
        lam[c] = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                  A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                  A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0])) / det;
    }
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
            v[bi[k]][a] += lam[k] * e[k][a] * im[bi[k]];
            v[bj[k]][a] -= lam[k] * e[k][a] * im[bj[k]];
        }
}

/* Rigid-water velocity Verlet step parts on whole arrays (FP64), the
 * sequence of IntegratorVerlet::run with FixRigid (postNVE_X settles positions
 * and places M, FixRigid.cu:1234-1279; preStepFinal distributes the M force,
 * settles velocities and places M, FixRigid.cu:1916-1984).  mol: (O, H1, H2,
 * M) indices, M = -1 for three-site water.  gamma: M = gamma O + (1-gamma)/2
 * (H1 + H2).  Forces f are the raw per-site forces; the M force is
 * distributed inside.  phase 1: kick + drift + SETTLE + velocity correction +
 * M; phase 2: kick + RATTLE. */
static void orc_fdist(const double *f, const int *m, double gamma, double fO[3], double fH1[3],
                      double fH2[3]) {
    for (int a = 0; a < 3; ++a) {
        double fm = m[3] >= 0 ? f[3 * m[3] + a] : 0.0;
        fO[a] = f[3 * m[0] + a] + gamma * fm;
        fH1[a] = f[3 * m[1] + a] + 0.5 * (1.0 - gamma) * fm;
        fH2[a] = f[3 * m[2] + a] + 0.5 * (1.0 - gamma) * fm;
    }
}

void orc_water_step_half(int phase, int nmol, const int *mol, double *x, double *v,
                         const double *f, double mO, double mH, double dOH, double dHH,
                         double gamma, double dt, double ftm, const double *L, const int *per) {
    double invL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nmol; ++i) {
        const int *m = mol + 4 * i;
        double fO[3], fH1[3], fH2[3];
        orc_fdist(f, m, gamma, fO, fH1, fH2);
        double *F[3] = {fO, fH1, fH2};
        double ms[3] = {mO, mH, mH};
        for (int s = 0; s < 3; ++s)
            for (int a = 0; a < 3; ++a) v[3 * m[s] + a] += 0.5 * dt * ftm * F[s][a] / ms[s];
        double rO[3], rH1[3], rH2[3];
        for (int a = 0; a < 3; ++a) {
            rO[a] = 0.0;
            rH1[a] = orc_mi(x[3 * m[1] + a] - x[3 * m[0] + a], L[a], invL[a], per[a]);
            rH2[a] = orc_mi(x[3 * m[2] + a] - x[3 * m[0] + a], L[a], invL[a], per[a]);
        }
        if (phase == 1) {
            double A1[3], B1[3], C1[3];
            for (int a = 0; a < 3; ++a) {
                A1[a] = dt * v[3 * m[0] + a];
                B1[a] = rH1[a] + dt * v[3 * m[1] + a];
                C1[a] = rH2[a] + dt * v[3 * m[2] + a];
            }
            double A3[3] = {A1[0], A1[1], A1[2]}, B3[3] = {B1[0], B1[1], B1[2]}, C3[3] = {C1[0], C1[1], C1[2]};
            orc_settle_one(rH1, rH2, A3, B3, C3, mO, mH, dOH, dHH);
            double xO[3];
            for (int a = 0; a < 3; ++a) {
                v[3 * m[0] + a] += (A3[a] - A1[a]) / dt;
                v[3 * m[1] + a] += (B3[a] - B1[a]) / dt;
                v[3 * m[2] + a] += (C3[a] - C1[a]) / dt;
                xO[a] = x[3 * m[0] + a] + A3[a];
            }
            for (int a = 0; a < 3; ++a) {
                x[3 * m[0] + a] = xO[a];
                x[3 * m[1] + a] = xO[a] + (B3[a] - A3[a]);
                x[3 * m[2] + a] = xO[a] + (C3[a] - A3[a]);
                if (m[3] >= 0)
                    x[3 * m[3] + a] = xO[a] + 0.5 * (1.0 - gamma) * ((B3[a] - A3[a]) + (C3[a] - A3[a]));
            }
        } else {
            orc_rattle_one(rO, rH1, rH2, v + 3 * m[0], v + 3 * m[1], v + 3 * m[2], mO, mH);
        }
        if (m[3] >= 0) v[3 * m[3]] = v[3 * m[3] + 1] = v[3 * m[3] + 2] = 0.0;
    }
}
