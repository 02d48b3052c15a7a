/*
 * mdx.h -- C ABI of libmdx, the MI355X-native per-step force path of a
 * DASH-style (dreid1991/md_engine) molecular-dynamics engine.
 *
 * Plain C types only: pointers to host arrays, sizes and scalars.  All arrays
 * that describe atoms are indexed by atom id 0..n-1 (the library keeps its own
 * cell-ordered copy on the device).  Every entry point returns 0 on success or
 * one of the MDX_ERR_* codes; mdx_last_error() gives the message.  No C++
 * exception crosses this boundary.  Work is ordered on the context's own HIP
 * stream; results reach the host only through the *_get calls.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the md_engine source tree).
 */
#ifndef MDX_H
#define MDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdx_ctx mdx_ctx;

#define MDX_OK 0
#define MDX_ERR_ARG 1      /* invalid argument (size, style, cutoff, ...) */
#define MDX_ERR_HIP 2      /* HIP / hipFFT runtime failure */
#define MDX_ERR_STATE 3    /* call out of order (e.g. no system set) */
#define MDX_ERR_INTERNAL 4

/* force terms for mdx_forces_compute / mdx_run */
#define MDX_PAIR 1      /* LJ + Ewald real space (FixLJCut with offloaded charge pairs) */
#define MDX_PPPM 2      /* P3M mesh part (FixChargeEwald) */
#define MDX_E3B 4       /* E3B two- and three-body (FixE3B) */
#define MDX_ENERGY 8    /* also accumulate energies */
#define MDX_BONDED 64   /* intramolecular terms of flexible water (mdx_flexible_water_set) */
#define MDX_VIRIAL 16   /* also accumulate the virial tensor */
#define MDX_SURROGATE 32 /* E3B with the learned three-body term (mdx_surrogate_set) */
#define MDX_VIRIAL_TRACE 4096 /* only the virial's trace where a term has a trace-only form (pair, E3B):
                                 what a barostat reads (FixPressureBerendsen's scalar pressure);
                                 read it with mdx_virial_trace_get */
#define MDX_PPPM_PER_BEAD 8192 /* ring polymers: every bead its own P3M instead of the centroid contraction */
#define MDX_RING_BEAD_LISTS 16384 /* ring polymers: every bead its own neighbor list instead of one on the centroids */

int mdx_version(void);

/* Context on HIP device `device`; replaces State construction plus
 * DeviceManager::setDevice (src/State.cpp:40-80, src/DeviceManager.cpp). */
int mdx_create(int device, mdx_ctx **out);
int mdx_destroy(mdx_ctx *ctx);
const char *mdx_last_error(const mdx_ctx *ctx);

/* Atoms and box.  Replaces State::prepareForRun's packing of the device
 * arrays (src/State.cpp:562-647, GPUData.h:15-71): x, v, inv_mass (vs.w),
 * q, type (xs.w).  x/v are n x 3 FP64, by id.  periodic[3] per axis. */
int mdx_system_set(mdx_ctx *ctx, int n, const double *x, const double *v, const double *inv_mass,
                   const double *q, const int32_t *type, const double lo[3], const double len[3],
                   const int32_t periodic[3]);
/* New positions (by id), velocities kept; the list is revalidated on use. */
int mdx_positions_set(mdx_ctx *ctx, const double *x);
/* Positions / velocities by id (either pointer may be NULL). */
int mdx_system_get(mdx_ctx *ctx, double *x, double *v);
/* Orthorhombic box (BoundsGPU, src/BoundsGPU.h:91-154). */
int mdx_box_set(mdx_ctx *ctx, const double lo[3], const double len[3]);
/* Current box (it changes under a barostat). */
int mdx_box_get(mdx_ctx *ctx, double lo[3], double len[3]);

/* Neighbor list.  Replaces GridGPU (src/GridGPU.cu:84-113, 824-1103):
 * r_list = rCut + padding, rebuild when any atom moved more than skin/2
 * (setBuildFlag, GridGPU.cu:755-774). */
int mdx_nlist_config(mdx_ctx *ctx, double r_list, double skin);
/* Exclusion CSR by id: offsets[n_ids+1], entries other_id | tag << 30 with
 * tag 1/2/3 = 1-2/1-3/1-4 (GridGPU::handleExclusionsDistance, GridGPU.cu:1377-1434). */
int mdx_exclusions_set(mdx_ctx *ctx, int n_ids, const int32_t *offsets, const uint32_t *tagged_ids);
/* Same, from bonds (pairs of ids) by breadth-first search to max_depth
 * (GridGPU::generateExclusionList, GridGPU.cu:1458-1518). */
int mdx_exclusions_from_bonds(mdx_ctx *ctx, int nbonds, const int32_t *bonds, int max_depth);
/* Rebuild now (force != 0) or when the skin criterion says so; *rebuilt set
 * to 1 when a build happened (periodicBoundaryConditions, GridGPU.cu:824-1103). */
int mdx_nlist_update(mdx_ctx *ctx, int force, int *rebuilt);
/* Per-id neighbor lists, ids ascending, each entry neighbor_id | tag << 30:
 * row r of tagged_ids (row_cap entries) belongs to atom id r.  *max_count
 * receives the longest row; rows longer than row_cap are truncated.
 * Replaces GridGPU::getNeighborList (GridGPU.cu:1520-1648). */
int mdx_nlist_get(mdx_ctx *ctx, int32_t *counts, uint32_t *tagged_ids, int row_cap, int *max_count);
/* Rebuild count since system_set, row capacity, cells per axis, and the
 * number of mdx_run steps whose list was older than the skin allows (a
 * displacement above skin/2 that the rebuild decision missed; 0 in a valid
 * run).  Any pointer may be NULL. */
int mdx_nlist_stats(mdx_ctx *ctx, int64_t *rebuilds, int *row_cap, int32_t cells[3],
                    int64_t *violations);

/* LJ type-pair tables, ntypes x ntypes row-major: rc^2, 24 eps, sigma^6
 * (FixLJCut::prepareForRun, src/Fixes/FixLJCut.cu:83-122). */
int mdx_lj_set(mdx_ctx *ctx, int ntypes, const double *rc2, const double *eps24, const double *sig6);
/* 1-2 / 1-3 / 1-4 multipliers (State::specialNeighborCoefs, State.cpp:55-57). */
int mdx_special_set(mdx_ctx *ctx, double c12, double c13, double c14);
/* Ewald real-space term fused into the pair kernel (ChargeEvaluatorEwald,
 * src/Evaluators/ChargeEvaluatorEwald.h:11-63).  alpha <= 0 takes the P3M
 * alpha once mdx_pppm_set is active. */
int mdx_ewald_real_set(mdx_ctx *ctx, int enable, double alpha, double qqr, double r_cut);

/* P3M mesh (FixChargeEwald::setParameters / setError, FixChargeEwald.cu:969-1085).
 * Mesh sizes must be powers of two in [32, 1024] as in the reference; order
 * 1, 3, 5 or 7; alpha <= 0 solves for it (find_optimal_parameters, 935-967). */
int mdx_pppm_set(mdx_ctx *ctx, int nx, int ny, int nz, int order, double r_cut, double qqr,
                 double alpha);
/* Current alpha and the RMS error estimate DeltaF_k + DeltaF_real. */
int mdx_pppm_info(mdx_ctx *ctx, double *alpha, double *rms_error);
/* Kernels the last mesh evaluation ran (no counterpart in the reference;
 * for tests and benchmarks): spread 2 owner-computes cell blocks, 1 LDS tiles
 * with global atomics, 0 global atomics; gather 2 LDS field tiles, 1 row
 * teams from the per-site mesh records, 0 row teams from the positions; fft
 * 1 the mesh solve in three LDS passes (power-of-two meshes whose planes fit
 * a workgroup), 0 hipFFT with separate field and conversion kernels; -1
 * before any evaluation.  Any pointer may be NULL. */
int mdx_pppm_paths(mdx_ctx *ctx, int *spread, int *gather, int *fft);
/* Error mode (FixChargeEwald::setError + setGridToErrorTolerance,
 * FixChargeEwald.cu:1024-1096): the mesh starts from the one set before
 * (else 64^3) and the axis with the fewest points per length is doubled until
 * DeltaF_k + DeltaF_real at the solved alpha is at most `error`; the search is
 * repeated whenever the box changed (at the start of a run or force call).
 * Meshes beyond 1024 points per axis are refused. */
int mdx_pppm_set_error(mdx_ctx *ctx, double error, double r_cut, int order, double qqr);
/* Long-range interval (FixChargeEwald::setLongRangeInterval,
 * FixChargeEwald.cu:1251, 1402-1451, 1614-1618): the mesh part runs on every
 * interval-th step of a run, counted from its first, and stores its forces by
 * atom id; the steps between add the stored forces (applyStoredForces,
 * FixChargeEwald.cu:804-814).  0 keeps the current value.  Energy and virial
 * requests always run the mesh.  Runs with an interval above 1 take eager
 * steps (captured steps evaluate the mesh every step). */
int mdx_pppm_interval_set(mdx_ctx *ctx, int interval);
/* Mesh in use (after an error-mode search) and the long-range interval. */
int mdx_pppm_mesh_get(mdx_ctx *ctx, int32_t sz[3], int *interval);

/* E3B water (FixE3B::addMolecule / createEvaluator, src/Fixes/FixE3B.cu:515-593,
 * 753-792).  mol_ids: 4 ids per molecule ordered O, H1, H2, M.  params:
 * E2, Ea, Eb, Ec (energy units), k2, k3, rs, rf (length units); r_two_body is
 * the O-O cutoff of the two-body term (rf + padding = 6.2 A in the reference),
 * r_max the three-body pair test (6.1572 A). */
int mdx_e3b_set(mdx_ctx *ctx, int nmol, const int32_t *mol_ids, const double params[8],
                double r_two_body, double r_max);

/* Rigid water (FixRigid::createRigid with a TIP4P/2005-style style,
 * src/Fixes/FixRigid.cu:1372-1379, 1631-1681): mol_ids 4 per molecule, O, H1,
 * H2, M (M = -1 for three-site water); masses, the O-H and H-H distances kept
 * by SETTLE, and gamma with M = gamma O + (1 - gamma)/2 (H1 + H2).  mdx_run
 * then integrates these sites with SETTLE / RATTLE in FP64 and places M. */
int mdx_rigid_set(mdx_ctx *ctx, int nmol, const int32_t *mol_ids, double m_O, double m_H, double d_OH,
                  double d_HH, double gamma);

/* Thermostat applied every `every` steps of mdx_run: tau <= 0 rescales the
 * velocities to `temperature` (FixNVTRescale, src/Fixes/FixNVTRescale.cu),
 * tau > 0 is Berendsen weak coupling with time constant tau.  Degrees of
 * freedom: 3 per massive site - 3 per rigid molecule - 3.  boltz and
 * mvv_to_eng are the unit constants (Units.cpp:5-32). */
int mdx_thermostat_set(mdx_ctx *ctx, int enable, double temperature, int every, double tau,
                       double boltz, double mvv_to_eng);
/* Instantaneous temperature and kinetic energy (FP64 reduction on the device;
 * DataComputerTemperature). */
int mdx_temperature(mdx_ctx *ctx, double boltz, double mvv_to_eng, double *temperature,
                    double *kinetic);

/* Zero and accumulate the requested force terms at the current positions
 * (IntegratorUtil::force, src/Integrators/IntegratorUtil.cpp:15-30). */
int mdx_forces_compute(mdx_ctx *ctx, int flags);
/* Same as mdx_forces_compute, without allocating or freeing device memory:
 * no counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops).  The P3M work arrays (charge
 * mesh, spectra, field) come from a library-wide device scratch that holds
 * meshes of up to 64^3 points (about 14 MiB per GPU, kept until the process
 * ends); larger meshes are rejected with MDX_ERR_ARG.  The context must be
 * set up by one mdx_forces_compute call first (list rows, P3M plans); a call
 * that would need new memory (first build, rows that overflow, a regridded
 * cell list) fails with MDX_ERR_STATE instead.  Results equal
 * mdx_forces_compute's.  Every call on a GPU shares its scratch: calls that
 * overlap on one GPU (contexts on different host threads) race. */
int mdx_forces_compute_noalloc(mdx_ctx *ctx, int flags);
/* Forces by id, n x 3 FP64 (converted from the 64-bit fixed-point buffer). */
int mdx_forces_get(mdx_ctx *ctx, double *f);
/* Energies of the last compute with MDX_ENERGY: e[0] LJ + real space,
 * e[1] P3M reciprocal incl. self term, e[2] E3B two-body, e[3] E3B three-body;
 * virial[6] xx yy zz xy xz yz (sum F (x) dr). */
int mdx_energy_get(mdx_ctx *ctx, double e[4], double virial[6]);
/* Trace of the virial of the last compute with MDX_VIRIAL or MDX_VIRIAL_TRACE
 * (xx + yy + zz of the tensor plus the trace-only terms), the scalar the
 * pressure computes from (the scalar virial sum of DataComputerPressure::computeScalar_GPU,
 * src/DataStorageUser/DataComputerPressure.cu:26-91). */
int mdx_virial_trace_get(mdx_ctx *ctx, double *trace);
/* Flexible q-TIP4P/F water, replacing the rigid model: FixBondQuartic
 * (src/Fixes/FixBondQuartic.cu; U = k2 dr^2 + k3 dr^3 + k4 dr^4 per O-H
 * bond, src/Evaluators/BondEvaluatorQuartic.h:9-36), FixAngleHarmonic
 * (src/Fixes/FixAngleHarmonic.cu; U = ktheta/2 (theta - theta0)^2 per H-O-H,
 * src/Evaluators/AngleEvaluatorHarmonic.h:12-78) and the massless M site of
 * FixTIP4PFlexible (src/Fixes/FixTIP4PFlexible.cu:92-300: placed at
 * gamma r_O + (1 - gamma)/2 (r_H1 + r_H2) before every force evaluation, its
 * force spread back onto O and H).  mol_ids: nmol x (O, H1, H2, M) ids (M -1:
 * three-site).  The intramolecular forces enter compute under MDX_BONDED;
 * their energy is read with mdx_bonded_energy_get. */
int mdx_flexible_water_set(mdx_ctx *ctx, int nmol, const int32_t *mol_ids, double k2, double k3, double k4,
                           double r0, double ktheta, double theta0, double gamma);
int mdx_bonded_energy_get(mdx_ctx *ctx, double *e);

/* Velocity Verlet (IntegratorVerlet::run, src/Integrators/IntegratorVerlet.cu:649-770):
 * dt and the force-to-acceleration unit factor (Units::ftm_to_v). */
int mdx_integrator_set(mdx_ctx *ctx, double dt, double ftm_to_v);
/* Advance nsteps with the force terms in `flags`.  With option "graph" = 1
 * the steps are recorded once into a hipGraph (8 steps per graph, rebuild
 * and thermostat decided on the device each step) and replayed. */
int mdx_run(mdx_ctx *ctx, int nsteps, int flags);
/* Batched replicas: advances every context (same device, each with its own
 * system) by nsteps with one captured graph per 8 steps that holds all of
 * them.  When the replicas run the same kernels (same force terms and set-up,
 * any sizes) every kernel of a step is ONE launch over all replicas, each
 * workgroup reading its replica's arguments from a table (option
 * "batch_merge" 0 on ctxs[0]: one stream per replica instead).  No
 * counterpart in the reference, whose integrator runs one system per
 * process; results per replica equal a single-context captured run bit for
 * bit.  Errors are reported on ctxs[0]. */
int mdx_run_batch(mdx_ctx *const *ctxs, int n, int nsteps, int flags);
/* 1 if the last batch led by ctx ran as one launch per kernel, else 0. */
int mdx_batch_merged(mdx_ctx *ctx, int *merged);
/* Path-integral MD: P bead contexts (the same atoms, box, time step and
 * force terms; flexible molecules, no barostat) advanced as one ring polymer
 * per atom, like IntegratorVerlet with nPerRingPoly = P
 * (src/Integrators/IntegratorVerlet.cu:63-304, 342-593): every step each
 * bead is kicked by half a step of its own forces, the free ring polymer of
 * every atom is propagated exactly for dt in normal modes (frequencies
 * w_k = 2 omegaP sin(k pi / P); the reference takes omegaP = kB T / hbar at
 * its bead temperature T = P T_physical, IntegratorVerlet.cu:794), every bead
 * evaluates its forces and takes the second half kick and its thermostat.
 * The mesh part is contracted to the ring's centroid as the reference does
 * (src/Fixes/FixChargeEwald.cu:1255-1266, 1395-1405): charges spread from
 * each atom's centroid (computeCentroid, :22-50) on bead 0's mesh, the field
 * interpolated there, the whole force on every bead; alpha is solved over
 * all P x n ring atoms with one bead's Q^2 (:866-900).  MDX_PPPM_PER_BEAD in
 * flags gives every bead its own uncontracted P3M instead.  The neighbor
 * list is built on the centroids as the reference builds it
 * (src/GridGPU.cu:163-189, 888-895): atoms are binned and paired by the
 * centroid of their ring, every bead evaluates the same id pairs at its own
 * positions (bead k with bead k, PairEvaluateIso.h:77-78, 123-124), and all
 * beads rebuild together when any bead has moved half the skin since the last
 * build (setBuildFlag over all P n atoms); no E3B then.  MDX_RING_BEAD_LISTS
 * gives every bead its own list on its own positions instead.  P: a power of
 * two from 2 to 64.  The beads run on bead 0's stream inside captured
 * graphs. */
int mdx_ring_run(mdx_ctx *const *beads, int P, int nsteps, int flags, double omegaP);
/* Forces of every bead of a ring at the current positions, as one step of
 * mdx_ring_run evaluates them (contracted mesh part unless
 * MDX_PPPM_PER_BEAD); read them per bead with mdx_forces_get.  The ring
 * force evaluation of IntegratorUtil::force with nPerRingPoly = P
 * (src/Integrators/IntegratorUtil.cpp:15-30). */
int mdx_ring_forces_compute(mdx_ctx *const *beads, int P, int flags);
/* Launch geometry knobs (Tunable nThreadPerAtom, src/Tunable.h:3-24):
 * lanes per atom in the list build and the pair kernel, lanes per molecule in
 * E3B.  0 keeps the current value.  Powers of two only. */
int mdx_teams_set(mdx_ctx *ctx, int build_team, int pair_team, int e3b_team);
/* Learned E3B three-body surrogate.  No counterpart in the reference (its
 * FixE3B, src/Fixes/FixE3B.cu, only has the analytic three-body term); a
 * model replaces that term under MDX_SURROGATE while the two-body E3B term
 * stays analytic.  Model: per-molecule descriptors (2 radial functions
 * exp(-kappa r) fc(r), kappa = k3 and 2 k3 - the E3B bond function f and
 * f^2 up to the switch - on the H..O distances in three channels: accepted,
 * H1-donated, H2-donated; three pair products of f over each neighbour:
 * both own H to one O, both H of one neighbour to the own O, mutual
 * donation), network input of 12 features (the channels, d1 + d2, d1 * d2,
 * the pair products and three products of the f channels; the E3B
 * three-body energy of a molecule is linear in them), MLP 12 -> 64 -> 64 -> 1
 * with SiLU on BF16 MFMA tiles, forces from its hand-written backward pass.
 * params: kSurNParam = 5083 floats packed as W1[12][64] b1[64] W2[64][64]
 * b2[64] w3[64] b3 mu[12] sd[12] escale eshift.
 * record: appends the current frame (raw channels + the E3B three-body energy
 * owned by each molecule, the training target) to the device sample set;
 * fit: Adam on the per-molecule energy MSE (normalize: set mu, sd, escale,
 * eshift from the samples first); eval: per-molecule energies (nmol) and raw
 * descriptor channels (nmol x 9), either may be NULL. */
int mdx_surrogate_set(mdx_ctx *ctx, const float *params, int nparams);
int mdx_surrogate_get(mdx_ctx *ctx, float *params, int nparams);
int mdx_surrogate_record(mdx_ctx *ctx, int reset, int capacity);
int mdx_surrogate_fit(mdx_ctx *ctx, int epochs, int batch, double lr, int seed, int normalize, double *loss);
int mdx_surrogate_eval(mdx_ctx *ctx, float *emol, float *feat);
/* E3B as the judge of the learned term: during mdx_run with MDX_SURROGATE,
 * every `every` steps (0: never) the analytic E3B three-body forces and the
 * surrogate's are both evaluated on the current frame; when their RMSE over
 * atoms and components exceeds max_rmse (force units) the rest of the run,
 * and later runs, fall back to E3B (MDX_E3B) until the judge is set again.
 * Steps run in chunks that end on the judge points.  stats: checks so far,
 * the last and the worst RMSE, whether the run fell back.  No counterpart in
 * the reference (its FixE3B has only the analytic term). */
int mdx_surrogate_judge_set(mdx_ctx *ctx, int every, double max_rmse);
int mdx_surrogate_judge_stats(mdx_ctx *ctx, int64_t *checks, double *last_rmse, double *worst_rmse, int *fell_back);
/* Gradient of the training loss (mean squared error of the standardised
 * per-molecule energy) over one mini-batch of recorded samples, rows
 * (bstart + s) * mul + add mod nsamp for s < bsize, as mdx_surrogate_fit
 * computes it before its Adam update: grad[0 .. 5057) for W1 b1 W2 b2 w3 b3
 * in the packed layout, grad[5057] the batch loss (n >= 5058).  For checking
 * the hand-written backward pass against an FP64 restatement. */
int mdx_surrogate_grad(mdx_ctx *ctx, int bstart, int bsize, uint32_t mul, uint32_t add, float *grad, int n);
/* E3B three-body energy owned by each molecule (donor of type A, centre of
 * type B, acceptor of type C; the per-molecule shares of
 * compute_E3B_energy_threebody, src/Evaluators/ThreeBodyE3B.h:553-918). */
int mdx_e3b_molecule_energies(mdx_ctx *ctx, float *emol);
/* Nose-Hoover chain thermostat (FixNoseHoover, src/Fixes/FixNoseHoover.cu:
 * 120-1098; setTemperature(T, timeConstant), chainLength 10 there): half a
 * chain step before the first kick and after the second, scaling all
 * velocities; chain state stays on the device.  Excludes mdx_thermostat_set.
 * chain_energy: sum_k Q_k v_k^2 / 2 + ndf kT xi_0 + sum_{k>0} kT xi_k, the
 * thermostat part of the conserved extended energy. */
int mdx_nose_hoover_set(mdx_ctx *ctx, int enable, double temperature, double time_constant, int chain_length,
                        double boltz, double mvv_to_eng);
int mdx_nose_hoover_energy(mdx_ctx *ctx, double *chain_energy);
/* Raw chain state (n >= 4 chain_length + 2 doubles): chain velocities,
 * forces, masses, positions, the current sum m v^2 (energy units) and the last
 * velocity scale factor; the thermVel / thermForce / thermMass vectors and
 * ke_current of FixNoseHoover (src/Fixes/FixNoseHoover.cu:795-866). */
int mdx_nose_hoover_state(mdx_ctx *ctx, double *state, int n);
/* Berendsen barostat (FixPressureBerendsen(state, handle, pressure, period,
 * applyEvery), src/Fixes/FixPressureBerendsen.cu; bulkModulus 10 and
 * maxDilation 1e-5 there): every `every` steps the molecular pressure sets an
 * isotropic dilation of the box and of the molecules' centres of mass.
 * Steps run eagerly while it is on.  mdx_pressure: instantaneous pressure of
 * the current state with the force terms in `flags` (molecular virial, same
 * units as nktv_to_press); mdx_barostat_last: the pressure of its last
 * application. */
int mdx_barostat_set(mdx_ctx *ctx, int enable, double pressure, double period, int every, double bulk_modulus,
                     double max_dilation, double nktv_to_press, double mvv_to_eng);
int mdx_pressure(mdx_ctx *ctx, int flags, double nktv_to_press, double mvv_to_eng, double *pressure);
int mdx_barostat_last(mdx_ctx *ctx, double *pressure);
/* Autotuner (Integrator::tune, src/Integrators/Integrator.cu:272-376, grid
 * search of Tunable knobs by timing): times `evals` runs of every list-build
 * kernel/team width, pair team width and E3B team width for the terms in
 * `flags`, keeps the fastest, and recomputes forces.  chosen[4] (may be NULL)
 * = build mode (2, the per-atom team build), build team, pair team, E3B team;
 * avg_ms[3] (may be NULL) = the winners' average build, pair, E3B times. */
int mdx_tune(mdx_ctx *ctx, int flags, int evals, int32_t *chosen, double *avg_ms);
/* Autotuner of a batch: like the reference's tuning during a run
 * (Integrator.cu:272-376, real steps timed between candidates), times
 * `steps` captured steps of the whole batch (mdx_run_batch, merged launches)
 * for every candidate of the pair team width, the E3B team width and the
 * list-build kernel (wave per atom, or teams of 8 or 16 lanes), one knob
 * after the other, and keeps the fastest for every replica, so the batch
 * still runs one launch per kernel whatever its size classes.  The timed
 * steps advance the replicas (2 x steps per candidate).  chosen[3] (may be
 * NULL) = build team (64: wave per atom), pair team, E3B team; ms_per_step
 * (may be NULL) = the winner's time per batch step. */
int mdx_tune_batch(mdx_ctx *const *ctxs, int n, int flags, int steps, int32_t *chosen, double *ms_per_step);
int mdx_sync(mdx_ctx *ctx);
/* Implementation switches (no counterpart in the reference): "tiled_spread"
 * 1 (default) accumulates the charge mesh in LDS tiles, 0 adds every weight
 * to the global mesh directly (bit-identical meshes); "p3m_spread" 1 (default)
 * LDS tiles with one global atomic per tile point, 2 owner-computes cell
 * blocks writing the real mesh directly (no global atomics), 0 global
 * atomics only (bit-identical meshes); "p3m_gather" 2 (default) the field
 * box of each cell-column tile staged in LDS ("gather_tile_kb" 8..64 KB of
 * LDS a tile, default 48; "gather_tile_team" 4, 8 (default) or 16 lanes a
 * site; "gather_tile_planar" 1 (default) loads the three planar field
 * components, 0 interleaves them first), 1 row teams from the per-site
 * mesh records of the position refresh, 0 row teams computing the mesh point
 * from the positions (identical forces); "spread_tile_kb" 8..64 KB of LDS a
 * spread tile (default 48: 24, 32 and 16 measured slower); "p3m_tile_cols" 1 (default) or 2
 * cell columns per tile side in the LDS-tiled spread and gather (2 x 2
 * columns cut the halo points but halve the workgroups: measured slower
 * here; identical forces); "p3m_fft" 1 (default) solves the mesh
 * in three hand-written passes (FFTs in LDS with the charge conversion, the
 * field and the k-space energy and virial fused; power-of-two meshes whose
 * planes fit a workgroup, else hipFFT), 0 hipFFT with separate conversion and
 * field kernels (forces equal to FP32 rounding); "mesh_stream" 1 (default) runs the
 * P3M chain of a captured step on a second stream beside the pair and E3B
 * kernels, 0 on one stream (identical trajectories); "graph" 1 (default)
 * runs mdx_run as captured 8-step graphs, 0 eagerly; "build_mode" 0 or 2
 * (both the per-atom team build); "cell_reach_xy" 1..2 and "cell_reach_z"
 * 1..16: cell edges >= rlist / reach per axis (identical lists, finer cells
 * test fewer candidates); "gather_team" 8 or 16 lanes per charged atom in
 * the P3M field gather; "sur_force_lanes" 8, 16, 32 or 64 lanes per molecule
 * in the learned term's force pass; "row_cap" >= 16 sets the neighbor-row capacity after
 * the first build (rows still grow on overflow); "replay_chunk" k >= 1 makes the
 * k-th 64-step chunk of every following mdx_run / mdx_run_batch call replay once
 * from its snapshot as if its rows had overflowed (0: off), for testing the
 * replay path. */
int mdx_option_set(mdx_ctx *ctx, const char *name, int value);

/* Force-path steps at fixed positions: nsteps evaluations of the terms in
 * `flags`, each preceded by the skin check, with a forced neighbor-list
 * rebuild every rebuild_every steps (0: only when the skin criterion fires).
 * Measures the force path in isolation, the part of IntegratorVerlet::run
 * between the half-kicks (IntegratorVerlet.cu:696-729). */
int mdx_force_steps(mdx_ctx *ctx, int nsteps, int flags, int rebuild_every);
/* Kernel timing with HIP events on the context stream (no counterpart in the
 * reference, whose only timer is the wall clock of IntegratorVerlet.cu:687,
 * 759-767).  Enabling clears earlier records.  Kernel ids: 0 list build,
 * 1 pair, 2 mesh spread, 3 mesh gather, 4 E3B pass 1, 5 E3B pass 2,
 * 6 FFTs, 7 field, 8 surrogate descriptors, 9 surrogate network, 10
 * surrogate forces. */
int mdx_profile_enable(mdx_ctx *ctx, int enable);
int mdx_profile_get(mdx_ctx *ctx, int kernel, int *count, double *total_ms);
/* Entries in the current pair list and E3B molecule list (sum of counts). */
int mdx_nlist_totals(mdx_ctx *ctx, int64_t *pair_entries, int64_t *mol_entries);
/* Work of the pair kernel on the current list: row entries it reads (charged
 * partners of charged atoms and of atoms with LJ to charged types, charge-free
 * partners of atoms with LJ to charge-free types) and atoms that read any.
 * No counterpart in the reference; the roofline's algorithmic bytes. */
int mdx_pair_entries_read(mdx_ctx *ctx, int64_t *entries, int64_t *atoms);

#ifdef __cplusplus
}
#endif

#endif /* MDX_H */
