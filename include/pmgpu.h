/*
 * pmgpu.h -- C ABI of the MI355X (gfx950) implementation of ProtoMol's
 * force-evaluation hot path: the periodic cell-list cutoff nonbonded system
 * force and the PME reciprocal/correction force.
 *
 * The entry points replace, one for one, what a ProtoMol SystemForce does for
 * these forces (ProtoMol 2.1 tree, framework/):
 *   - force construction by id and positional parameters
 *       ForceFactory::make -> Force::make -> doMake
 *       (frontend/ForceFactoryDetails.h:60, forces/Force.cpp:15-21,
 *        forces/NonbondedCutoffForce.h:113-129,
 *        forces/NonbondedPMEwaldSystemForce.h:1078-1139)
 *   - evaluation
 *       SystemForce::evaluate(topo, positions, forces, energies)
 *       (forces/SystemForce.h:39), driven by ForceGroup::evaluateSystemForces
 *       (forces/ForceGroup.cpp:43-96)
 *   - Force::uncache (forces/Force.h, NonbondedPMEwaldSystemForce.h:118-121)
 * plus the topology data that GenericTopology holds (atoms, types, LJ table,
 * exclusions, boundary conditions, cell manager; topology/GenericTopology.h,
 * topology/Topology.h) and an on-device MD driver that replaces
 * LeapfrogIntegrator::run / LangevinLeapfrogIntegrator
 * (integrators/LeapfrogIntegrator.cpp:107-121,
 *  integrators/LangevinLeapfrogIntegrator.cpp:58-79).
 *
 * Conventions (identical to ProtoMol): positions/forces are AoS double[3N]
 * (base/Vector3DBlock.h), Angstrom, kcal/mol, amu; charges are pre-scaled by
 * sqrt(332.0636); forces and energies are ACCUMULATED into (+=), the caller
 * zeroes them (StandardIntegrator.cpp:61).  energies[] is the prefix of
 * ScalarStructure (base/ScalarStructure.h:25-45):
 *   [0] COULOMB, [1] LENNARDJONES, [2] BOND, [3] ANGLE, [4] DIHEDRAL,
 *   [5] IMPROPER, [6] OTHER, [7..15] VIRIALXX XY XZ YX YY YZ ZX ZY ZZ.
 * All functions return PM_OK (0) or a PM_ERR_* code; pm_last_error() gives the
 * message, worded as the reference's errMsg/report text where one exists.
 */
#ifndef PMGPU_H
#define PMGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define PM_ABI_VERSION 1

enum {
  PM_OK = 0,
  PM_ERR_ARGUMENT = 1,     /* invalid parameter (reference: make() returns NULL + errMsg) */
  PM_ERR_STATE = 2,        /* call out of order (no topology, no boundary, ...) */
  PM_ERR_DEVICE = 3,       /* HIP runtime failure */
  PM_ERR_CAPACITY = 4,     /* a device buffer bound was exceeded (message says which) */
  PM_ERR_UNSUPPORTED = 5,  /* force id not implemented on the GPU */
  PM_ERR_NUMERIC = 6       /* a position was not finite, or a force term left the
                              64-bit fixed-point range (+-2^23 kcal/mol/A): the
                              forces and energies of the call are not valid */
};

enum { PM_PRECISION_MIXED = 0, PM_PRECISION_DOUBLE = 1 };
/* PM_EVAL_VIRIAL: accumulate the atomic virial (ScalarStructure::virial()).
 * PM_EVAL_MOLVIRIAL: the molecular virial (ScalarStructure::molecularVirial(),
 * OneAtomPair.h:175-183) -- not computed on the GPU: pm_evaluate returns
 * PM_ERR_UNSUPPORTED, so an adapter never hands back a run without it. */
enum { PM_EVAL_VIRIAL = 1, PM_EVAL_MOLVIRIAL = 2 };
enum { PM_INTEGRATOR_LEAPFROG = 0, PM_INTEGRATOR_LANGEVIN = 1 };
enum { PM_ENERGY_COUNT = 16 };
enum {
  PM_T_BIN = 0, PM_T_LIST = 1, PM_T_PAIR = 2, PM_T_PME = 3, PM_T_CORR = 4,
  PM_T_BONDED = 5, PM_T_REDUCE = 6, PM_T_INTEGRATE = 7, PM_TIMING_CATEGORIES = 8
};

typedef struct pm_context pm_context;

/* Version of this ABI (PM_ABI_VERSION). */
int pm_abi_version(void);

/* Last error message of this context (never NULL). */
const char* pm_last_error(const pm_context* ctx);

/* Create a context on HIP device `device`.  No reference counterpart: the
 * reference keeps this state in the Force objects and the topology. */
int pm_create(int device, pm_context** out);
void pm_destroy(pm_context* ctx);

/* Arithmetic mode of every kernel: FP32 pair/reciprocal arithmetic with
 * 64-bit fixed-point force accumulation (MIXED), or FP64 throughout (DOUBLE). */
int pm_set_precision(pm_context* ctx, int precision);

/* Topology arrays as frontend/buildTopology.cpp builds them: scaled charges
 * (:148), atom types, LJ A/B/A14/B14 tables ntypes x ntypes row-major
 * (:643-664, LennardJonesParameterTable), molecule of each atom (:733-905),
 * exclusion pairs and their ExclusionClass 0 NONE / 1 MODIFIED / 2 FULL
 * (:963-1011, ExclusionTable::add), masses, coulombScalingFactor. */
int pm_set_topology(pm_context* ctx, int n, const double* scaled_charge, const int* type,
                    int ntypes, const double* ljA, const double* ljB, const double* ljA14,
                    const double* ljB14, const int* molecule, const double* mass, int nexcl,
                    const int* ex_a, const int* ex_b, const int* ex_class,
                    double coulomb_scaling_factor);

/* Per-atom Lennard-Jones parameters in place of the type table: sigma (A) and
 * epsilon (kcal/mol) of every atom, and their 1-4 values (NULL: the same).
 * Pairs combine them as frontend/buildTopology.cpp:643-664 combines the type
 * parameters: R = sigma 2^(-5/6), A = (Ri+Rj)^12 sqrt(eps_i eps_j),
 * B = 2 (Ri+Rj)^6 sqrt(eps_i eps_j).  sigma == NULL returns to the type
 * table.  Needed for gradients with respect to per-atom sigma and epsilon
 * (pm_param_gradients); the reference's own route is one atom type per atom. */
int pm_set_lj_per_atom(pm_context* ctx, const double* sigma, const double* epsilon, const double* sigma14,
                       const double* epsilon14);

/* Bonded terms of the topology (frontend/buildTopology.cpp fills
 * GenericTopology::bonds and ::angles from PSF/PAR): bond atoms, spring
 * constant (kcal/mol/A^2), rest length (A); angle atoms (a2 is the vertex),
 * force constant (kcal/mol/rad^2), rest angle (rad), Urey-Bradley constant
 * and rest length.  Used by the "Bond" and "Angle" forces
 * (forces/BondSystemForce.h:111-145, forces/AngleSystemForce.h:113-200). */
int pm_set_bonded(pm_context* ctx, int nbonds, const int* bond_a, const int* bond_b,
                  const double* bond_k, const double* bond_r0, int nangles, const int* angle_a,
                  const int* angle_b, const int* angle_c, const double* angle_k,
                  const double* angle_theta0, const double* angle_kub, const double* angle_rub);

/* Dihedrals (kind 0) or impropers (kind 1) of the topology, as
 * frontend/buildTopology.cpp:348-566 fills GenericTopology::dihedrals /
 * ::impropers (Torsion records): torsion t has atoms[4t .. 4t+3] and the
 * terms term_row[t] .. term_row[t+1]-1 (term_row[0] = 0) with force constant
 * (kcal/mol), periodicity and phase shift (radians).  Used by the "Dihedral"
 * and "Improper" forces (forces/DihedralSystemForce.h,
 * forces/ImproperSystemForce.h through MTorsionSystemForce::calcTorsion,
 * forces/MTorsionSystemForce.h:34-153).  Call after pm_set_topology. */
int pm_set_torsions(pm_context* ctx, int kind, int ntorsions, const int* atoms, const int* term_row,
                    const double* k, const int* period, const double* phase);

/* Periodic boundary conditions and cubic cell manager: cellBasisVector1-3,
 * cellOrigin and cellsize (topology/PeriodicBoundaryConditions.cpp:34-77,
 * topology/CubicCellManager.cpp:25-42).  Only orthogonal boxes (the reference
 * PME requires them, NonbondedPMEwaldSystemForce.h:542-543). */
int pm_set_boundary(pm_context* ctx, const double* e1, const double* e2, const double* e3,
                    const double* origin, double cellsize);

/* Create a force from its ProtoMol id and positional parameter values, e.g.
 *   "LennardJones CoulombEwaldReal -algorithm NonbondedCutoff
 *    -switchingFunction C2 -switchingFunction Cutoff"
 *      values = {switchon, cutoff (C2), alpha, cutoff (Cutoff), algorithm cutoff}
 *   "Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline"
 *      values = {nx, ny, nz, cutoff, order, accuracy, alpha}   (alias "CoulombPME")
 *   "Bond", "Angle", "Dihedral", "Improper"   (no values; the topology's bonded terms)
 * Replaces ForceFactory::make + doMake (forces/NonbondedCutoffForce.h:123-129,
 * forces/NonbondedPMEwaldSystemForce.h:1100-1139); the same checks, the same
 * messages.  *handle identifies the force in later calls. */
int pm_force_create(pm_context* ctx, const char* force_id, const double* values, int nvalues,
                    int* handle);

/* Force::uncache: drop cached precomputation (PME alpha/tables, cell plans). */
int pm_force_uncache(pm_context* ctx, int handle);

/* SystemForce::evaluate for a group of forces sharing one cell-list build
 * (ForceGroup::evaluateSystemForces).  Host buffers; positions double[3n],
 * forces double[3n] += , energies double[PM_ENERGY_COUNT] += .
 * flags: PM_EVAL_VIRIAL to accumulate the atomic virial; PM_EVAL_MOLVIRIAL
 * (the molecular virial) returns PM_ERR_UNSUPPORTED; PM_ERR_NUMERIC reports
 * a non-finite position or a force outside the fixed-point range. */
int pm_evaluate(pm_context* ctx, const int* handles, int nhandles, const double* positions,
                double* forces, double* energies, int flags);

/* Cell binning diagnostics of topology/buildCellLists.cpp:11-38 and the
 * enumerator CellListEnumerator_periodicBoundaries.h: cell grid dims[3],
 * cell of every atom, per-cell head/next linked lists in the reference's
 * order, and per first-cell candidate / in-cutoff (r^2 <= pair_cutoff^2, not
 * FULL-excluded) pair counts over the enumerator's cell pairs built from
 * algorithm_cutoff.  Any output pointer may be NULL; head/cand/inrc need
 * dims[0]*dims[1]*dims[2] entries (query with head == NULL first). */
int pm_cell_diagnostics(pm_context* ctx, const double* positions, double algorithm_cutoff,
                        double pair_cutoff, int* dims, int* cell_of_atom, int* head, int* next,
                        long long* cand, long long* inrc);

/* On-device MD driver (replaces LeapfrogIntegrator::run and the Langevin
 * half kick).  dt in fs, temperature in K, gamma in 1/ps, seed for the
 * Langevin noise.  Positions/velocities stay in HBM between calls. */
int pm_md_setup(pm_context* ctx, const int* handles, int nhandles, int integrator, double dt,
                double temperature, double gamma, unsigned long long seed);
int pm_md_set_state(pm_context* ctx, const double* x, const double* v);
/* LeapfrogIntegrator::run(nsteps): half kick + drift, forces, (kick + drift,
 * forces) x (nsteps-1), half kick. */
int pm_md_run(pm_context* ctx, int nsteps);
/* Copy back x, v (double[3n], may be NULL) and the energies of the last force
 * evaluation (double[PM_ENERGY_COUNT] overwritten, may be NULL). */
int pm_md_get_state(pm_context* ctx, double* x, double* v, double* energies);

/* SHAKE and RATTLE bond constraints in pm_md_run, attached as the reference
 * application attaches them (protomol.cpp:553-578): SHAKE
 * (ModifierShake::doExecute, integrators/ModifierShake.cpp:21-103) after
 * every drift, RATTLE (ModifierRattle::doExecute,
 * integrators/ModifierRattle.cpp:20-85) after every step; with RATTLE a
 * leapfrog step is half kick, RATTLE, half kick, drift, SHAKE
 * (LeapfrogIntegrator.cpp:72-84).  The constraints are
 * buildRattleShakeBondConstraintList's (topology/topologyutilities.cpp:782-865):
 * every bond of pm_set_bonded at its rest length, then for every angle whose
 * end atoms both weigh less than 5 the end atoms' distance from the two bonds'
 * rest lengths and the rest angle.  A modifier is on when its epsilon > 0 and
 * max_iter > 0 (the reference's shakeEpsilon / shakeMaxIter, rattleEpsilon /
 * rattleMaxIter; protomol's defaults are 1e-5 and 30); 0 turns it off.  Call
 * after pm_set_bonded; SHAKE's reference positions are the positions at this
 * call, or at pm_md_set_state.  The degrees of freedom are not reduced here
 * (the reference subtracts the constraint count, pm_md_constraint_count, from
 * the temperature's degrees of freedom). */
int pm_md_set_constraints(pm_context* ctx, double shake_eps, int shake_max_iter, double rattle_eps,
                          int rattle_max_iter);
int pm_md_constraint_count(pm_context* ctx, int* count);

/* One SHAKE application outside MD (ModifierShake::doExecute): x and v are
 * corrected in place, x_last holds the positions after the previous
 * application (the caller keeps the corrected x as the next x_last); dt_fs is
 * the timestep in fs.  *sweeps (may be null) = the sweeps run.  pm_rattle:
 * one RATTLE application (ModifierRattle::doExecute) on v at positions x. */
int pm_shake(pm_context* ctx, double* x, double* v, const double* x_last, double dt_fs, double eps, int max_iter,
             int* sweeps);
int pm_rattle(pm_context* ctx, const double* x, double* v, double dt_fs, double eps, int max_iter, int* sweeps);
/* pm_shake without allocating: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as pm_shake, but its work space (device copies of x, v and x_last,
 * the probe's copy and per-sweep error sums, the sweep count) lies in one
 * fixed 16 MiB array per GPU that the library holds until the process ends,
 * so it needs 4 * 24 n + 8 (max_iter + 1) * waves bytes (waves = 4 per 256
 * constraint groups) plus alignment within 16 MiB -- about 160,000 atoms at
 * max_iter 30 -- and returns PM_ERR_ARGUMENT beyond that.  The constraint
 * list must exist (pm_md_set_constraints, or one pm_shake call, made it);
 * otherwise PM_ERR_STATE.  Every call on one GPU uses that GPU's array:
 * calls that overlap in time on one GPU (other threads, other contexts on
 * the same device) race and must be serialised by the caller. */
int pm_shake_noalloc(pm_context* ctx, double* x, double* v, const double* x_last, double dt_fs, double eps,
                     int max_iter, int* sweeps);

/* Choose this context's layout options for the MD steps set up with
 * pm_md_setup by timing them: cluster size 8 or 4 (PM_OPT_CLUSTER_SIZE)
 * crossed with pair-list skins 0, 0.4, 0.8 and 1.2 A (PM_OPT_LIST_SKIN),
 * each run for trial_steps steps after a warm-up from the current state.
 * The state (positions, velocities, Langevin counter) is restored afterwards,
 * so the trajectory does not advance; the fastest options stay set.  Any of
 * the outputs may be NULL: the chosen cluster size, skin (1e-3 A) and the
 * measured ms per step.  No reference counterpart (the reference has one
 * fixed code path). */
int pm_md_autotune(pm_context* ctx, int trial_steps, int* cluster_size, int* list_skin, double* ms_per_step);

/* Kernel statistics of the last evaluation (for roofline reporting):
 * stats[0] pair-list entries (i-cluster x j-atom), stats[1] pairs evaluated
 * (lanes issued), stats[2] clusters, stats[3] padded slots, stats[4] pair-list
 * builds of reusing evaluations since the context was created (PM_OPT_LIST_SKIN),
 * stats[5] prunes of the inner lists without a build (PM_OPT_PRUNE_SKIN).
 * With inner lists, stats[0] and [1] count the inner lists. */
int pm_last_stats(pm_context* ctx, long long* stats, int nstats);

/* Kernel timing with HIP events recorded on the context's stream around the
 * kernel groups selected by mask (bit k = category k):
 *   0 binning/slots (and folding per-slot pair forces into atoms),
 *   1 pair lists, 2 pair kernel, 3 PME reciprocal
 *   (spread, FFTs, influence function, gather), 4 PME correction,
 *   5 bond/angle, 6 energy reduction, 7 position/velocity update.
 * No reference counterpart (the reference's TimerStatistic wraps whole force
 * evaluations, base/TimerStatistic.h:22-32).  pm_set_timing resets the
 * totals; pm_get_timing returns total milliseconds and launch counts. */
int pm_set_timing(pm_context* ctx, unsigned mask);
int pm_get_timing(pm_context* ctx, double* ms, long long* count, int ncategories);

/* Hessian-vector product of the energy of a force group, hv = H v (double[3n],
 * overwritten), at the given positions: the nonbonded cutoff pairs (the
 * derivative of exactly the forces pm_evaluate computes, switching included),
 * the PME reciprocal term (B-spline spread of the directional charge
 * derivative, a second convolution, second-derivative gather) and the PME
 * correction term.  Bond and Angle forces are rejected.  Computed by an
 * adjoint kernel that re-evaluates every pair, storing no Hessian.  The
 * reference assembles a dense Hessian for LJ and Coulomb cutoff forces only
 * (integrators/Hessian.cpp:456-624, forces/ReducedHessLennardJones.cpp,
 * ReducedHessCoulomb.cpp). */
int pm_hessian_vec(pm_context* ctx, const int* handles, int nhandles, const double* positions, const double* v,
                   double* hv);
/* Gradients of the energy of a force group with respect to every atom's scaled
 * charge (dE_dq, double[n]), and, with per-atom LJ parameters set, its sigma and
 * epsilon (dE_dsigma, dE_depsilon; pairs that are not MODIFIED; 0 where
 * epsilon is 0).  Any output may be NULL.  No reference counterpart. */
int pm_param_gradients(pm_context* ctx, const int* handles, int nhandles, const double* positions, double* dE_dq,
                       double* dE_dsigma, double* dE_depsilon);

/* Replace every atom's scaled charge (double[n], the same order and scaling as
 * pm_set_topology), keeping types, exclusions and caches.  The reference
 * changes charges by writing Atom::scaledCharge in the topology
 * (base/Atom.h:79); any later evaluation sees the new values. */
int pm_set_charges(pm_context* ctx, const double* scaled_charge);

/* Lanczos iteration with full reorthogonalisation on the Hessian of a force
 * group (the same terms as pm_hessian_vec), or with mass_weighted on
 * M^-1/2 H M^-1/2, the matrix the reference diagonalises for normal modes
 * (integrators/NormalModeDiagonalize.cpp:204-229, Hessian::evaluate with mass
 * re-weighting; the reference calls LAPACK dsyevr on the dense matrix).  Runs
 * `steps` Hessian-vector products from v0 (double[3n], normalised here); the
 * Krylov basis, Gram-Schmidt coefficients and norms stay in device memory.
 * Outputs: ritz_values[steps] ascending (NaN past an invariant subspace);
 * ritz_vectors[nvec][3n] for the nvec lowest (may be NULL when nvec is 0);
 * residuals[steps] (optional), the bound |beta_m s_mi| on |A y - theta y|;
 * alpha_beta[2 steps] (optional), the tridiagonal diagonal then
 * off-diagonal.  1 <= steps <= min(PM_LANCZOS_MAX_STEPS, 3n). */
#define PM_LANCZOS_MAX_STEPS 256
int pm_lanczos(pm_context* ctx, const int* handles, int nhandles, const double* positions, const double* v0,
               int steps, int mass_weighted, int nvec, double* ritz_values, double* ritz_vectors,
               double* residuals, double* alpha_beta);

/* Layout and kernel options, chosen per context (no reference counterpart:
 * they change speed, never results beyond the mixed-precision tolerance):
 *   PM_OPT_CLUSTER_SIZE  atoms per i-cluster of the pair kernels: 4, 8, or
 *                        0 for the default (8).
 *   PM_OPT_GRAPH         1 (default): pm_md_run replays captured HIP graphs
 *                        of up to 8 steps; 0: every kernel launched directly.
 *                        Steps are captured only while no kernel timing is
 *                        requested (pm_set_timing mask 0).
 *   PM_OPT_LIST_SKIN     pair-list buffer in 1e-3 A: -1 (default) automatic
 *                        (2200 mixed, 1000 FP64), 0 off (cells and lists rebuilt at every
 *                        evaluation, as the reference rebuilds its cell list,
 *                        ForceGroup.cpp:50), 1..3000.  Lists then hold the j
 *                        atoms within the pair cutoff + skin and are rebuilt
 *                        on the device when an atom has moved more than
 *                        skin / 2 since the last build; every evaluation still
 *                        computes exactly the pairs within the cutoff.  Used
 *                        for one pair pass whose algorithm cutoff covers its
 *                        pair cutoff (any box: boxes narrower than 2 (cutoff
 *                        + skin + cell diagonal) list by the exact per-atom
 *                        distance); otherwise lists are rebuilt each time.
 *   PM_OPT_PME_FFT       the y and x transforms of the PME grid: 0 (default)
 *                        automatic (separate passes), 1 separate y and x
 *                        passes over all lines, 2 one kernel per kz plane
 *                        (y and x transforms and the influence function
 *                        with the plane held in LDS; used where a plane of
 *                        the grid fits, separate passes elsewhere: faster
 *                        alone, slower beside the pair kernel).
 *   PM_OPT_PME_PRIORITY  issue priority (s_setprio, 0..3) of the PME spread,
 *                        transform and gather waves over the waves of the
 *                        pair kernels that run beside them on the other
 *                        stream: -1 (default) automatic (2), 0 equal share.
 *   PM_OPT_PRUNE_SKIN    inner pair lists, in 1e-3 A: with list reuse, the
 *                        pair kernel prunes the lists to the j atoms within
 *                        the pair cutoff + this skin of some atom of the
 *                        cluster, whenever an atom has moved more than half
 *                        of it since the last prune (and on every build), and
 *                        walks the pruned lists in between; every evaluation
 *                        still computes exactly the pairs within the cutoff.
 *                        -1 (default) automatic, 0 off, 1..3000; used when
 *                        below the list skin.
 *   PM_OPT_LIST_TEST     the pair-list distance test: -1 (default) automatic
 *                        (1), 0 the distance from the cluster's bounding box
 *                        in boxes wider than 2 (cutoff + skin + cell
 *                        diagonal) and the exact per-atom distance in smaller
 *                        ones, 1 the exact per-atom distance in every box
 *                        (fewer entries, a longer build).
 *   PM_OPT_PME_SPREAD    the PME charge spread: -1 (default) automatic (1),
 *                        1 fused with the z transform (a block owns whole z
 *                        lines and adds the stencil points of the atoms of
 *                        the cell columns that reach them, no global
 *                        atomics; used where a block's columns and lines fit
 *                        its LDS), 0 by bricks of cells into LDS tiles added
 *                        to the grid with atomics, then the z transform.
 *                        Both give the same grid bit for bit.
 *   PM_OPT_PAIR_LANES    the pair kernel's lane layout: -1 (default)
 *                        automatic (1), 1 one lane per list entry (the
 *                        cluster's atoms in scalar registers, the entry's j
 *                        force whole in its lane), 0 a row of one lane per
 *                        cluster atom for each entry (j forces summed over
 *                        the row).  The lane-per-entry kernel needs the LJ
 *                        type table plus 8 KB of LDS per workgroup: FP64
 *                        contexts with over 41 atom types use rows.
 * pm_get_option returns the value in effect (after the automatic choice). */
enum { PM_OPT_CLUSTER_SIZE = 0, PM_OPT_GRAPH = 1, PM_OPT_LIST_SKIN = 2, PM_OPT_PME_FFT = 3, PM_OPT_PME_PRIORITY = 4,
       PM_OPT_PRUNE_SKIN = 5, PM_OPT_LIST_TEST = 6, PM_OPT_PME_SPREAD = 7, PM_OPT_PAIR_LANES = 8 };
int pm_set_option(pm_context* ctx, int option, int value);
int pm_get_option(pm_context* ctx, int option, int* value);

/* Wait for all work queued by this context.  Like pm_md_get_state, it
 * reports what the kernels of the queued work flagged (PM_ERR_CAPACITY,
 * PM_ERR_NUMERIC): pm_md_run queues its steps and returns at once. */
int pm_synchronize(pm_context* ctx);

/* Batches of independent systems (replicas), each in its own context set up
 * for Langevin dynamics (pm_md_setup with PM_INTEGRATOR_LANGEVIN), advanced
 * together: their steps are captured into one HIP graph (every replica on its
 * own streams) and replayed.  Replica exchange follows
 * applications/rem-app/rem.cpp: neighbours (k, k+1) of the temperature ladder
 * with k of the attempt's parity (getNextAddress :949-957) are tested with
 * delta = (1/(kB Tk) - 1/(kB Tk+1)) (E(k+1) - E(k)) (calcDelta :941-946) and
 * accepted when delta <= 0 or u <= exp(-delta) (canDoEvenSwitch :841-880);
 * an accepted pair swaps bath temperatures, and the velocities are scaled by
 * sqrt(T_new / T_old) or left to the caller (pm_batch_set_exchange_velocities).
 * E is the potential energy of the last step.  No
 * reference counterpart for the batch itself: the reference runs one replica
 * per MPI process. */
typedef struct pm_batch pm_batch;
int pm_batch_create(pm_context* const* contexts, int nreplicas, pm_batch** out);
void pm_batch_destroy(pm_batch* batch);
const char* pm_batch_last_error(const pm_batch* batch);
/* nsteps Langevin steps of every replica (energies of the last one). */
int pm_batch_run(pm_batch* batch, int nsteps);
/* One exchange attempt; u from (seed, attempt, k).  accepted[k] and deltas[k]
 * (nreplicas entries each, may be NULL) are written for the tested pairs. */
int pm_batch_exchange(pm_batch* batch, int attempt, unsigned long long seed, int* accepted, double* deltas);
/* What an accepted exchange does to the two replicas' velocities.
 * PM_EXCHANGE_RESCALE (default): scaled by sqrt(T_new / T_old), doSwitch's
 * rescaling branch (rem.cpp:930-935).  PM_EXCHANGE_KEEP: left as they are,
 * for the caller to redraw at the new temperature and remove the momentum,
 * which is what doSwitch runs (redrawVelocities = 1, rem.cpp:912, :923-928:
 * randomVelocity, removeLinearMomentum, removeAngularMomentum); the library
 * has no atom names to tell the solute from water, which momentum removal
 * needs, so the host driver does it (pmrun's REM mode does). */
enum { PM_EXCHANGE_RESCALE = 0, PM_EXCHANGE_KEEP = 1 };
int pm_batch_set_exchange_velocities(pm_batch* batch, int mode);
/* ladder[k]: the replica now at ladder temperature temps[k]. */
int pm_batch_ladder(pm_batch* batch, int* ladder, double* temps);

/* How pm_batch_run advances the replicas.  PM_BATCH_AUTO (default): every
 * kernel of an MD step runs as one launch over all replicas (each block finds
 * its replica from a table of the replicas' arguments; the kernels are the
 * single-context ones), replayed from a captured graph, when every replica
 * records the same kernel sequence (same forces, precision and cluster size,
 * no PME); otherwise, and with PM_BATCH_CHAINS, one chain of captured steps
 * per replica on its own stream.  Both give the same trajectories.
 * pm_batch_info: batched = 1 when the last pm_batch_run used batched launches;
 * why (may be NULL) receives the reason when it did not. */
enum { PM_BATCH_AUTO = 0, PM_BATCH_CHAINS = 1 };
int pm_batch_set_mode(pm_batch* batch, int mode);
int pm_batch_info(pm_batch* batch, int* batched, char* why, int why_len);

#ifdef __cplusplus
}
#endif
#endif
