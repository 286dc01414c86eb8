/*
 * pmoracle.h -- FP64 CPU restatement of ProtoMol's force-evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in protomol_amd/.  It is linked or loaded only by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py; the product
 * path never calls it.
 *
 * Every routine follows the loop order and arithmetic of the reference file
 * cited beside it (paths relative to the ProtoMol 2.1 tree, framework/).
 * Units: Angstrom, kcal/mol, amu; charges pre-scaled by sqrt(332.0636)
 * (frontend/buildTopology.cpp:148).
 *
 * Pinning (what checks each function, see tests/test_oracle*.py):
 *   orc_cells / orc_delta_list / orc_cutoff   -> compiled reference
 *       (oracle/_ref), committed golden fixtures, brute-force O(N^2) pair sets
 *   orc_pme (reciprocal, correction, self, charged) -> compiled reference,
 *       golden fixtures, direct Ewald summation (numpy), naive DFT for the FFT
 *   orc_pme_real                               -> compiled reference
 *   orc_bspline                                -> partition of unity, closed
 *       form of the cubic B-spline, compiled reference (through orc_pme)
 *   orc_leapfrog / orc_langevin_half_kick      -> closed form on one particle
 *       (harmonic oscillator) and the reference formulas
 *   orc_hessian_pairs                          -> central finite differences
 *       of orc_cutoff forces; compiled reference Hessian (Hessian.cpp)
 *   orc_rem_delta                              -> closed form
 *   orc_constraint_list                        -> law of cosines on the TIP3
 *       geometry, bond list order (tests/test_oracle_constraints.py)
 *   orc_shake / orc_rattle                     -> the converged SHAKE and
 *       RATTLE solutions of one molecule solved independently (Newton on the
 *       three multipliers; the 3x3 linear system), error bound after return
 */
#ifndef PMORACLE_H
#define PMORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ExclusionClass, topology/ExclusionTable.h:20-25 */
enum { ORC_EXCL_NONE = 0, ORC_EXCL_MODIFIED = 1, ORC_EXCL_FULL = 2 };

/* pair potentials (forces/LennardJonesForce.h, CoulombForce.h,
 * CoulombEwaldRealForce.h) */
enum { ORC_POT_LJ = 0, ORC_POT_COULOMB = 1, ORC_POT_EWALDREAL = 2 };

/* switching functions (topology/CutoffSwitchingFunction.h,
 * C1SwitchingFunction.h, C2SwitchingFunction.h) */
enum { ORC_SWF_CUTOFF = 0, ORC_SWF_C1 = 1, ORC_SWF_C2 = 2 };

/* energies[] layout: the ScalarStructure prefix COULOMB, LENNARDJONES, BOND,
 * ANGLE, DIHEDRAL, IMPROPER, OTHER, then the nine atomic virial entries
 * (base/ScalarStructure.h:25-45) */
enum {
  ORC_E_COULOMB = 0, ORC_E_LJ = 1, ORC_E_BOND = 2, ORC_E_ANGLE = 3, ORC_E_DIHEDRAL = 4, ORC_E_IMPROPER = 5,
  ORC_E_VXX = 7, ORC_E_VXY, ORC_E_VXZ, ORC_E_VYX, ORC_E_VYY, ORC_E_VYZ,
  ORC_E_VZX, ORC_E_VZY, ORC_E_VZZ,
  ORC_E_COUNT
};

/* one OneAtomPair term: potential + switching function + parameters */
typedef struct {
  int pot;          /* ORC_POT_* */
  int swf;          /* ORC_SWF_* */
  double alpha;     /* Ewald splitting (EwaldReal only) */
  double switchon;  /* C2 only */
  double cutoff;    /* switching-function cutoff */
} orc_term;

/* Cell binning (topology/buildCellLists.cpp:11-38).  Outputs:
 *   dims[3]   cell-grid dimensions
 *   realcell[3] real cell size
 *   cell[n]   linear cell index (x*ny+y)*nz+z of every atom
 *   head[ncells] first atom of every cell (-1 if empty), next[n] list link.
 * Returns the number of cells, or -1 if head is too small (ncap). */
int orc_cells(int n, const double* pos, const double* box, const double* origin,
              double cellsize, int* dims, double* realcell, int* cell,
              int* head, int* next, int ncap);

/* Deduplicated, sorted delta list of the periodic enumerator
 * (topology/CellListEnumerator_periodicBoundaries.h:37-108).  out holds
 * 3*cap ints.  Returns the count (or -count if cap was too small). */
int orc_delta_list(const double* box, const double* origin, double cellsize,
                   double cutoff, int* out, int cap);

/* Nonbonded cutoff system force (forces/NonbondedCutoffSystemForce.h:64-74,
 * NonbondedCutoffForce.h:86-109, OneAtomPair.h:127-188,
 * OneAtomPairTwo.h:131-199).  nterms is 1 (OneAtomPair) or 2
 * (OneAtomPairTwo).  forces (3n) and energies (ORC_E_COUNT) are ACCUMULATED
 * into.  Optional per-cell statistics (ncells entries each, may be NULL):
 *   cand[c]  candidate pairs visited with c as the first cell of the pair
 *   inrc[c]  of those, pairs with r^2 <= rc^2 that are not fully excluded
 * Returns 0 on success, negative on error. */
int orc_cutoff(int n, const double* pos, const double* q, const int* type,
               int ntypes, const double* ljA, const double* ljB,
               const double* ljA14, const double* ljB14, const int* molecule,
               int nexcl, const int* ex_a, const int* ex_b, const int* ex_cls,
               double csf, const double* box, const double* origin,
               double cellsize, int nterms, const orc_term* terms,
               double algo_cutoff, int want_virial, double* forces,
               double* energies, long long* cand, long long* inrc);

/* B-spline weights of order p at fractional offset w (base/BSpline.cpp:64-117). */
void orc_bspline(int order, double w, double* theta, double* dtheta);

/* Ewald splitting parameter by bisection
 * (forces/NonbondedPMEwaldSystemForce.h:564-586). */
double orc_pme_alpha(double rc, double accuracy);

/* PME reciprocal + correction + self + charged-system terms
 * (forces/NonbondedPMEwaldSystemForce.h:271-349,499-648,879-981;
 *  base/Grid.h:138-228,276-389,557-665).  flags: bit0 real, bit1 reciprocal,
 * bit2 correction.  alpha<=0 selects the bisection.  Accumulates into
 * forces/energies.  parts (may be NULL) receives
 * {real, reciprocal, intra(correction), pointSelf, chargedSystem, alpha}.
 * qgrid (may be NULL, nx*ny*nz doubles) receives the spread charge grid. */
int orc_pme(int n, const double* pos, const double* q, const int* molecule,
            int nexcl, const int* ex_a, const int* ex_b, const int* ex_cls,
            double csf, const double* box, const double* origin,
            double cellsize, int nx, int ny, int nz, int order, double rc,
            double accuracy, double alpha, int flags, int want_virial,
            double* forces, double* energies, double* parts, double* qgrid);

/* In-place unnormalised 3D complex FFT, data[nx][ny][nz] interleaved re,im.
 * sign -1: exp(-i...) ("backward" in base/FFTComplex.cpp:79-82), +1: exp(+i).
 * Factors 2,3,4,5 natively, other primes by direct DFT. */
void orc_fft3d(double* data, int nx, int ny, int nz, int sign);

/* Leapfrog kick-drift over nsteps with the force callback replaced by a
 * caller-provided force array sequence is not possible in C ABI; instead
 * orc_leapfrog_kickdrift does one velocity+position update
 * (integrators/LeapfrogIntegrator.cpp:46-104): v += F*h*fac/m; x += v*h with
 * h = dt*INV_TIMEFACTOR; fac = 0.5 for the half kick, 1 for the full kick. */
void orc_leapfrog_kickdrift(int n, double dt, double fac, const double* mass,
                            const double* f, double* v, double* x);
/* Half kick only (LeapfrogIntegrator::doHalfKick via STSIntegrator). */
void orc_half_kick(int n, double dt, const double* mass, const double* f,
                   double* v);

/* Langevin BBK-style half kick (integrators/LangevinLeapfrogIntegrator.cpp:58-79)
 * with caller-provided Gaussian numbers xi[3n]; gamma in 1/ps. */
void orc_langevin_half_kick(int n, double dt, double temperature, double gamma,
                            const double* mass, const double* f,
                            const double* xi, double* v);

/* Kinetic energy (topology/topologyutilities.cpp:182-191). */
double orc_kinetic_energy(int n, const double* mass, const double* v);

/* Dense pair Hessian-vector product of the cutoff force set, all pairs under
 * the minimal image that are not fully excluded
 * (integrators/Hessian.cpp:456-624, forces/ReducedHessLennardJones.cpp:7-50,
 * ReducedHessCoulomb.cpp:8-42, C1/C2SwitchingFunction.cpp hessian()).
 * hv[3n] = H v. */
int orc_hessian_vec(int n, const double* pos, const double* q, const int* type,
                    int ntypes, const double* ljA, const double* ljB,
                    const double* ljA14, const double* ljB14,
                    const int* molecule, int nexcl, const int* ex_a,
                    const int* ex_b, const int* ex_cls, double csf,
                    const double* box, const double* origin, int nterms,
                    const orc_term* terms, const double* v, double* hv);

/* Bond and angle system forces (forces/BondSystemForce.h:111-145,
 * forces/AngleSystemForce.h:113-200).  which: bit0 bonds, bit1 angles. */
void orc_bonded(int n, const double* pos, const double* box, const double* origin, int nb,
                const int* ba, const int* bb, const double* bk, const double* br0, int na,
                const int* aa, const int* ab, const int* ac, const double* ak, const double* at0,
                const double* akub, const double* arub, int which, int want_virial,
                double* forces, double* energies);

/* Dihedral (improper = 0) or improper (improper = 1) system force
 * (forces/DihedralSystemForce.h, forces/ImproperSystemForce.h through
 * MTorsionSystemForce::calcTorsion, forces/MTorsionSystemForce.h:34-153):
 * torsion t has atoms[4t..4t+3] and terms term_row[t] .. term_row[t+1]-1
 * (force constant, periodicity, phase in radians); periodicity > 0 gives
 * k (1 + cos(n phi + phase)), periodicity 0 the harmonic k (phi - phase)^2
 * with the difference wrapped into [-pi, pi].  Forces and energies ACCUMULATE. */
void orc_torsions(int n, const double* pos, const double* box, const double* origin, int nt, const int* atoms,
                  const int* term_row, const double* k, const int* period, const double* phase, int improper,
                  int want_virial, double* forces, double* energies);

/* Replica-exchange acceptance exponent (applications/rem-app/rem.cpp:941-946). */
double orc_rem_delta(double ti, double tj, double ei, double ej);

/* buildRattleShakeBondConstraintList (topology/topologyutilities.cpp:782-865):
 * every bond at its rest length, then for every angle whose end atoms both
 * weigh less than 5 the distance sqrt(b^2 + c^2 - 2 b c cos(theta0)) of the
 * first pair of bonds joining them to the vertex.  c1, c2, cr0 hold room for
 * nb + na entries; returns the constraint count. */
int orc_constraint_list(int n, const double* mass, int nb, const int* ba, const int* bb, const double* br0,
                        int na, const int* aa, const int* ab, const int* ac, const double* at0, int* c1, int* c2,
                        double* cr0);
/* ModifierShake::doExecute (integrators/ModifierShake.cpp:21-103) with
 * ModifierMetaShake::calcError; x, v corrected, x_last set to x; returns the
 * sweeps run.  dt_fs: the timestep in fs. */
int orc_shake(int m, const int* c1, const int* c2, const double* cr0, const double* mass, double dt_fs,
              double eps, int max_iter, double* x, double* v, double* x_last);
/* ModifierRattle::doExecute (integrators/ModifierRattle.cpp:20-85) with
 * ModifierMetaRattle::calcError; v corrected; returns the sweeps run. */
int orc_rattle(int m, const int* c1, const int* c2, const double* mass, double dt_fs, double eps, int max_iter,
               const double* x, double* v);

/* OpenMP-parallel nonbonded cutoff force over cell packages with per-thread
 * force buffers (the multi-core CPU baseline; same pairs as orc_cutoff, the
 * summation order differs).  Returns the thread count used. */
int orc_cutoff_omp(int nthreads, int n, const double* pos, const double* q,
                   const int* type, int ntypes, const double* ljA,
                   const double* ljB, const double* ljA14, const double* ljB14,
                   const int* molecule, int nexcl, const int* ex_a,
                   const int* ex_b, const int* ex_cls, double csf,
                   const double* box, const double* origin, double cellsize,
                   int nterms, const orc_term* terms, double algo_cutoff,
                   double* forces, double* energies);

#ifdef __cplusplus
}
#endif
#endif
