// pm_internal.h -- structures shared by the host side and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Fixed-point scales.  Forces: 2^40 units per kcal/mol/A in int64 (range
// +-2^23 kcal/mol/A, resolution 9.1e-13).  PME charge grid: 2^40 units per
// scaled charge (range +-2^23, resolution 9e-13).
#define PM_FORCE_FIX 1099511627776.0
#define PM_FORCE_UNFIX (1.0 / 1099511627776.0)
#define PM_GRID_FIX 1099511627776.0
#define PM_GRID_UNFIX (1.0 / 1099511627776.0)

// ExclusionClass, topology/ExclusionTable.h:20-25
#define PM_EXCL_NONE 0
#define PM_EXCL_MODIFIED 1
#define PM_EXCL_FULL 2

// Pair-kernel potential variants.  LJ switching: 0 none (no LJ term),
// 1 C1, 2 C2, 3 Cutoff.  Coulomb kind: 0 none, 1 Coulomb, 2 CoulombEwaldReal.
// Coulomb switching: 1 C1, 2 C2, 3 Cutoff.
#define PM_SWF_NONE 0
#define PM_SWF_C1 1
#define PM_SWF_C2 2
#define PM_SWF_CUTOFF 3
#define PM_COUL_NONE 0
#define PM_COUL_PLAIN 1
#define PM_COUL_EWALD 2

// Atoms per i-cluster: 8 or 16 (chosen per box from the cell occupancy);
// a wave evaluates 64 / M list entries against the cluster per iteration
// (lane = jslot * M + ii).
#define PM_MMAX 16

// Bounds that keep degenerate inputs from turning into unbounded loops.
#define PM_SORT_MAX 512
#define PM_CAND_FACTOR 16
#define PM_DIAG_MAX 4096

// Layout of an energy record: the prefix of ScalarStructure
// (base/ScalarStructure.h:25-45) up to the atomic virial.
#define PM_E_COUL 0
#define PM_E_LJ 1
#define PM_E_BOND 2
#define PM_E_ANGLE 3
#define PM_E_DIHEDRAL 4
#define PM_E_IMPROPER 5
#define PM_E_VIR 7
#define PM_NE 16
// LDS bytes per wave of the lane-per-entry pair kernel (k_pair_jl), after
// the block's LJ type table: small, so that the kernel leaves most of a CU's
// LDS to the PME kernels that run beside it on the other stream
#define PM_JL_LDS 2048

// Error flags the kernels raise in the context's flag word (atomicOr); the
// host reads and clears the word after each evaluation or MD run and turns a
// raised flag into an error return.
#define PM_FLAG_CAPACITY 1      // a pair list exceeded its capacity
#define PM_FLAG_NONFINITE 2     // a position is not finite (or beyond 2^31 box lengths)
#define PM_FLAG_FORCE_RANGE 4   // a force term outside the fixed-point range +-2^23 kcal/mol/A
#define PM_FORCE_LIMIT 8388608.0

__device__ inline void pm_raise(int* err, int bit) {
  if (err) atomicOr(err, bit);
}

// 64-bit fixed point of one force component; a value beyond the range (or
// not finite) raises PM_FLAG_FORCE_RANGE and contributes 0.
__device__ inline long long pm_force_fix(double v, int* err) {
  if (!(__builtin_fabs(v) < PM_FORCE_LIMIT)) {
    pm_raise(err, PM_FLAG_FORCE_RANGE);
    return 0;
  }
  return __double2ll_rn(v * PM_FORCE_FIX);
}

// Box description used by binning, the pair kernels and PME.
struct PmBox {
  double L[3];        // edge lengths (e1.x, e2.y, e3.z)
  double er[3];       // myE1r.x, myE2r.y, myE3r.z (PeriodicBoundaryConditions.cpp:59-64)
  double origin[3];
  double h[3];        // myH = max - min
  double h2[3];       // myH2 = myH * 0.5
  double d;           // myD = min((L/2)^2)
  double delta[3];    // origin - min (buildCellLists.cpp:15)
  double rreal[3];    // 1 / real cell size (CubicCellManager.cpp:33-35)
  int n[3];           // cell grid dimensions
  double to_fix[3];   // 2^32 / L
  double fix_to_A[3]; // L / 2^32
  double cell0_fix[3];  // fixed point of the cell grid's lower corner (min), in [0, 2^32)
  double real_fix[3];   // real cell size in fixed-point units
  int* err;             // the context's error flag word (PM_FLAG_*), device memory
};

// Switched pair potentials of one (possibly fused) nonbonded force.
struct PmPairParams {
  int lj_swf, coul_kind, coul_swf;
  int modify;          // apply switching (any term's MODIFY flag)
  double cut2;         // max of the terms' switching cutoffs squared
  // LJ switching
  double lj_cut2, lj_on2, lj_sw1, lj_sw2, lj_sw3, lj_c15_1, lj_c05_3, lj_c15_3;
  // Coulomb switching
  double c_cut2, c_on2, c_sw1, c_sw2, c_sw3, c_c15_1, c_c05_3, c_c15_3;
  double alpha, two_alpha_pi;
  double csf;
};

struct PmNbLaunch {
  // slots
  const int4* slot_pos;      // X, Y, Z fixed point, w = charge (float bits)
  const longlong4* slot_pos64; // 64-bit fixed point + charge (double bits), FP64 mode
  const int* slot_meta;      // type | orig << 8
  const int* cluster_cell;
  const int* counts;         // [0] nslot, [1] nclusters
  int ncl_max;
  int dummy_slot;
  // list reuse: the list kernel runs only when *rebuild != 0 (null: always),
  // listing j atoms within the pair cutoff + skin of the cluster's box
  const int* rebuild;
  double skin;
  // cells
  const int* cell_slot;
  const int* cell_count;
  const int4* stencil;       // cell columns (x, y offsets) of the pair lists
  int nstencil;
  // lists
  int* jlist;
  int* jcount;
  int lcap;
  int* overflow;
  int bbox_filter;
  int max_delta;             // largest |a-b| over exclusion pairs
  // exclusions: resolved classes of both atoms of every pair, partner | class << 24
  const int* xs_row;
  const int* xs_part;
  int xs_tcap;               // LDS entries per cluster: M * largest partner count
  unsigned int* jmask;       // per list entry (flagged entries): FULL bit ii, MODIFIED bit 16 + ii
  int m;                     // atoms per cluster (4 or 8)
  int jlanes;                // pair kernel: 1 one lane per list entry (k_pair_jl), 0 rows of m lanes (k_pair)
  // topology
  const int* molecule;
  const int* ex_row;
  const int* ex_partner;
  const int* ex_cls;
  int ntypes1;               // types + 1 dummy
  const float4* lj_f;
  const double4* lj_d;
  // per-atom LJ parameters (when set, instead of the type table): R = Rmin/2,
  // e = sqrt(eps), R14, e14 per atom; A = (Ri+Rj)^12 ei ej, B = 2 (Ri+Rj)^6 ei ej
  const float4* atom_lj_f;
  const double4* atom_lj_d;
  // outputs
  long long* force_fix;      // [3n] fixed point, accumulated (+=) for every real atom
  long long* force_slot;     // [3 * slots] fixed point per slot (pair kernel), folded into force_fix
  double* epart;             // [ncl_max * PM_NE]
  int want_virial;
  int want_energy;           // write the per-cluster energy records
  // inner (pruned) lists, with list reuse: rebuild[1] is the prune flag, set
  // when an atom has moved more than half the inner skin since the last
  // prune.  On an evaluation with either flag set k_prune writes, per
  // cluster, the entries of the full list with some i atom within
  // sqrt(cut2_prune) (the pair cutoff + the inner skin) into ilist / imask /
  // icount; the pair kernel walks only those.
  int inner;
  float cut2_prune;
  int* ilist;
  unsigned int* imask;
  int* icount;
};

// ---------------------------------------------------------------- launchers
// Binning (reference-exact) and slot layout.
// `rebuild` (device flag, may be null = always): the kernels return at once
// when it is 0, keeping the layout of the last build.
// The binning counts atoms per cell into bin_count (zero before every
// binning); the scan copies them into the layout's cell_count and clears
// bin_count (also on steps that keep the layout).
void pm_launch_bin(hipStream_t s, const PmBox& box, int n, const double* pos, int* cell_of,
                   int* bin_count, int* rank, float* key, const int* rebuild);
// Slots: the atoms of a cell column (the nz cells that share x and y) in
// consecutive slots, cell after cell, the column padded to a multiple of m
// after its last cell (nz <= 8192).  aux: 2 ints per chunk of whole columns
// of at most 8192 cells (at most 2 (ncells / 4096 + 2) in all); with more
// than two chunks (and not recording) the scan runs over many blocks,
// otherwise in one
void pm_launch_scan_cells(hipStream_t s, int ncells, int nz, int* bin_count, int* cell_count, int* cell_start,
                          int* cell_slot, int* counts, int* cluster_cell, int ncl_max, int m, const int* rebuild,
                          int* aux);
void pm_launch_scatter(hipStream_t s, int n, const int* cell_of, const int* rank, const int* cell_start,
                       int* cell_atoms, const int* rebuild);
// x_prune (may be null): the positions of the last inner-list prune, written
// on builds and, on kept layouts, when the prune flag rebuild[1] is set
void pm_launch_fill_slots(hipStream_t s, const PmBox& box, int ncells, const double* pos,
                          const double* q, const int* type, const int* cell_count,
                          const int* cell_start, const int* cell_slot, int* cell_atoms,
                          const float* key, int4* slot_pos, longlong4* slot_pos64, int* slot_meta,
                          int* atom_of_slot, int dummy_type, int dummy_slot, int m, const int* rebuild,
                          int4* x_build, int4* x_prune);
// list reuse: flag[0] = 1 when an atom moved more than sqrt(lim2) A from its
// position at the last build (x_build, 32-bit fixed point, w = 1 when valid);
// with x_prune, flag[1] = 1 when one moved more than sqrt(lim2_in) A from its
// position at the last prune of the inner lists
void pm_launch_check_disp(hipStream_t s, const PmBox& box, int n, const double* pos, const int4* x_build,
                          float lim2, int* flag, const int4* x_prune = nullptr, float lim2_in = 0.f);
void pm_launch_set_flag(hipStream_t s, int* flag, int v);
void pm_launch_invalidate_build(hipStream_t s, int n, int4* x_build);
// force_fix[atom] += force_slot[slot] for every real slot (then the slot
// cleared when `clear`); with `flag` (list reuse) also the end of the
// evaluation: builds[0] += (flag[0] != 0), builds[1] += (only flag[1] != 0:
// an inner-list prune without a build), flag[0] = flag[1] = 0
void pm_launch_fold_slots(hipStream_t s, int nslot_max, const int* counts, const int* atom_of_slot,
                          long long* force_slot, long long* force_fix, int* flag,
                          unsigned long long* builds, int clear);
// entries and issued lanes of the lists whose lengths are `jcount`
void pm_launch_list_stats(hipStream_t s, const int* jcount, const int* counts, unsigned long long* out, int m,
                          int jlanes);
void pm_launch_list(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p);
// inner lists from the full ones (a.inner set; returns at once unless a flag is set)
void pm_launch_prune(hipStream_t s, const PmBox& box, const PmNbLaunch& a);
// Adjoint pair kernel (clusters of 8): mode 1 H v, mode 2 (dE/dq, dE/dR, dE/de)
// per slot into `out` (64-bit fixed point, 3 per slot, accumulated).
void pm_launch_adjoint(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p, int precision,
                       int mode, const double4* slot_v, long long* out);
void pm_launch_slot_vec(hipStream_t s, const int* counts, const int* atom_of_slot, const double* v, double4* sv,
                        int nslot_max);
void pm_launch_pair(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p,
                    int precision);
void pm_launch_diag(hipStream_t s, const PmBox& box, int n, const double* pos, const int* cell_of,
                    const int* cell_start, const int* cell_atoms, const int4* half_stencil,
                    int nhalf, const int* molecule, const int* ex_row, const int* ex_partner,
                    const int* ex_cls, double pair_cut2, int ncells, const int* cell_count,
                    long long* cand, long long* inrc);

// PME
struct PmPmeLaunch {
  int n;
  int nx, ny, nz, nzc;       // nzc = nz/2 + 1
  int order;
  double hr[3];              // n / L per axis
  const double* q;           // scaled charges (atom order)
  const int4* slot_pos;
  const int* atom_of_slot;
  const int* counts;
  int nslot_max;
  unsigned long long* grid_fix;  // [nx*ny*nz] int64 charge grid (cleared by the z pass)
  void* grid_r;              // real grid (float or double) [nx*ny*nz]
  void* grid_c;              // complex half grid, nx*ny*nzc points (float2 / double2):
                             // [x][y][kz], or [kz][x][y] when plane is set
  int plane;                 // 1: y and x transforms and the influence function in one
                             // kernel per kz plane (the plane fits in LDS)
  const double* bmod;        // [nx + ny + nz] B(m) = 1/|b(m)|^2 (Grid::dftmod)
  const double* expf_tab;    // [nx + ny + nzc] exp(-pi^2 (m/L)^2 / alpha^2)
  double inv_pi_v;           // 1/(pi V)
  double fac;                // pi^2/alpha^2
  double Lr[3];              // 1/L
  const double* tw;          // twiddles exp(-2 pi i m/n), interleaved re, im
  long long* force_fix;
  double* epart;             // [pm_pme_xpass_blocks() * PM_NE]
  int want_virial;
  int want_energy;           // 0: no energy records (then neither is the virial wanted)
  // theta(m) of every point of the line-major half grid ([x][y][kz], float
  // or double as the grid), the factor the x pass applies when it writes no
  // energy records
  const void* theta;
  int* err;                  // error flag word (PM_FLAG_FORCE_RANGE)
  int prio;                  // wave priority (s_setprio) of the PME kernels: 0 .. 3
  // spread by cell bricks staged in LDS tiles
  const int* cell_slot;
  const int* cell_count;
  int ncell[3];              // cell grid
  int brick[3];              // cells per brick
  int nbrick[3];             // bricks per axis
  int tile[3];               // LDS tile extent (grid points) per axis
  double cell_lo_u[3];       // grid coordinate of cell 0's lower face: frac((origin - L/2)/L) * n
  double cell_du[3];         // grid points per cell
  // spread fused with the z forward transform (k_spread_zfwd): the z lines
  // a block owns along y (4, 2 or 1), 0 when off;
  // drift_u: how far (grid points per axis) an atom may lie outside the cell
  // it was binned in at the last layout build
  int spread_zfwd;
  double drift_u[3];
};

#define PM_FFT_MAXST 10
struct PmFftPlan {
  int n;
  int nst;
  int rad[PM_FFT_MAXST];
  int tw;  // offset of this length's twiddles in the table (complex entries)
};
struct PmAtomTail;
// spread, transforms and gather; tail (may be null) adds the correction and
// bonded forces of each atom in the gather kernel's thread for that atom
bool pm_launch_pme_plans(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans,
                         const longlong4* pos64, int precision, const PmAtomTail* tail);
// the lines along y (4, 2 or 1) of a k_spread_zfwd block whose 4 x ly z
// lines fit 64 KB of LDS, 0 when none does or its cell columns exceed its
// LDS lists (needs nz, order, ncell, cell_du, drift_u)
int pm_pme_spread_zfwd_fits(const PmPmeLaunch& a, int precision);
// a.theta (written through a const_cast) from bmod, expf_tab, inv_pi_v and Lr
void pm_launch_pme_theta(hipStream_t s, const PmPmeLaunch& a, int precision);
// energy records written by the reciprocal term (one per x-pass block, or
// one per kz plane)
int pm_pme_xpass_blocks(int nx, int ny, int nz, int precision, int plane);
// 1 when a kz plane of an nx x ny grid fits the plane kernel's LDS
int pm_pme_plane_fits(int nx, int ny, int precision, int lds_max);
// PME adjoint: mode 0 out[n] += dE/dq (reciprocal potential), mode 1
// out[3n] += (H v) of the reciprocal energy; slot_v = v in slot order
bool pm_launch_pme_adjoint(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans, const longlong4* pos64,
                           int precision, int mode, const double4* slot_v, double* out);

// PME real-space term with lattice images, for boxes no wider than twice the
// cutoff: lat = 3 * nlat lattice vectors ((0,0,0) first, nlat <= 1024), one
// energy record per block (epart may be null: no records)
void pm_launch_ewald_lattice(hipStream_t s, const PmBox& box, int n, const double* pos, const double* q,
                             const int* xs_row, const int* xs_part, const double* lat, int nlat, double rc2,
                             double alpha, double csf, int want_virial, long long* force_fix, double* epart,
                             int nblk);

// Correction (excluded pairs), reductions, conversions and integrators
void pm_launch_correction(hipStream_t s, const PmBox& box, int n, const double* pos, const double* q,
                          const int* row, const int* partner, const int* cls, double alpha, double csf,
                          int want_virial, long long* force_fix, int ntab, const int* tab_a,
                          const int* tab_b, const int* tab_c, double* epart, int nblk);
void pm_launch_corr_adjoint(hipStream_t s, const PmBox& box, int n, const double* pos, const double* q,
                            const int* row, const int* partner, const int* cls, double alpha, double csf, int mode,
                            const double* v, double* out);
void pm_launch_add_fix_to(hipStream_t s, int m, const long long* fixed, double* out);

// Lanczos vector kernels (pm_krylov.hip); `partial` holds nv * pm_krylov_dot_blocks(n3) doubles
int pm_krylov_dot_blocks(int n3);
void pm_launch_scale_copy(hipStream_t s, int n3, const double* in, const double* w, double* out);
void pm_launch_fix_combine(hipStream_t s, int n3, const long long* fix, const double* dbl, const double* w,
                           double* out);
void pm_launch_dots(hipStream_t s, int n3, int nv, const double* V, const double* x, double* partial, double* out,
                    int acc_row, double* acc, int acc_set, int take_sqrt);
void pm_launch_gs_update(hipStream_t s, int n3, int nv, const double* V, const double* c, double* w);
void pm_launch_normalise(hipStream_t s, int n3, const double* w, const double* nrm, double tiny, double* out);
void pm_launch_ritz(hipStream_t s, int n3, int nv, const double* V, const double* S, int ny, double* Y);
void pm_launch_reduce_energy(hipStream_t s, const double* parts, int nparts, double constant_coul,
                             double constant_vir_diag, double* out, int overwrite, double* scratch);
int pm_reduce_scratch_records(int nparts);
void pm_launch_zero_ll(hipStream_t s, long long* p, long long count);
void pm_launch_zero_i(hipStream_t s, int* p, long long count);
void pm_launch_zero_d(hipStream_t s, double* p, long long count);
void pm_launch_zero_force_cells(hipStream_t s, const int* rebuild, long long* f, long long nf, long long* fs,
                                long long nfs, int* cc, int nc);
void pm_launch_fix_to_double(hipStream_t s, int n3, const long long* f, double* out);
void pm_launch_add_fix(hipStream_t s, int n3, const long long* f, double* out);
// The start of the next evaluation, run by the MD position update on the
// positions it writes: the list-reuse displacement test (pm_launch_check_disp's
// test; flag null: none) and the binning (pm_launch_bin's; bin_count null:
// none), so that the evaluation launches neither.
struct PmStepWork {
  double to_fix[3], fix_to_A[3];
  const int4* x_build;
  float lim2;
  int* flag;                 // flag[0]: rebuild; with x_prune, flag[1]: prune the inner lists
  const int4* x_prune;
  float lim2_in;
  PmBox box;
  int* cell_of; int* rank; float* key; int* bin_count;
};
// leapfrog kick (+ drift); zero_f: clear the forces after reading them (the
// next evaluation then needs no clearing of its own)
void pm_launch_kick(hipStream_t s, int n, int mode, double h, const double* mass, long long* f,
                    double* v, double* x, const PmStepWork* w = nullptr, int zero_f = 0);
// MD state the integrator kernels read from device memory
struct PmMdDevice {
  unsigned long long counter;   // Langevin noise counter (2 per step)
  double temperature;           // Langevin bath temperature (K)
};
void pm_launch_langevin_kick(hipStream_t s, int n, double dt_akma, double gamma_akma, const double* mass,
                             long long* f, double* v, unsigned long long seed, const PmMdDevice* md,
                             int offset, int zero_f = 0);
void pm_launch_md_advance(hipStream_t s, PmMdDevice* md);
void pm_launch_scale(hipStream_t s, int n3, double factor, double* v);
// x += h v for n atoms
void pm_launch_drift(hipStream_t s, int n, double h, const double* v, double* x, const PmStepWork* w = nullptr);

// SHAKE / RATTLE bond constraints (pm_shake.hip): constraints ordered by
// connected group (gstart), each group's in the reference's list order
#define PM_CONS_WATER (-2)
struct PmConstraints {
  int ngroup;
  int ncons;                 // constraints in total (the mean error's divisor)
  int maxit;                 // maxIter: at most maxit + 1 sweeps
  const int* gstart;         // [ngroup + 1]
  const int* a1;
  const int* a2;
  const int* la;             // small groups: local indices of a1, a2 (la1 | la2 << 2)
  const int4* gatom;         // [ngroup] atoms of groups of at most 4 atoms and 6
                             // constraints (-1 unused; w = PM_CONS_WATER: 3 atoms with
                             // constraints (0,1), (0,2), (1,2) in that order); x = -1
                             // for other groups
  const double* r0;          // rest lengths
  const double* rmass;       // 1 / scaledMass per atom
  double eps;
  double dt;                 // timestep / TIMEFACTOR
  double* err;               // [(maxit + 1) * waves] wave sums of the errors after s sweeps
  int* sweeps;               // [1] sweeps run (may be null)
  double* scratch;           // [3n] probe copy of positions (SHAKE) or velocities (RATTLE)
};
// SHAKE (rattle 0: x, v corrected, xl the reference positions, then set to
// x) or RATTLE (rattle 1: v corrected); two launches
bool pm_launch_constraints(hipStream_t s, const PmConstraints& c, int rattle, double* x, double* v, double* xl);
// The allocation-free entry point's work space: a file-scope device array of
// PM_NOALLOC_SCRATCH_BYTES in pm_shake.hip.  pm_scratch_init (with `device`
// current; once per device, later calls return at once) looks up its device
// address; pm_scratch_address returns it (nullptr before the lookup).
#define PM_NOALLOC_SCRATCH_BYTES (16u << 20)
#define PM_MAX_DEVICES 64
int pm_scratch_init(int device);
void* pm_scratch_address(int device);

// Bonded terms (Bond / Angle system forces), per-atom gather formulation
struct PmBondedLaunch {
  int n;
  const double* pos;
  // per-atom CSR of the bonds and angles an atom belongs to
  const int* b_row; const int* b_idx;     // bond index
  const int* a_row; const int* a_idx;     // angle index
  int nb, na;
  const int* b_a; const int* b_b; const double* b_k; const double* b_r0;
  const int* a_a; const int* a_b; const int* a_c; const double* a_k; const double* a_t0;
  const double* a_kub; const double* a_rub;
  int do_bond, do_angle;
  // torsions, kind 0 dihedrals and kind 1 impropers: per-atom CSR of the
  // torsions an atom belongs to, the four atoms, and each torsion's terms
  // term[t] .. term[t+1]-1 (force constant, periodicity, phase in radians)
  const int* t_row[2]; const int* t_idx[2];
  int nt[2];
  const int4* t_atoms[2]; const int* t_term[2];
  const double* t_k[2]; const int* t_n[2]; const double* t_phase[2];
  int do_tors[2];
  long long* force_fix;
  double* epart;        // records: bond, angle, dihedral, improper blocks
  int want_virial;
  int want_energy;      // write the energy records (skipped on steps that report none)
};
// Energy-record blocks of each kind (bond, angle, dihedral, improper) for a.
void pm_bonded_blocks(const PmBondedLaunch& a, int nblk[4]);
void pm_launch_bonded(hipStream_t s, const PmBox& box, const PmBondedLaunch& a);

// Per-atom force terms that the PME gather kernel adds for its atoms, each
// rounded to fixed point on its own as the separate kernels round it (the
// integer sum is then the same as theirs).
struct PmAtomTail {
  PmBox box;
  int corr;                    // PME excluded-pair correction
  const double* pos;
  const double* q;
  const int* row; const int* partner; const int* cls;
  double alpha, csf;
  int bonded;                  // bonded terms (b.do_bond / do_angle / do_tors)
  PmBondedLaunch b;
};

// ------------------------------------------------- batched launch records
// Replica batches run one launch per kernel over every replica.  While a
// recorder is set on the calling thread, the launchers of the MD step append
// a record (kernel kind, template variant, blocks, LDS bytes, the kernel's
// arguments) instead of launching; the batch checks that every replica
// recorded the same kernel sequence, uploads the argument tables and launches
// each kind once: a block finds its replica from the table's prefix of block
// counts and runs the same kernel body on that replica's arguments.
enum PmRecKind {
  PM_REC_KICK_DRIFT = 1, PM_REC_LANGEVIN_KICK, PM_REC_DRIFT, PM_REC_MD_ADVANCE, PM_REC_ZERO_FORCE_CELLS,
  PM_REC_CHECK_DISP, PM_REC_BIN, PM_REC_SCAN_CELLS, PM_REC_SCATTER, PM_REC_FILL_SLOTS,
  PM_REC_LIST, PM_REC_PAIR, PM_REC_FOLD_SLOTS, PM_REC_BONDED_FORCE, PM_REC_BONDED_ENERGY,
  PM_REC_REDUCE1, PM_REC_REDUCE2, PM_REC_PRUNE
};
#define PM_REC_ARGMAX 1024
struct PmRecord {
  int kind;
  int variant;          // template variant (pair: potential/switching combination and precision)
  unsigned blocks;      // blocks of this replica's launch (0: nothing to do for it)
  unsigned threads;     // threads per block
  unsigned lds;         // dynamic LDS bytes
  unsigned argbytes;
  alignas(16) unsigned char args[PM_REC_ARGMAX];
};
// the calling thread's recorder (null: launch normally)
void pm_set_recorder(void* rec);
bool pm_recording();
void pm_record(int kind, int variant, unsigned blocks, unsigned threads, unsigned lds, const void* args,
               unsigned argbytes);
// One launch of `kind` over nrep replicas: args_table holds nrep argument
// structs of argbytes each, prefix nrep + 1 block offsets (prefix[nrep] =
// total blocks).  Returns false for a kind/variant this file does not launch.
bool pm_launch_batched_md(hipStream_t s, int kind, int variant, const void* args_table, const int* prefix,
                          int nrep, unsigned total_blocks, unsigned threads, unsigned lds);
bool pm_launch_batched_nb(hipStream_t s, int kind, int variant, const void* args_table, const int* prefix,
                          int nrep, unsigned total_blocks, unsigned threads, unsigned lds);
bool pm_launch_batched_bonded(hipStream_t s, int kind, int variant, const void* args_table, const int* prefix,
                              int nrep, unsigned total_blocks, unsigned threads, unsigned lds);

// the replica of block `b` (prefix: nrep + 1 non-decreasing block offsets)
__device__ inline int pm_batch_replica(const int* __restrict__ prefix, int nrep, int b) {
  int lo = 0, hi = nrep - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}
