// ref_harness.cpp -- C entry points around the compiled ProtoMol reference.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's own sources (read in place from the reference tree) into
// oracle/_ref/libpmref.so.  The library instantiates the reference's force
// templates exactly as factories/registerForceExemplarsCutoff.cpp and
// registerForceExemplarsFastElectrostatic.cpp register them, builds a
// Topology<PeriodicBoundaryConditions,CubicCellManager> from plain arrays the
// way frontend/buildTopology.cpp fills it, and evaluates forces the way
// applications/mdEnergy-app/mdEnergy.cpp:387-392 does (clear energies, zero
// forces, evaluate).  Its entry points mirror oracle/pmoracle.h so that the
// tests can compare the restatement against the reference call for call.
#include "Report.h"
// base/Grid.h names report/error/endr unqualified; the reference includes it
// after a using-directive (forces/NonbondedPMEwaldSystemForce.h:18).
using namespace ProtoMol::Report;
#include "Topology.h"
#include "PeriodicBoundaryConditions.h"
#include "CubicCellManager.h"
#include "CellListEnumerator_periodicBoundaries.h"
#include "NonbondedCutoffSystemForce.h"
#include "OneAtomPair.h"
#include "OneAtomPairTwo.h"
#include "LennardJonesForce.h"
#include "CoulombForce.h"
#include "CoulombEwaldRealForce.h"
#include "C1SwitchingFunction.h"
#include "C2SwitchingFunction.h"
#include "CutoffSwitchingFunction.h"
#include "NonbondedPMEwaldSystemForce.h"
#include "BSpline.h"
#include "ScalarStructure.h"
#include "Vector3DBlock.h"
#include "ForceGroup.h"
#include "Hessian.h"
#include "BondSystemForce.h"
#include "AngleSystemForce.h"
#include "DihedralSystemForce.h"
#include "ImproperSystemForce.h"
#include "PSFReader.h"
#include "PARReader.h"
#include "PSF.h"
#include "PAR.h"
#include "buildTopology.h"
#include "DCDTrajectoryWriter.h"
#include "XYZWriter.h"
#include "topologyutilities.h"

#include <vector>
#include <map>

using namespace ProtoMol;

namespace {

typedef Topology<PeriodicBoundaryConditions, CubicCellManager> Topo;

enum { POT_LJ = 0, POT_COULOMB = 1, POT_EWALDREAL = 2 };
enum { SWF_CUTOFF = 0, SWF_C1 = 1, SWF_C2 = 2 };

struct Term {
  int pot;
  int swf;
  double alpha;
  double switchon;
  double cutoff;
};

// Counters fed by the constraint hooks of OneAtomPair(Two)
// (forces/oneAtomContraints.h): every candidate pair passes PRE_CHECK, every
// pair that reaches the force computation passes POST_CHECK.
std::vector<int>* g_cellOf = 0;
long long* g_cand = 0;
long long* g_inrc = 0;

struct CountConstraint {
  enum { PRE_CHECK = 1, POST_CHECK = 1 };
  static bool check(const GenericTopology*, int i, int) {
    if (g_cand) g_cand[(*g_cellOf)[i]]++;
    return true;
  }
  static bool check(const GenericTopology*, int) { return true; }
  static std::string getPrefixId() { return ""; }
  static std::string getPostfixId() { return "Count"; }
  static void check(const GenericTopology*, int i, int, const Vector3D&, Real, const Vector3D&) {
    if (g_inrc) g_inrc[(*g_cellOf)[i]]++;
  }
};

Topo* makeTopo(int n, const double* pos, const double* q, const int* type, int ntypes,
               const double* ljA, const double* ljB, const double* ljA14, const double* ljB14,
               const int* molecule, int nexcl, const int* ea, const int* eb, const int* ec,
               double csf, const double* box, const double* origin, double cellsize,
               Vector3DBlock& positions) {
  // frontend/buildTopology.cpp fills the same members from PSF/PAR data.
  Topo* t = new Topo(csf, ExclusionType(ExclusionType::ONE4MODIFIED),
                     PeriodicBoundaryConditions(Vector3D(box[0], 0, 0), Vector3D(0, box[1], 0),
                                                Vector3D(0, 0, box[2]),
                                                Vector3D(origin[0], origin[1], origin[2])),
                     CubicCellManager(cellsize));
  int nmol = 0;
  for (int i = 0; i < n; i++) nmol = std::max(nmol, molecule[i] + 1);
  t->atomTypes.resize(ntypes);
  t->atoms.resize(n);
  t->molecules.resize(nmol);
  for (int i = 0; i < n; i++) {
    t->atoms[i].type = type[i];
    t->atoms[i].scaledCharge = q[i];
    t->atoms[i].scaledMass = 1.0;
    t->atoms[i].molecule = molecule[i];
    t->molecules[molecule[i]].atoms.push_back(i);
  }
  t->lennardJonesParameters.resize(ntypes);
  for (int a = 0; a < ntypes; a++)
    for (int b = 0; b < ntypes; b++)
      t->lennardJonesParameters.set(a, b, LennardJonesParameters(ljA[a * ntypes + b], ljB[a * ntypes + b],
                                                                 ljA14[a * ntypes + b], ljB14[a * ntypes + b]));
  t->exclusions.resize(n);
  for (int k = 0; k < nexcl; k++) t->exclusions.add(ea[k], eb[k], (ExclusionClass)ec[k]);
  t->exclusions.optimize();
  positions.resize(n);
  for (int i = 0; i < n; i++) positions[i] = Vector3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  return t;
}

void readEnergies(const ScalarStructure& e, double* energies) {
  energies[0] += e[ScalarStructure::COULOMB];
  energies[1] += e[ScalarStructure::LENNARDJONES];
  energies[2] += e[ScalarStructure::BOND];
  energies[3] += e[ScalarStructure::ANGLE];
  const ScalarStructure::Index v[9] = {ScalarStructure::VIRIALXX, ScalarStructure::VIRIALXY,
                                       ScalarStructure::VIRIALXZ, ScalarStructure::VIRIALYX,
                                       ScalarStructure::VIRIALYY, ScalarStructure::VIRIALYZ,
                                       ScalarStructure::VIRIALZX, ScalarStructure::VIRIALZY,
                                       ScalarStructure::VIRIALZZ};
  for (int k = 0; k < 9; k++) energies[7 + k] += e[v[k]];
}

// Switching-function and potential objects from a Term, as
// registerForceExemplarsCutoff.cpp instantiates them.
template <class S> S makeSwf(const Term& t);
template <> C2SwitchingFunction makeSwf<C2SwitchingFunction>(const Term& t) {
  return C2SwitchingFunction(t.switchon, t.cutoff);
}
template <> C1SwitchingFunction makeSwf<C1SwitchingFunction>(const Term& t) {
  return C1SwitchingFunction(t.cutoff);
}
template <> CutoffSwitchingFunction makeSwf<CutoffSwitchingFunction>(const Term& t) {
  return CutoffSwitchingFunction(t.cutoff);
}
template <class P> P makePot(const Term& t);
template <> LennardJonesForce makePot<LennardJonesForce>(const Term&) { return LennardJonesForce(); }
template <> CoulombForce makePot<CoulombForce>(const Term&) { return CoulombForce(); }
template <> CoulombEwaldRealForce makePot<CoulombEwaldRealForce>(const Term& t) {
  return CoulombEwaldRealForce(t.alpha);
}

template <class Pair>
SystemForce* makeCutoff(double algoCutoff, const Pair& p) {
  return new NonbondedCutoffSystemForce<CubicCellManager, Pair>(algoCutoff, p);
}

template <class S1, class P1, bool Count>
SystemForce* make1(const Term* t, double algoCutoff) {
  if (Count) {
    typedef OneAtomPair<PeriodicBoundaryConditions, S1, P1, CountConstraint> Pair;
    return makeCutoff(algoCutoff, Pair(makePot<P1>(t[0]), makeSwf<S1>(t[0])));
  }
  typedef OneAtomPair<PeriodicBoundaryConditions, S1, P1> Pair;
  return makeCutoff(algoCutoff, Pair(makePot<P1>(t[0]), makeSwf<S1>(t[0])));
}

template <class S1, class P1, class S2, class P2, bool Count>
SystemForce* make2(const Term* t, double algoCutoff) {
  if (Count) {
    typedef OneAtomPairTwo<PeriodicBoundaryConditions, S1, P1, S2, P2, CountConstraint> Pair;
    return makeCutoff(algoCutoff, Pair(makePot<P1>(t[0]), makeSwf<S1>(t[0]), makePot<P2>(t[1]), makeSwf<S2>(t[1])));
  }
  typedef OneAtomPairTwo<PeriodicBoundaryConditions, S1, P1, S2, P2> Pair;
  return makeCutoff(algoCutoff, Pair(makePot<P1>(t[0]), makeSwf<S1>(t[0]), makePot<P2>(t[1]), makeSwf<S2>(t[1])));
}

template <class P1, bool Count>
SystemForce* dispatch1Swf(const Term* t, double c) {
  switch (t[0].swf) {
    case SWF_C1: return make1<C1SwitchingFunction, P1, Count>(t, c);
    case SWF_C2: return make1<C2SwitchingFunction, P1, Count>(t, c);
    default: return make1<CutoffSwitchingFunction, P1, Count>(t, c);
  }
}

template <bool Count>
SystemForce* dispatch1(const Term* t, double c) {
  switch (t[0].pot) {
    case POT_LJ: return dispatch1Swf<LennardJonesForce, Count>(t, c);
    case POT_COULOMB: return dispatch1Swf<CoulombForce, Count>(t, c);
    default: return dispatch1Swf<CoulombEwaldRealForce, Count>(t, c);
  }
}

template <class S1, class P1, class P2, bool Count>
SystemForce* dispatch2Swf2(const Term* t, double c) {
  switch (t[1].swf) {
    case SWF_C1: return make2<S1, P1, C1SwitchingFunction, P2, Count>(t, c);
    case SWF_C2: return make2<S1, P1, C2SwitchingFunction, P2, Count>(t, c);
    default: return make2<S1, P1, CutoffSwitchingFunction, P2, Count>(t, c);
  }
}

template <class P1, class P2, bool Count>
SystemForce* dispatch2Swf1(const Term* t, double c) {
  switch (t[0].swf) {
    case SWF_C1: return dispatch2Swf2<C1SwitchingFunction, P1, P2, Count>(t, c);
    case SWF_C2: return dispatch2Swf2<C2SwitchingFunction, P1, P2, Count>(t, c);
    default: return dispatch2Swf2<CutoffSwitchingFunction, P1, P2, Count>(t, c);
  }
}

template <bool Count>
SystemForce* dispatch2(const Term* t, double c) {
  // The registered OneAtomPairTwo ids put LennardJones first
  // (registerForceExemplarsCutoff.cpp:75-142).
  if (t[0].pot != POT_LJ) return 0;
  switch (t[1].pot) {
    case POT_COULOMB: return dispatch2Swf1<LennardJonesForce, CoulombForce, Count>(t, c);
    case POT_EWALDREAL: return dispatch2Swf1<LennardJonesForce, CoulombEwaldRealForce, Count>(t, c);
    default: return 0;
  }
}

template <bool R, bool Q, bool C>
SystemForce* makePME(int nx, int ny, int nz, int order, double rc, double acc, double alpha) {
  return new NonbondedPMEwaldSystemForce<PeriodicBoundaryConditions, CubicCellManager, R, Q, C,
                                         BSpline, CutoffSwitchingFunction>(nx, ny, nz, order, rc, acc,
                                                                           alpha, 3.0);
}

void cellOfAtoms(const Topo* t, std::vector<int>& cellOf) {
  const ArrayCellListStructure& cl = t->cellLists;
  int ny = cl.getDimY(), nz = cl.getDimZ();
  cellOf.assign(t->atoms.size(), -1);
  for (ArrayCellListStructure::const_iterator it = cl.begin(); it != cl.end(); ++it) {
    if (it->second < 0) continue;
    int c = (it->first.x * ny + it->first.y) * nz + it->first.z;
    for (int a = it->second; a != -1; a = t->atoms[a].cellListNext) cellOf[a] = c;
  }
}

}  // namespace

extern "C" {

int ref_cells(int n, const double* pos, const double* box, const double* origin, double cellsize,
              int* dims, double* realcell, int* cell, int* head, int* next, int ncap) {
  std::vector<double> q(n, 0.0), lj(1, 0.0);
  std::vector<int> type(n, 0), mol(n);
  for (int i = 0; i < n; i++) mol[i] = i;
  Vector3DBlock positions;
  Topo* t = makeTopo(n, pos, &q[0], &type[0], 1, &lj[0], &lj[0], &lj[0], &lj[0], &mol[0], 0, 0, 0, 0,
                     1.0, box, origin, cellsize, positions);
  t->uncacheCellList();
  t->updateCellLists(&positions);
  const ArrayCellListStructure& cl = t->cellLists;
  int nx = cl.getDimX(), ny = cl.getDimY(), nz = cl.getDimZ();
  int nc = nx * ny * nz;
  if (dims) { dims[0] = nx; dims[1] = ny; dims[2] = nz; }
  if (realcell) {
    Vector3D r = t->cellManager.getRealCellSize();
    realcell[0] = r.x; realcell[1] = r.y; realcell[2] = r.z;
  }
  if (head && nc <= ncap) {
    for (int c = 0; c < nc; c++) head[c] = -1;
    // iterate the whole array (begin() is trimmed to the first non-empty cell)
    for (ArrayCellListStructure::const_iterator it = cl.begin(); it != cl.end(); ++it) {
      int c = (it->first.x * ny + it->first.y) * nz + it->first.z;
      head[c] = it->second;
    }
    for (int i = 0; i < n; i++) if (next) next[i] = t->atoms[i].cellListNext;
    if (cell) {
      std::vector<int> co;
      cellOfAtoms(t, co);
      for (int i = 0; i < n; i++) cell[i] = co[i];
    }
  }
  delete t;
  return nc > ncap ? -1 : nc;
}

int ref_cutoff(int n, const double* pos, const double* q, const int* type, int ntypes,
               const double* ljA, const double* ljB, const double* ljA14, const double* ljB14,
               const int* molecule, int nexcl, const int* ea, const int* eb, const int* ec, double csf,
               const double* box, const double* origin, double cellsize, int nterms, const Term* terms,
               double algoCutoff, int wantVirial, double* forces, double* energies, long long* cand,
               long long* inrc) {
  Vector3DBlock positions;
  Topo* t = makeTopo(n, pos, q, type, ntypes, ljA, ljB, ljA14, ljB14, molecule, nexcl, ea, eb, ec, csf,
                     box, origin, cellsize, positions);
  bool count = (cand != 0 || inrc != 0);
  SystemForce* f = 0;
  if (nterms == 1) f = count ? dispatch1<true>(terms, algoCutoff) : dispatch1<false>(terms, algoCutoff);
  else if (nterms == 2) f = count ? dispatch2<true>(terms, algoCutoff) : dispatch2<false>(terms, algoCutoff);
  if (!f) { delete t; return -1; }
  std::vector<int> cellOf;
  if (count) {
    t->uncacheCellList();
    t->updateCellLists(&positions);
    cellOfAtoms(t, cellOf);
    int nc = t->cellLists.getDimX() * t->cellLists.getDimY() * t->cellLists.getDimZ();
    if (cand) for (int c = 0; c < nc; c++) cand[c] = 0;
    if (inrc) for (int c = 0; c < nc; c++) inrc[c] = 0;
    g_cellOf = &cellOf;
    g_cand = cand;
    g_inrc = inrc;
  }
  Vector3DBlock fo;
  fo.resize(n);
  fo.zero();
  ScalarStructure e;
  e.clear();
  e.virial(wantVirial != 0);
  t->uncacheCellList();     // ForceGroup::evaluateSystemForces (ForceGroup.cpp:50)
  f->evaluate(t, &positions, &fo, &e);
  g_cellOf = 0; g_cand = 0; g_inrc = 0;
  for (int i = 0; i < n; i++) {
    forces[3 * i] += fo[i].x; forces[3 * i + 1] += fo[i].y; forces[3 * i + 2] += fo[i].z;
  }
  readEnergies(e, energies);
  delete f;
  delete t;
  return 0;
}

int ref_pme(int n, const double* pos, const double* q, const int* molecule, int nexcl, const int* ea,
            const int* eb, const int* ec, double csf, const double* box, const double* origin,
            double cellsize, int nx, int ny, int nz, int order, double rc, double accuracy, double alpha,
            int flags, int wantVirial, double* forces, double* energies) {
  std::vector<double> lj(1, 0.0);
  std::vector<int> type(n, 0);
  Vector3DBlock positions;
  Topo* t = makeTopo(n, pos, q, &type[0], 1, &lj[0], &lj[0], &lj[0], &lj[0], molecule, nexcl, ea, eb, ec,
                     csf, box, origin, cellsize, positions);
  SystemForce* f = 0;
  switch (flags & 7) {
    case 1: f = makePME<true, false, false>(nx, ny, nz, order, rc, accuracy, alpha); break;
    case 2: f = makePME<false, true, false>(nx, ny, nz, order, rc, accuracy, alpha); break;
    case 3: f = makePME<true, true, false>(nx, ny, nz, order, rc, accuracy, alpha); break;
    case 4: f = makePME<false, false, true>(nx, ny, nz, order, rc, accuracy, alpha); break;
    case 5: f = makePME<true, false, true>(nx, ny, nz, order, rc, accuracy, alpha); break;
    case 6: f = makePME<false, true, true>(nx, ny, nz, order, rc, accuracy, alpha); break;
    case 7: f = makePME<true, true, true>(nx, ny, nz, order, rc, accuracy, alpha); break;
    default: delete t; return -1;
  }
  Vector3DBlock fo;
  fo.resize(n);
  fo.zero();
  ScalarStructure e;
  e.clear();
  e.virial(wantVirial != 0);
  t->uncacheCellList();
  f->evaluate(t, &positions, &fo, &e);
  for (int i = 0; i < n; i++) {
    forces[3 * i] += fo[i].x; forces[3 * i + 1] += fo[i].y; forces[3 * i + 2] += fo[i].z;
  }
  readEnergies(e, energies);
  delete f;
  delete t;
  return 0;
}

// BondSystemForce / AngleSystemForce over topo->bonds / topo->angles
// (forces/BondSystemForce.h:99-145, forces/AngleSystemForce.h:98-200).
int ref_bonded(int n, const double* pos, const double* box, const double* origin, int nb, const int* ba,
               const int* bb, const double* bk, const double* br0, int na, const int* aa, const int* ab,
               const int* ac, const double* ak, const double* at0, const double* akub, const double* arub,
               int which, int wantVirial, double* forces, double* energies) {
  std::vector<double> q(n, 0.0), lj(1, 0.0);
  std::vector<int> type(n, 0), mol(n, 0);
  Vector3DBlock positions;
  Topo* t = makeTopo(n, pos, &q[0], &type[0], 1, &lj[0], &lj[0], &lj[0], &lj[0], &mol[0], 0, 0, 0, 0, 1.0,
                     box, origin, 5.0, positions);
  for (int k = 0; k < nb; k++) t->bonds.push_back(Bond(ba[k], bb[k], bk[k], br0[k]));
  for (int k = 0; k < na; k++) {
    Angle a;
    a.atom1 = aa[k]; a.atom2 = ab[k]; a.atom3 = ac[k];
    a.forceConstant = ak[k]; a.restAngle = at0[k];
    a.ureyBradleyConstant = akub[k]; a.ureyBradleyRestLength = arub[k];
    t->angles.push_back(a);
  }
  Vector3DBlock fo;
  fo.resize(n);
  fo.zero();
  ScalarStructure e;
  e.clear();
  e.virial(wantVirial != 0);
  if (which & 1) { BondSystemForce<PeriodicBoundaryConditions> f; f.evaluate(t, &positions, &fo, &e); }
  if (which & 2) { AngleSystemForce<PeriodicBoundaryConditions> f; f.evaluate(t, &positions, &fo, &e); }
  for (int i = 0; i < n; i++) {
    forces[3 * i] += fo[i].x; forces[3 * i + 1] += fo[i].y; forces[3 * i + 2] += fo[i].z;
  }
  readEnergies(e, energies);
  delete t;
  return 0;
}

void ref_bspline(int order, double w, double* theta, double* dtheta) {
  BSpline b(order, w);
  for (int k = 0; k < order; k++) { theta[k] = b.theta[k]; dtheta[k] = b.dTheta[k]; }
}

// Dense Hessian of LJ + Coulomb forces through Hessian::findForces/evaluate
// (integrators/Hessian.cpp:88-267, 269-624), contracted with v.
int ref_hessian_vec(int n, const double* pos, const double* q, const int* type, int ntypes,
                    const double* ljA, const double* ljB, const double* ljA14, const double* ljB14,
                    const int* molecule, int nexcl, const int* ea, const int* eb, const int* ec,
                    double csf, const double* box, const double* origin, double cellsize, int nterms,
                    const Term* terms, const double* v, double* hv) {
  Vector3DBlock positions;
  Topo* t = makeTopo(n, pos, q, type, ntypes, ljA, ljB, ljA14, ljB14, molecule, nexcl, ea, eb, ec, csf,
                     box, origin, cellsize, positions);
  ForceGroup* fg = new ForceGroup();
  for (int k = 0; k < nterms; k++) {
    SystemForce* f = dispatch1<false>(&terms[k], terms[k].cutoff);
    if (!f) { delete fg; delete t; return -1; }
    fg->addSystemForce(f);
  }
  unsigned int sz = 3 * n;
  Hessian h(sz);
  h.findForces(fg);
  h.clear();
  h.evaluate(&positions, t, false);
  for (unsigned int r = 0; r < sz; r++) {
    double s = 0.0;
    for (unsigned int c = 0; c < sz; c++) s += h.hessM[(size_t)r * sz + c] * v[c];
    hv[r] = s;
  }
  delete fg;
  delete t;
  return 0;
}


// Dihedral / Improper system forces over topo->dihedrals / ->impropers
// (forces/DihedralSystemForce.h, forces/ImproperSystemForce.h, both through
// MTorsionSystemForce::calcTorsion).  Torsion t has atoms[4t..4t+3] and the
// terms term_row[t] .. term_row[t+1]-1.
int ref_torsions(int n, const double* pos, const double* box, const double* origin, int nt, const int* atoms,
                 const int* term_row, const double* k, const int* period, const double* phase, int improper,
                 int wantVirial, double* forces, double* energies) {
  std::vector<double> q(n, 0.0), lj(1, 0.0);
  std::vector<int> type(n, 0), mol(n, 0);
  Vector3DBlock positions;
  Topo* t = makeTopo(n, pos, &q[0], &type[0], 1, &lj[0], &lj[0], &lj[0], &lj[0], &mol[0], 0, 0, 0, 0, 1.0,
                     box, origin, 5.0, positions);
  for (int i = 0; i < nt; i++) {
    Torsion tor;
    tor.atom1 = atoms[4 * i]; tor.atom2 = atoms[4 * i + 1]; tor.atom3 = atoms[4 * i + 2]; tor.atom4 = atoms[4 * i + 3];
    tor.multiplicity = term_row[i + 1] - term_row[i];
    for (int u = term_row[i]; u < term_row[i + 1]; u++) {
      tor.forceConstant.push_back(k[u]);
      tor.periodicity.push_back(period[u]);
      tor.phaseShift.push_back(phase[u]);
    }
    (improper ? t->impropers : t->dihedrals).push_back(tor);
  }
  Vector3DBlock fo;
  fo.resize(n);
  fo.zero();
  ScalarStructure e;
  e.clear();
  e.virial(wantVirial != 0);
  if (improper) { ImproperSystemForce<PeriodicBoundaryConditions> f; f.evaluate(t, &positions, &fo, &e); }
  else { DihedralSystemForce<PeriodicBoundaryConditions> f; f.evaluate(t, &positions, &fo, &e); }
  for (int i = 0; i < n; i++) {
    forces[3 * i] += fo[i].x; forces[3 * i + 1] += fo[i].y; forces[3 * i + 2] += fo[i].z;
  }
  readEnergies(e, energies);
  energies[4] = e[ScalarStructure::DIHEDRAL];
  energies[5] = e[ScalarStructure::IMPROPER];
  delete t;
  return 0;
}

// The topology the reference's protomol application builds from a PSF and a
// PAR file (applications/protomol-app/protomol.cpp:340-403): PSFReader,
// PARReader, buildTopology with the exclusion type and the dihedral
// multiplicity rule (-1: the application's default).  info[] receives
// atoms, types, bonds, angles, dihedrals, dihedral terms, impropers,
// improper terms, molecules, degrees of freedom, the detected PAR dialect
// (19 or 28) and the multiplicity rule used.
void* ref_topo_load(const char* psfPath, const char* parPath, const char* exclude, int dmult, int* info) {
  PSFReader pr;
  if (!pr.open(psfPath)) return 0;
  PSF psf;
  if (!(pr >> psf)) return 0;
  PARReader ar;
  if (!ar.open(parPath)) return 0;
  PAR par;
  if (!(ar >> par)) return 0;
  const bool mult = dmult < 0 ? (ar.getCharmmTypeDetected() != PAR::CHARMM28) : (dmult != 0);
  Topo* t = new Topo(1.0, ExclusionType(std::string(exclude)),
                     PeriodicBoundaryConditions(Vector3D(1000, 0, 0), Vector3D(0, 1000, 0), Vector3D(0, 0, 1000),
                                                Vector3D(0, 0, 0)),
                     CubicCellManager(5.0));
  t->exclude = ExclusionType(std::string(exclude));
  buildTopology(t, psf, par, mult);
  int dt = 0, it = 0;
  for (size_t i = 0; i < t->dihedrals.size(); i++) dt += t->dihedrals[i].multiplicity;
  for (size_t i = 0; i < t->impropers.size(); i++) it += t->impropers[i].multiplicity;
  info[0] = (int)t->atoms.size(); info[1] = (int)t->atomTypes.size(); info[2] = (int)t->bonds.size();
  info[3] = (int)t->angles.size(); info[4] = (int)t->dihedrals.size(); info[5] = dt;
  info[6] = (int)t->impropers.size(); info[7] = it; info[8] = (int)t->molecules.size();
  info[9] = (int)t->degreesOfFreedom; info[10] = ar.getCharmmTypeDetected() == PAR::CHARMM28 ? 28 : 19;
  info[11] = mult ? 1 : 0;
  return t;
}

void ref_topo_free(void* h) { delete (Topo*)h; }

int ref_topo_atoms(void* h, double* q, double* mass, int* type, int* molecule, int* water) {
  Topo* t = (Topo*)h;
  for (size_t i = 0; i < t->atoms.size(); i++) {
    q[i] = t->atoms[i].scaledCharge; mass[i] = t->atoms[i].scaledMass;
    type[i] = t->atoms[i].type; molecule[i] = t->atoms[i].molecule;
    water[i] = t->molecules[t->atoms[i].molecule].water ? 1 : 0;
  }
  return 0;
}

int ref_topo_type_name(void* h, int k, char* buf, int cap) {
  Topo* t = (Topo*)h;
  std::strncpy(buf, t->atomTypes[k].name.c_str(), cap - 1);
  buf[cap - 1] = 0;
  return 0;
}

int ref_topo_lj(void* h, double* A, double* B, double* A14, double* B14) {
  Topo* t = (Topo*)h;
  const int nt = (int)t->atomTypes.size();
  for (int a = 0; a < nt; a++)
    for (int b = 0; b < nt; b++) {
      const LennardJonesParameters& p = t->lennardJonesParameters(a, b);
      A[a * nt + b] = p.A; B[a * nt + b] = p.B; A14[a * nt + b] = p.A14; B14[a * nt + b] = p.B14;
    }
  return 0;
}

int ref_topo_bonds(void* h, int* a, int* b, double* k, double* r0) {
  Topo* t = (Topo*)h;
  for (size_t i = 0; i < t->bonds.size(); i++) {
    a[i] = t->bonds[i].atom1; b[i] = t->bonds[i].atom2; k[i] = t->bonds[i].springConstant; r0[i] = t->bonds[i].restLength;
  }
  return 0;
}

int ref_topo_angles(void* h, int* a, int* b, int* c, double* k, double* t0, double* kub, double* rub) {
  Topo* t = (Topo*)h;
  for (size_t i = 0; i < t->angles.size(); i++) {
    const Angle& g = t->angles[i];
    a[i] = g.atom1; b[i] = g.atom2; c[i] = g.atom3; k[i] = g.forceConstant; t0[i] = g.restAngle;
    kub[i] = g.ureyBradleyConstant; rub[i] = g.ureyBradleyRestLength;
  }
  return 0;
}

int ref_topo_torsions(void* h, int improper, int* atoms, int* term_row, double* k, int* period, double* phase) {
  Topo* t = (Topo*)h;
  const std::vector<Torsion>& L = improper ? t->impropers : t->dihedrals;
  int u = 0;
  for (size_t i = 0; i < L.size(); i++) {
    atoms[4 * i] = L[i].atom1; atoms[4 * i + 1] = L[i].atom2; atoms[4 * i + 2] = L[i].atom3; atoms[4 * i + 3] = L[i].atom4;
    term_row[i] = u;
    for (int m = 0; m < L[i].multiplicity; m++, u++) {
      k[u] = L[i].forceConstant[m]; period[u] = L[i].periodicity[m]; phase[u] = L[i].phaseShift[m];
    }
  }
  term_row[L.size()] = u;
  return 0;
}

// ExclusionTable::check for the given pairs
int ref_topo_exclusion_class(void* h, int np, const int* a, const int* b, int* cls) {
  Topo* t = (Topo*)h;
  for (int k = 0; k < np; k++) cls[k] = (int)t->exclusions.check(a[k], b[k]);
  return 0;
}

// removeLinearMomentum then removeAngularMomentum (topologyutilities.cpp),
// as protomol.cpp:429-444 applies them at start-up
int ref_topo_remove_momentum(void* h, const double* pos, double* vel, int linear, int angular) {
  Topo* t = (Topo*)h;
  const int n = (int)t->atoms.size();
  Vector3DBlock p, v;
  p.resize(n);
  v.resize(n);
  for (int i = 0; i < n; i++) {
    p[i] = Vector3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    v[i] = Vector3D(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
  }
  if (linear) removeLinearMomentum(&v, t);
  if (angular) removeAngularMomentum(&p, &v, t);
  for (int i = 0; i < n; i++) { vel[3 * i] = v[i].x; vel[3 * i + 1] = v[i].y; vel[3 * i + 2] = v[i].z; }
  return 0;
}

// randomVelocity (topologyutilities.cpp:24-58) on the loaded topology.  The
// reference's generator is seeded on its first use in the process only.
int ref_topo_random_velocity(void* h, double temperature, unsigned int seed, double* vel) {
  Topo* t = (Topo*)h;
  Vector3DBlock v;
  randomVelocity(temperature, t, &v, seed);
  for (size_t i = 0; i < t->atoms.size(); i++) { vel[3 * i] = v[i].x; vel[3 * i + 1] = v[i].y; vel[3 * i + 2] = v[i].z; }
  return 0;
}

// DCDTrajectoryWriter (io/DCDTrajectoryWriter.cpp) with its defaults
// (time step 1, first step 1), one write() per frame; XYZWriter
// (io/XYZWriter.cpp) of one frame with the given names ('\n'-separated).
int ref_write_dcd(const char* path, int n, int nframes, const double* xyz) {
  DCDTrajectoryWriter w(path);
  for (int f = 0; f < nframes; f++) {
    Vector3DBlock c;
    c.resize(n);
    for (int i = 0; i < n; i++) {
      const double* p = xyz + 3 * ((size_t)f * n + i);
      c[i] = Vector3D(p[0], p[1], p[2]);
    }
    if (!w.write(c)) return -1;
  }
  return 0;
}

int ref_write_xyz(const char* path, int n, const double* xyz, const char* names) {
  std::vector<std::string> nm;
  std::string cur;
  for (const char* p = names; *p; p++) {
    if (*p == '\n') { nm.push_back(cur); cur.clear(); } else cur += *p;
  }
  if (!cur.empty()) nm.push_back(cur);
  Vector3DBlock c;
  c.resize(n);
  for (int i = 0; i < n; i++) c[i] = Vector3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  XYZWriter w(path);
  return w.write(c, nm) ? 0 : -1;
}

}  // extern "C"
