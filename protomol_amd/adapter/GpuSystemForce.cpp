// GpuSystemForce.cpp -- the ProtoMol side of the drop-in boundary: the
// SystemForce adapter, the topology hand-off to the library context, and
// the registration over the CPU exemplars.  Compiled with the ProtoMol
// framework headers (base/, topology/, forces/, frontend/) and linked
// against libpmgpu.so.
#include "GpuSystemForce.h"

#include "ForceFactory.h"
#include "ForceGroup.h"
#include "PeriodicBoundaryConditions.h"
#include "CubicCellManager.h"
#include "Topology.h"
#include "ScalarStructure.h"
#include "Vector3DBlock.h"
#include "Report.h"

#include <algorithm>
#include <set>
#include <utility>

using namespace ProtoMol::Report;

namespace ProtoMol {

  namespace {
    typedef Topology<PeriodicBoundaryConditions, CubicCellManager> PeriodicTopology;

    struct GpuState {
      pm_context* ctx = 0;
      bool tried = false;
      const GenericTopology* topo = 0;
      unsigned int natoms = 0;
      int precision = PM_PRECISION_MIXED;
      std::string err;
    };
    GpuState& state() {
      static GpuState s;
      return s;
    }

    // The ordered exclusion adds that rebuild `table`: its entries in order
    // (the pairs ExclusionTable::getTable() keeps, as the PME correction term
    // visits them), then every pair once more with the class check() gives
    // now, so that the last add of each pair is the class the nonbonded
    // forces see (topology/ExclusionTable.cpp:120-139).
    void exclusionAdds(const ExclusionTable& table, std::vector<int>& a, std::vector<int>& b, std::vector<int>& c) {
      const std::vector<ExclusionPair>& t = table.getTable();
      std::set<std::pair<int, int> > seen;
      for (unsigned int k = 0; k < t.size(); k++) {
        a.push_back(t[k].a1);
        b.push_back(t[k].a2);
        c.push_back(static_cast<int>(t[k].excl));
      }
      for (unsigned int k = 0; k < t.size(); k++) {
        if (!seen.insert(std::make_pair(t[k].a1, t[k].a2)).second) continue;
        a.push_back(t[k].a1);
        b.push_back(t[k].a2);
        c.push_back(static_cast<int>(table.check(t[k].a1, t[k].a2)));
      }
    }

    // Everything the topology holds that the library needs
    // (frontend/buildTopology.cpp fills these members from PSF/PAR data).
    int upload(pm_context* ctx, const PeriodicTopology* t) {
      const unsigned int n = t->atoms.size();
      const int nt = static_cast<int>(t->atomTypes.size());
      std::vector<double> q(n), mass(n), A(nt * nt), B(nt * nt), A14(nt * nt), B14(nt * nt);
      std::vector<int> type(n), mol(n);
      for (unsigned int i = 0; i < n; i++) {
        q[i] = t->atoms[i].scaledCharge;
        mass[i] = t->atoms[i].scaledMass;
        type[i] = t->atoms[i].type;
        mol[i] = t->atoms[i].molecule;
      }
      for (int x = 0; x < nt; x++)
        for (int y = 0; y < nt; y++) {
          const LennardJonesParameters& p = t->lennardJonesParameters(x, y);
          A[x * nt + y] = p.A;
          B[x * nt + y] = p.B;
          A14[x * nt + y] = p.A14;
          B14[x * nt + y] = p.B14;
        }
      std::vector<int> ea, eb, ec;
      exclusionAdds(t->exclusions, ea, eb, ec);
      int r = pm_set_topology(ctx, static_cast<int>(n), &q[0], &type[0], nt, &A[0], &B[0], &A14[0], &B14[0], &mol[0],
                              &mass[0], static_cast<int>(ea.size()), ea.empty() ? 0 : &ea[0],
                              eb.empty() ? 0 : &eb[0], ec.empty() ? 0 : &ec[0], t->coulombScalingFactor);
      if (r != PM_OK) return r;
      // bonds and angles (BondSystemForce / AngleSystemForce inputs)
      std::vector<int> ba, bb, aa, ab, ac;
      std::vector<double> bk, br, ak, at, au, ar;
      for (unsigned int k = 0; k < t->bonds.size(); k++) {
        ba.push_back(t->bonds[k].atom1);
        bb.push_back(t->bonds[k].atom2);
        bk.push_back(t->bonds[k].springConstant);
        br.push_back(t->bonds[k].restLength);
      }
      for (unsigned int k = 0; k < t->angles.size(); k++) {
        aa.push_back(t->angles[k].atom1);
        ab.push_back(t->angles[k].atom2);
        ac.push_back(t->angles[k].atom3);
        ak.push_back(t->angles[k].forceConstant);
        at.push_back(t->angles[k].restAngle);
        au.push_back(t->angles[k].ureyBradleyConstant);
        ar.push_back(t->angles[k].ureyBradleyRestLength);
      }
      if (!ba.empty() || !aa.empty()) {
        r = pm_set_bonded(ctx, static_cast<int>(ba.size()), ba.empty() ? 0 : &ba[0], bb.empty() ? 0 : &bb[0],
                          bk.empty() ? 0 : &bk[0], br.empty() ? 0 : &br[0], static_cast<int>(aa.size()),
                          aa.empty() ? 0 : &aa[0], ab.empty() ? 0 : &ab[0], ac.empty() ? 0 : &ac[0],
                          ak.empty() ? 0 : &ak[0], at.empty() ? 0 : &at[0], au.empty() ? 0 : &au[0],
                          ar.empty() ? 0 : &ar[0]);
        if (r != PM_OK) return r;
      }
      // dihedrals (kind 0) and impropers (kind 1): Torsion records with their terms
      for (int kind = 0; kind < 2; kind++) {
        const std::vector<Torsion>& tv = kind == 0 ? t->dihedrals : t->impropers;
        if (tv.empty()) continue;
        std::vector<int> atoms, row(1, 0), per;
        std::vector<double> k, ph;
        for (unsigned int m = 0; m < tv.size(); m++) {
          const Torsion& x = tv[m];
          atoms.push_back(x.atom1); atoms.push_back(x.atom2); atoms.push_back(x.atom3); atoms.push_back(x.atom4);
          for (int u = 0; u < x.multiplicity; u++) {
            k.push_back(x.forceConstant[u]);
            per.push_back(x.periodicity[u]);
            ph.push_back(x.phaseShift[u]);
          }
          row.push_back(static_cast<int>(k.size()));
        }
        r = pm_set_torsions(ctx, kind, static_cast<int>(tv.size()), &atoms[0], &row[0], k.empty() ? 0 : &k[0],
                            per.empty() ? 0 : &per[0], ph.empty() ? 0 : &ph[0]);
        if (r != PM_OK) return r;
      }
      // PeriodicBoundaryConditions::set (topology/PeriodicBoundaryConditions.cpp:34-77) and the cell size
      const Vector3D e1 = t->boundaryConditions.e1(), e2 = t->boundaryConditions.e2(),
                     e3 = t->boundaryConditions.e3(), o = t->boundaryConditions.origin();
      const double v1[3] = {e1.x, e1.y, e1.z}, v2[3] = {e2.x, e2.y, e2.z}, v3[3] = {e3.x, e3.y, e3.z},
                   vo[3] = {o.x, o.y, o.z};
      return pm_set_boundary(ctx, v1, v2, v3, vo, t->cellManager.getCellSize());
    }
  }

  const std::string& gpuLastError() { return state().err; }

  void gpuSetPrecision(int precision) {
    GpuState& s = state();
    s.precision = precision;
    if (s.ctx) pm_set_precision(s.ctx, precision);
  }

  void gpuReset() {
    GpuState& s = state();
    if (s.ctx) pm_destroy(s.ctx);
    s = GpuState();
  }

  pm_context* gpuContext(const GenericTopology* topo) {
    GpuState& s = state();
    if (!s.ctx) {
      if (s.tried) return 0;
      s.tried = true;
      if (pm_create(0, &s.ctx) != PM_OK) {
        s.ctx = 0;
        s.err = "no HIP device for the GPU forces (pm_create failed)";
        return 0;
      }
      pm_set_precision(s.ctx, s.precision);
    }
    if (topo && (topo != s.topo || topo->atoms.size() != s.natoms)) {
      const PeriodicTopology* t = dynamic_cast<const PeriodicTopology*>(topo);
      if (!t) {
        s.err = "the GPU forces need a periodic topology with a cubic cell manager";
        return 0;
      }
      if (upload(s.ctx, t) != PM_OK) {
        s.err = pm_last_error(s.ctx);
        return 0;
      }
      s.topo = topo;
      s.natoms = topo->atoms.size();
    }
    return s.ctx;
  }

  GpuSystemForce::GpuSystemForce(const Force* cpu, int handle) : myCpu(cpu), myHandle(handle) {}

  void GpuSystemForce::evaluate(const GenericTopology* topo, const Vector3DBlock* positions, Vector3DBlock* forces,
                                ScalarStructure* energies) {
    pm_context* ctx = gpuContext(topo);
    if (!ctx) {
      report << error << "GPU force " << getId() << ": " << gpuLastError() << endr;
      return;
    }
    const unsigned int n = positions->size();
    myForces.assign(3 * n, 0.0);
    double e[PM_ENERGY_COUNT] = {0};
    const int flags = (energies->virial() ? PM_EVAL_VIRIAL : 0) | (energies->molecularVirial() ? PM_EVAL_MOLVIRIAL : 0);
    // Vector3DBlock holds Vector3D{x, y, z} contiguously: the AoS double[3n] of the ABI
    if (pm_evaluate(ctx, &myHandle, 1, &(*positions)[0].x, &myForces[0], e, flags) != PM_OK) {
      report << error << "GPU force " << getId() << ": " << pm_last_error(ctx) << endr;
      return;
    }
    for (unsigned int i = 0; i < n; i++) {
      (*forces)[i].x += myForces[3 * i];
      (*forces)[i].y += myForces[3 * i + 1];
      (*forces)[i].z += myForces[3 * i + 2];
    }
    // energies[0..15] is the ScalarStructure prefix COULOMB .. VIRIALZZ
    for (int k = 0; k < PM_ENERGY_COUNT; k++) (*energies)[static_cast<ScalarStructure::Index>(k)] += e[k];
  }

  // One process evaluates the whole force (the device holds every atom).
  void GpuSystemForce::parallelEvaluate(const GenericTopology* topo, const Vector3DBlock* positions,
                                        Vector3DBlock* forces, ScalarStructure* energies) {
    evaluate(topo, positions, forces, energies);
  }

  void GpuSystemForce::uncache() {
    pm_context* ctx = gpuContext(0);
    if (ctx && myHandle >= 0) pm_force_uncache(ctx, myHandle);
  }

  std::string GpuSystemForce::getKeyword() const { return myCpu->getKeyword(); }
  std::string GpuSystemForce::getIdNoAlias() const { return myCpu->getIdNoAlias(); }
  void GpuSystemForce::getParameters(std::vector<Parameter>& parameters) const { myCpu->getParameters(parameters); }
  unsigned int GpuSystemForce::getParameterSize() const { return myCpu->getParameterSize(); }
  bool GpuSystemForce::adjustWithDefaultParameters(std::vector<Value>& values, const Configuration* config) const {
    return myCpu->adjustWithDefaultParameters(values, config);
  }

  // Force::make has checked the values against the exemplar's parameters
  // (forces/Force.cpp:15-21); the library repeats the force's own checks and
  // answers with the reference's messages.
  Force* GpuSystemForce::doMake(std::string& errMsg, std::vector<Value> values) const {
    std::vector<double> v;
    for (unsigned int i = 0; i < values.size(); i++) {
      Real x = 0;
      values[i].get(x);
      v.push_back(x);
    }
    pm_context* ctx = gpuContext(0);
    if (!ctx) {
      errMsg += gpuLastError();
      return 0;
    }
    int h = -1;
    if (pm_force_create(ctx, getIdNoAlias().c_str(), v.empty() ? 0 : &v[0], static_cast<int>(v.size()), &h) != PM_OK) {
      errMsg += pm_last_error(ctx);
      return 0;
    }
    GpuSystemForce* f = new GpuSystemForce(myCpu, h);
    f->myValues = v;
    return f;
  }

  int registerGpuForceExemplars() {
    // the periodic exemplars of the GPU path
    // (factories/registerForceExemplarsCutoff.cpp:39,40,45,57,62,63,67,76,88,90,91,131,
    //  factories/registerForceExemplarsFastElectrostatic.cpp:50-55)
    static const char* ids[] = {
      "Coulomb -switchingFunction C1 -algorithm NonbondedCutoff",
      "Coulomb -switchingFunction C2 -algorithm NonbondedCutoff",
      "Coulomb -switchingFunction Cutoff -algorithm NonbondedCutoff",
      "CoulombEwaldReal -switchingFunction Cutoff -algorithm NonbondedCutoff",
      "LennardJones -switchingFunction C1 -algorithm NonbondedCutoff",
      "LennardJones -switchingFunction C2 -algorithm NonbondedCutoff",
      "LennardJones -switchingFunction Cutoff -algorithm NonbondedCutoff",
      "LennardJones CoulombEwaldReal -switchingFunction C1 -switchingFunction Cutoff -algorithm NonbondedCutoff",
      "LennardJones Coulomb -switchingFunction C2 -switchingFunction C1 -algorithm NonbondedCutoff",
      "LennardJones CoulombEwaldReal -switchingFunction C2 -switchingFunction Cutoff -algorithm NonbondedCutoff",
      "LennardJones Coulomb -switchingFunction C2 -switchingFunction Cutoff -algorithm NonbondedCutoff",
      "LennardJones CoulombEwaldReal -switchingFunction Cutoff -switchingFunction Cutoff -algorithm NonbondedCutoff",
      "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline",
      "Coulomb -algorithm PMEwald -real -interpolation BSpline",
      "Coulomb -algorithm PMEwald -real -correction -interpolation BSpline",
      "Coulomb -algorithm PMEwald -reciprocal -interpolation BSpline",
      "Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline",
      "Coulomb -algorithm PMEwald -correction -interpolation BSpline",
      0};
    int count = 0;
    for (int k = 0; ids[k]; k++) {
      const Force* cpu = ForceFactory::find(ids[k]);
      if (!cpu || dynamic_cast<const GpuSystemForce*>(cpu)) continue;
      std::vector<std::string> aliases;
      if (std::string(ids[k]) == "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline")
        aliases.push_back("CoulombPME");
      ForceFactory::registerExemplar(new GpuSystemForce(cpu), aliases, ids[k]);
      count++;
    }
    return count;
  }
}
