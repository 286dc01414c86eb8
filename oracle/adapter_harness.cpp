// adapter_harness.cpp -- C entry points that drive the GpuSystemForce adapter
// (protomol_amd/adapter/) through ProtoMol's own factory and force group.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `adapter`) compiles this
// file and the adapter with the reference's headers, links them with the
// reference's objects (base/, topology/, forces/, the periodic cutoff and
// fast-electrostatics exemplar registrations, the force factory) and with
// libpmgpu.so into oracle/_ref/libpmadapter.so.  A test then checks, on the
// reference's side of the boundary, that:
//   * registerForceExemplarsCutoff + registerForceExemplarsFastElectrostatic
//     followed by registerGpuForceExemplars serve the GPU path's ids with
//     adapters that keep the CPU exemplars' ids and parameter lists;
//   * ForceFactory::make reports bad parameters the reference's way;
//   * ForceGroup::evaluateSystemForces over adapters gives the CPU
//     exemplars' forces and energies.
#include "Report.h"
using namespace ProtoMol::Report;
#include "Topology.h"
#include "PeriodicBoundaryConditions.h"
#include "CubicCellManager.h"
#include "ForceFactory.h"
#include "ForceFactoryDetails.h"
#include "ForceGroup.h"
#include "SystemForce.h"
#include "ScalarStructure.h"
#include "Vector3DBlock.h"
#include "registerForceExemplarsCutoff.h"
#include "registerForceExemplarsFastElectrostatic.h"
#include "GpuSystemForce.h"

#include <cstring>
#include <string>
#include <vector>

using namespace ProtoMol;

namespace {

typedef Topology<PeriodicBoundaryConditions, CubicCellManager> Topo;

void copyText(const std::string& s, char* buf, int cap) {
  if (!buf || cap <= 0) return;
  std::strncpy(buf, s.c_str(), cap - 1);
  buf[cap - 1] = 0;
}

// Values for exemplar `proto` from plain numbers, typed as its parameters are.
bool makeValues(const Force* proto, const double* v, int n, std::vector<Value>& out) {
  std::vector<Parameter> p;
  proto->getParameters(p);
  if ((int)p.size() != n) return false;
  out.clear();
  for (int i = 0; i < n; i++) {
    Value x(p[i].value);
    if (p[i].value.getType() == ValueType::INT) x.set(static_cast<int>(v[i]));
    else if (p[i].value.getType() == ValueType::UINT) x.set(static_cast<unsigned int>(v[i]));
    else x.set(static_cast<Real>(v[i]));
    out.push_back(x);
  }
  return true;
}

Topo* makeTopo(int n, const double* q, const int* type, int ntypes, const double* ljA, const double* ljB,
               const double* ljA14, const double* ljB14, const int* molecule, const double* mass, int nexcl,
               const int* ea, const int* eb, const int* ec, double csf, const double* box, const double* origin,
               double cellsize) {
  // the members frontend/buildTopology.cpp fills from PSF/PAR data
  Topo* t = new Topo(csf, ExclusionType(ExclusionType::ONE4MODIFIED),
                     PeriodicBoundaryConditions(Vector3D(box[0], 0, 0), Vector3D(0, box[1], 0),
                                                Vector3D(0, 0, box[2]), Vector3D(origin[0], origin[1], origin[2])),
                     CubicCellManager(cellsize));
  int nmol = 0;
  for (int i = 0; i < n; i++) nmol = std::max(nmol, molecule[i] + 1);
  t->atomTypes.resize(ntypes);
  t->atoms.resize(n);
  t->molecules.resize(nmol);
  for (int i = 0; i < n; i++) {
    t->atoms[i].type = type[i];
    t->atoms[i].scaledCharge = q[i];
    t->atoms[i].scaledMass = mass[i];
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
  return t;
}

}  // namespace

extern "C" {

// The CPU exemplars of the periodic cutoff and PME paths, then (gpu != 0)
// the adapters over them.  Returns the number of adapters registered.
int adp_register(int gpu) {
  const PeriodicBoundaryConditions* bc = 0;
  const CubicCellManager* cm = 0;
  registerForceExemplarsCutoff(bc, cm);
  registerForceExemplarsFastElectrostatic(bc, cm);
  return gpu ? registerGpuForceExemplars() : 0;
}

// What the factory holds under `id`: 1 if found; *is_gpu whether it is an
// adapter, *nparams its parameter count, idbuf its getIdNoAlias(), and
// cpu_id / cpu_nparams those of the CPU exemplar it stands in for.
int adp_describe(const char* id, int* is_gpu, int* nparams, char* idbuf, int cap, char* cpu_id, int cpu_cap,
                 int* cpu_nparams, char* keyword, int kcap) {
  const Force* f = ForceFactory::find(id);
  if (!f) return 0;
  const GpuSystemForce* g = dynamic_cast<const GpuSystemForce*>(f);
  *is_gpu = g ? 1 : 0;
  *nparams = (int)f->getParameterSize();
  copyText(f->getIdNoAlias(), idbuf, cap);
  copyText(f->getKeyword(), keyword, kcap);
  const Force* cpu = g ? g->cpuExemplar() : f;
  copyText(cpu->getIdNoAlias(), cpu_id, cpu_cap);
  *cpu_nparams = (int)cpu->getParameterSize();
  return 1;
}

// ForceFactory::make(id, values): 1 and the made force's kind (1 adapter,
// 0 CPU) on success; 0 and the factory's errMsg on a NULL result.
int adp_make(const char* id, const double* values, int n, int* is_gpu, char* err, int cap) {
  const Force* proto = ForceFactory::find(id);
  std::vector<Value> v;
  if (!proto || !makeValues(proto, values, n, v)) {
    copyText("no exemplar with that id and parameter count", err, cap);
    return 0;
  }
  std::string errMsg;
  Force* f = ForceFactory::make(errMsg, id, v);
  if (!f) {
    copyText(errMsg, err, cap);
    return 0;
  }
  *is_gpu = dynamic_cast<GpuSystemForce*>(f) ? 1 : 0;
  delete f;
  return 1;
}

// One ForceGroup::evaluateSystemForces (forces/ForceGroup.cpp:43-96) over
// forces made by ForceFactory::make from (ids[k], values[off[k] .. off[k+1])),
// on a Topology<PeriodicBoundaryConditions, CubicCellManager> built from the
// arrays; forces[3n] and energies[16] (the ScalarStructure prefix up to the
// virial) are overwritten.  Returns 0, or -1 with errMsg in err.
int adp_evaluate(int n, const double* pos, const double* q, const int* type, int ntypes, const double* ljA,
                 const double* ljB, const double* ljA14, const double* ljB14, const int* molecule, const double* mass,
                 int nexcl, const int* ea, const int* eb, const int* ec, double csf, const double* box,
                 const double* origin, double cellsize, int nforces, const char** ids, const int* off,
                 const double* values, int virial, double* forces, double* energies, int* gpu_forces, char* err,
                 int cap) {
  Topo* t = makeTopo(n, q, type, ntypes, ljA, ljB, ljA14, ljB14, molecule, mass, nexcl, ea, eb, ec, csf, box,
                     origin, cellsize);
  ForceGroup* group = new ForceGroup();
  int ngpu = 0;
  for (int k = 0; k < nforces; k++) {
    const Force* proto = ForceFactory::find(ids[k]);
    std::vector<Value> v;
    if (!proto || !makeValues(proto, values + off[k], off[k + 1] - off[k], v)) {
      copyText(std::string("no exemplar for ") + ids[k], err, cap);
      delete group;
      delete t;
      return -1;
    }
    std::string errMsg;
    Force* f = ForceFactory::make(errMsg, ids[k], v);
    if (!f) {
      copyText(errMsg, err, cap);
      delete group;
      delete t;
      return -1;
    }
    if (dynamic_cast<GpuSystemForce*>(f)) ngpu++;
    f->addToForceGroup(group);
  }
  Vector3DBlock x, f;
  x.resize(n);
  f.resize(n);
  for (int i = 0; i < n; i++) {
    x[i] = Vector3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    f[i] = Vector3D(0, 0, 0);
  }
  ScalarStructure e;
  e.clear();
  // as the applications set both flags from virialCalc / molVirialCalc
  // (applications/rem-app/remMT.cpp:479-480); ScalarStructure's constructor
  // leaves the molecular virial on, which the GPU forces refuse
  e.virial(virial != 0);
  e.molecularVirial(false);
  group->evaluateSystemForces(t, &x, &f, &e);
  for (int i = 0; i < n; i++) {
    forces[3 * i] = f[i].x;
    forces[3 * i + 1] = f[i].y;
    forces[3 * i + 2] = f[i].z;
  }
  for (int k = 0; k < 16; k++) energies[k] = e[static_cast<ScalarStructure::Index>(k)];
  if (gpu_forces) *gpu_forces = ngpu;
  delete group;
  delete t;
  return 0;
}

// Drop the adapters' library context (between tests).
void adp_reset() { gpuReset(); }

// Arithmetic of every GPU force made afterwards (PM_PRECISION_MIXED / _DOUBLE).
void adp_set_precision(int precision) { gpuSetPrecision(precision); }

}  // extern "C"
