// GpuSystemForce.h -- ProtoMol SystemForce adapter over the C ABI of
// libpmgpu (include/pmgpu.h).
//
// Compiled inside a ProtoMol build (it includes the framework's own headers):
// each adapter stands in for one CPU exemplar of the periodic cell-list
// cutoff or PME path, keeps that exemplar's keyword, id and parameter list
// (so Force::make checks the same parameters and a .conf line selects it
// unchanged), and forwards
//   Force::make -> doMake          to pm_force_create   (forces/Force.cpp:15-21)
//   SystemForce::evaluate          to pm_evaluate       (forces/SystemForce.h:39)
//   Force::uncache                 to pm_force_uncache  (forces/Force.h:65)
// ForceGroup::evaluateSystemForces (forces/ForceGroup.cpp:43-96) and every
// integrator then pick the GPU path up with no other change.
#ifndef PROTOMOL_GPU_SYSTEM_FORCE_H
#define PROTOMOL_GPU_SYSTEM_FORCE_H

#include "SystemForce.h"
#include "pmgpu.h"

#include <string>
#include <vector>

namespace ProtoMol {
  class GenericTopology;

  /// The library context of the process (device 0), with the topology of
  /// `topo` uploaded whenever the topology object or its atom count changes
  /// (topo may be NULL: the context only).  NULL when no device is usable;
  /// gpuLastError() then says why.
  pm_context* gpuContext(const GenericTopology* topo);
  const std::string& gpuLastError();
  /// PM_PRECISION_MIXED (default) or PM_PRECISION_DOUBLE for every GPU force.
  void gpuSetPrecision(int precision);
  /// Forget the context (and its forces): the next gpuContext creates a new one.
  void gpuReset();

  class GpuSystemForce : public SystemForce {
  public:
    /// cpu: the exemplar this force replaces (it keeps its keyword, id and
    /// parameter list); handle: the library's force, -1 for the exemplar.
    explicit GpuSystemForce(const Force* cpu, int handle = -1);

    virtual void evaluate(const GenericTopology* topo, const Vector3DBlock* positions, Vector3DBlock* forces,
                          ScalarStructure* energies);
    virtual void parallelEvaluate(const GenericTopology* topo, const Vector3DBlock* positions,
                                  Vector3DBlock* forces, ScalarStructure* energies);
    virtual void uncache();

    virtual std::string getKeyword() const;
    virtual std::string getIdNoAlias() const;
    virtual void getParameters(std::vector<Parameter>& parameters) const;
    virtual unsigned int getParameterSize() const;
    virtual bool adjustWithDefaultParameters(std::vector<Value>& values, const Configuration* config) const;

    int handle() const { return myHandle; }
    const Force* cpuExemplar() const { return myCpu; }

  private:
    virtual Force* doMake(std::string& errMsg, std::vector<Value> values) const;

    const Force* myCpu;
    int myHandle;
    std::vector<double> myValues;
    mutable std::vector<double> myForces;
  };

  /// Register one adapter over each CPU exemplar of the GPU path that this
  /// ProtoMol build registers (call after registerForceExemplarsCutoff and
  /// registerForceExemplarsFastElectrostatic: Factory::registerExemplar
  /// overwrites an id that is already registered, frontend/Factory.h:153-166).
  /// Returns the number of ids now served by the GPU.
  int registerGpuForceExemplars();
}

#endif /* PROTOMOL_GPU_SYSTEM_FORCE_H */
