// pm_bonded.hip -- FP64 bond and angle system forces on gfx950.
//
// BondSystemForce::calcBond (forces/BondSystemForce.h:111-145) and
// AngleSystemForce::calcAngle (forces/AngleSystemForce.h:113-200).  Forces
// are gathered per atom (each atom evaluates the terms it belongs to and keeps
// its own share), so no atomics are needed and the result is deterministic;
// energies and virials are summed per term in fixed-order block reductions.
#include "pm_internal.h"
#include "pm_terms.h"

#include <math.h>

namespace {

__device__ inline void bonded_force_body(const PmBox& box, const PmBondedLaunch& a, int bx) {
  const int i = bx * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double f[3];
  bonded_atom_force(box, a, i, f);
  a.force_fix[3 * i] += __double2ll_rn(f[0] * PM_FORCE_FIX);
  a.force_fix[3 * i + 1] += __double2ll_rn(f[1] * PM_FORCE_FIX);
  a.force_fix[3 * i + 2] += __double2ll_rn(f[2] * PM_FORCE_FIX);
}
__global__ void k_bonded_force(PmBox box, PmBondedLaunch a) { bonded_force_body(box, a, blockIdx.x); }

struct BondedArgs { PmBox box; PmBondedLaunch a; };
__global__ void kb_bonded_force(const BondedArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  bonded_force_body(t[r].box, t[r].a, blockIdx.x - pre[r]);
}

// per-term energies and virials; records: bond blocks first, then angle blocks
__device__ inline void bonded_energy_body(PmBox box, PmBondedLaunch a, int nbb, int bx) {
  __shared__ double s_red[4][7];
  const bool is_bond = (int)bx < nbb;
  const int t = (is_bond ? bx : bx - nbb) * blockDim.x + threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (is_bond && t < a.nb) {
    double f1[3], e, r12[3];
    bond_term(box, a, t, f1, e, r12);
    v[0] = e;
    if (a.want_virial) {
      v[1] = f1[0] * r12[0]; v[2] = f1[0] * r12[1]; v[3] = f1[0] * r12[2];
      v[4] = f1[1] * r12[1]; v[5] = f1[1] * r12[2]; v[6] = f1[2] * r12[2];
    }
  } else if (!is_bond && t < a.na) {
    double f1[3], f2[3], f3[3], e, r12[3], r32[3];
    angle_term(box, a, t, f1, f2, f3, e, r12, r32);
    v[0] = e;
    if (a.want_virial) {
      v[1] = f1[0] * r12[0] + f3[0] * r32[0];
      v[2] = f1[0] * r12[1] + f3[0] * r32[1];
      v[3] = f1[0] * r12[2] + f3[0] * r32[2];
      v[4] = f1[1] * r12[1] + f3[1] * r32[1];
      v[5] = f1[1] * r12[2] + f3[1] * r32[2];
      v[6] = f1[2] * r12[2] + f3[2] * r32[2];
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int u = 0; u < 7; u++) {
    double x = v[u];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) s_red[wv][u] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s[7];
    for (int u = 0; u < 7; u++) s[u] = s_red[0][u] + s_red[1][u] + s_red[2][u] + s_red[3][u];
    double* ep = a.epart + (long long)bx * PM_NE;
    for (int u = 0; u < PM_E_VIR; u++) ep[u] = 0.0;
    ep[is_bond ? PM_E_BOND : PM_E_ANGLE] = s[0];
    ep[PM_E_VIR + 0] = s[1]; ep[PM_E_VIR + 1] = s[2]; ep[PM_E_VIR + 2] = s[3];
    ep[PM_E_VIR + 3] = s[2]; ep[PM_E_VIR + 4] = s[4]; ep[PM_E_VIR + 5] = s[5];
    ep[PM_E_VIR + 6] = s[3]; ep[PM_E_VIR + 7] = s[5]; ep[PM_E_VIR + 8] = s[6];
  }
}
__global__ __launch_bounds__(256) void k_bonded_energy(PmBox box, PmBondedLaunch a, int nbb) { bonded_energy_body(box, a, nbb, blockIdx.x); }
struct BondedEnergyArgs { PmBox box; PmBondedLaunch a; int nbb; };
__global__ __launch_bounds__(256) void kb_bonded_energy(const BondedEnergyArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const BondedEnergyArgs& q = t[r];
  bonded_energy_body(q.box, q.a, q.nbb, blockIdx.x - pre[r]);
}


}  // namespace

void pm_launch_bonded(hipStream_t s, const PmBox& box, const PmBondedLaunch& a, int* nrec_bond, int* nrec_angle) {
  const int nbb = a.do_bond ? (a.nb + 255) / 256 : 0;
  const int nab = a.do_angle ? (a.na + 255) / 256 : 0;
  *nrec_bond = nbb;
  *nrec_angle = nab;
  if (pm_recording()) {
    const BondedArgs b{box, a};
    pm_record(PM_REC_BONDED_FORCE, 0, a.n > 0 && (a.do_bond || a.do_angle) ? (a.n + 255) / 256 : 0, 256, 0, &b,
              sizeof(b));
    if (a.want_energy) {
      const BondedEnergyArgs e{box, a, nbb};
      pm_record(PM_REC_BONDED_ENERGY, 0, (unsigned)(nbb + nab), 256, 0, &e, sizeof(e));
    }
    return;
  }
  if (a.n > 0 && (a.do_bond || a.do_angle))
    hipLaunchKernelGGL(k_bonded_force, dim3((a.n + 255) / 256), dim3(256), 0, s, box, a);
  if (nbb + nab > 0 && a.want_energy)
    hipLaunchKernelGGL(k_bonded_energy, dim3(nbb + nab), dim3(256), 0, s, box, a, nbb);
}

bool pm_launch_batched_bonded(hipStream_t s, int kind, int variant, const void* t, const int* pre, int nrep,
                              unsigned total, unsigned threads, unsigned lds) {
  (void)variant; (void)lds;
  if (kind == PM_REC_BONDED_FORCE) {
    hipLaunchKernelGGL(kb_bonded_force, dim3(total), dim3(threads), 0, s, (const BondedArgs*)t, pre, nrep);
    return true;
  }
  if (kind == PM_REC_BONDED_ENERGY) {
    hipLaunchKernelGGL(kb_bonded_energy, dim3(total), dim3(threads), 0, s, (const BondedEnergyArgs*)t, pre, nrep);
    return true;
  }
  return false;
}
