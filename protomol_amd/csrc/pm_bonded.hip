// pm_bonded.hip -- FP64 bonded system forces on gfx950: bonds, angles,
// dihedrals and impropers.
//
// BondSystemForce::calcBond (forces/BondSystemForce.h:111-145),
// AngleSystemForce::calcAngle (forces/AngleSystemForce.h:113-200) and
// MTorsionSystemForce::calcTorsion (forces/MTorsionSystemForce.h:34-153) for
// DihedralSystemForce and ImproperSystemForce.  Forces
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

// per-term energies and virials; records: bond, angle, dihedral and
// improper blocks in that order (kind of block bx from the prefix off[])
struct BondedBlocks { int off[5]; };
__device__ inline void bonded_energy_body(PmBox box, PmBondedLaunch a, BondedBlocks bl, int bx) {
  __shared__ double s_red[4][7];
  int kind = 0;
  while (kind < 3 && bx >= bl.off[kind + 1]) kind++;
  const int t = (bx - bl.off[kind]) * blockDim.x + threadIdx.x;
  const int count = kind == 0 ? a.nb : kind == 1 ? a.na : a.nt[kind - 2];
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (t < count) {
    double f1[3], f2[3] = {0, 0, 0}, f3[3] = {0, 0, 0}, e, r12[3], r23[3] = {0, 0, 0}, r34[3] = {0, 0, 0};
    if (kind == 0) {
      bond_term(box, a, t, f1, e, r12);
    } else if (kind == 1) {
      // the angle's virial is f1 r12 + f3 r32 (AngleSystemForce.h:180-198)
      angle_term(box, a, t, f1, f2, f3, e, r12, r34);
      f2[0] = f2[1] = f2[2] = 0.0;
    } else {
      torsion_term(box, a, kind - 2, t, f1, f2, f3, e, r12, r23, r34);
    }
    v[0] = e;
    if (a.want_virial) {
      v[1] = f1[0] * r12[0] + f2[0] * r23[0] + f3[0] * r34[0];
      v[2] = f1[0] * r12[1] + f2[0] * r23[1] + f3[0] * r34[1];
      v[3] = f1[0] * r12[2] + f2[0] * r23[2] + f3[0] * r34[2];
      v[4] = f1[1] * r12[1] + f2[1] * r23[1] + f3[1] * r34[1];
      v[5] = f1[1] * r12[2] + f2[1] * r23[2] + f3[1] * r34[2];
      v[6] = f1[2] * r12[2] + f2[2] * r23[2] + f3[2] * r34[2];
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
    const int slot[4] = {PM_E_BOND, PM_E_ANGLE, PM_E_DIHEDRAL, PM_E_IMPROPER};
    ep[slot[kind]] = s[0];
    ep[PM_E_VIR + 0] = s[1]; ep[PM_E_VIR + 1] = s[2]; ep[PM_E_VIR + 2] = s[3];
    ep[PM_E_VIR + 3] = s[2]; ep[PM_E_VIR + 4] = s[4]; ep[PM_E_VIR + 5] = s[5];
    ep[PM_E_VIR + 6] = s[3]; ep[PM_E_VIR + 7] = s[5]; ep[PM_E_VIR + 8] = s[6];
  }
}
__global__ __launch_bounds__(256) void k_bonded_energy(PmBox box, PmBondedLaunch a, BondedBlocks bl) {
  bonded_energy_body(box, a, bl, blockIdx.x);
}
struct BondedEnergyArgs { PmBox box; PmBondedLaunch a; BondedBlocks bl; };
__global__ __launch_bounds__(256) void kb_bonded_energy(const BondedEnergyArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const BondedEnergyArgs& q = t[r];
  bonded_energy_body(q.box, q.a, q.bl, blockIdx.x - pre[r]);
}

bool any_bonded(const PmBondedLaunch& a) { return a.do_bond || a.do_angle || a.do_tors[0] || a.do_tors[1]; }

}  // namespace

void pm_bonded_blocks(const PmBondedLaunch& a, int nblk[4]) {
  nblk[0] = a.do_bond ? (a.nb + 255) / 256 : 0;
  nblk[1] = a.do_angle ? (a.na + 255) / 256 : 0;
  nblk[2] = a.do_tors[0] ? (a.nt[0] + 255) / 256 : 0;
  nblk[3] = a.do_tors[1] ? (a.nt[1] + 255) / 256 : 0;
}

void pm_launch_bonded(hipStream_t s, const PmBox& box, const PmBondedLaunch& a) {
  int nblk[4];
  pm_bonded_blocks(a, nblk);
  BondedBlocks bl;
  bl.off[0] = 0;
  for (int k = 0; k < 4; k++) bl.off[k + 1] = bl.off[k] + nblk[k];
  const int total = bl.off[4];
  const bool forces = a.n > 0 && any_bonded(a);
  if (pm_recording()) {
    const BondedArgs b{box, a};
    pm_record(PM_REC_BONDED_FORCE, 0, forces ? (a.n + 255) / 256 : 0, 256, 0, &b, sizeof(b));
    if (a.want_energy) {
      const BondedEnergyArgs e{box, a, bl};
      pm_record(PM_REC_BONDED_ENERGY, 0, (unsigned)total, 256, 0, &e, sizeof(e));
    }
    return;
  }
  if (forces) hipLaunchKernelGGL(k_bonded_force, dim3((a.n + 255) / 256), dim3(256), 0, s, box, a);
  if (total > 0 && a.want_energy) hipLaunchKernelGGL(k_bonded_energy, dim3(total), dim3(256), 0, s, box, a, bl);
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
