// pm_bonded.hip -- FP64 bond and angle system forces on gfx950.
//
// BondSystemForce::calcBond (forces/BondSystemForce.h:111-145) and
// AngleSystemForce::calcAngle (forces/AngleSystemForce.h:113-200).  Forces
// are gathered per atom (each atom evaluates the terms it belongs to and keeps
// its own share), so no atomics are needed and the result is deterministic;
// energies and virials are summed per term in fixed-order block reductions.
#include "pm_internal.h"

#include <math.h>

namespace {

__device__ inline void min_diff(const PmBox& b, const double* c1, const double* c2, double* d) {
#pragma clang fp contract(off)
  double x = c2[0] - c1[0], y = c2[1] - c1[1], z = c2[2] - c1[2];
  double s = x * x + y * y + z * z;
  if (s > b.d) {
    if (x < -b.h2[0]) x += b.h[0]; else if (x > b.h2[0]) x -= b.h[0];
    if (y < -b.h2[1]) y += b.h[1]; else if (y > b.h2[1]) y -= b.h[1];
    if (z < -b.h2[2]) z += b.h[2]; else if (z > b.h2[2]) z -= b.h[2];
    s = x * x + y * y + z * z;
    if (s > b.d) {
      x -= b.L[0] * __builtin_rint(b.er[0] * x);
      y -= b.L[1] * __builtin_rint(b.er[1] * y);
      z -= b.L[2] * __builtin_rint(b.er[2] * z);
    }
  }
  d[0] = x; d[1] = y; d[2] = z;
}

__device__ inline double norm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// one bond: force on atom1 (atom2 gets the opposite), energy, r12
__device__ inline void bond_term(const PmBox& b, const PmBondedLaunch& a, int k, double* f1, double& e,
                                 double* r12) {
  const int a1 = a.b_a[k], a2 = a.b_b[k];
  const double r0 = a.b_r0[k], kb = a.b_k[k];
  min_diff(b, a.pos + 3 * a2, a.pos + 3 * a1, r12);   // vector from atom 2 to atom 1
  const double r = norm3(r12);
  const double dpotdr = 2.0 * kb * (r - r0);
  const double s = -dpotdr / r;
  f1[0] = r12[0] * s; f1[1] = r12[1] * s; f1[2] = r12[2] * s;
  e = kb * (r - r0) * (r - r0);
}

// one angle (a2 is the vertex): forces on the three atoms, energy, r12/r32
__device__ inline void angle_term(const PmBox& b, const PmBondedLaunch& a, int k, double* f1, double* f2,
                                  double* f3, double& e, double* r12, double* r32) {
  const int a1 = a.a_a[k], a2 = a.a_b[k], a3 = a.a_c[k];
  const double t0 = a.a_t0[k], kt = a.a_k[k], kub = a.a_kub[k], rub = a.a_rub[k];
  double r13[3];
  min_diff(b, a.pos + 3 * a2, a.pos + 3 * a1, r12);
  min_diff(b, a.pos + 3 * a2, a.pos + 3 * a3, r32);
  min_diff(b, a.pos + 3 * a3, a.pos + 3 * a1, r13);
  const double d12 = norm3(r12), d32 = norm3(r32), d13 = norm3(r13);
  double cr[3] = {r12[1] * r32[2] - r12[2] * r32[1], r12[2] * r32[0] - r12[0] * r32[2],
                  r12[0] * r32[1] - r12[1] * r32[0]};
  const double theta = atan2(norm3(cr), r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2]);
  const double st = sin(theta), ct = cos(theta);
  const double dpot = 2.0 * kt * (theta - t0);
  double u1[3], u3[3], ub1[3];
  for (int c = 0; c < 3; c++) {
    const double dr12 = r12[c] / d12, dr32 = r32[c] / d32, dr13 = r13[c] / d13;
    const double dth1 = (dr12 * ct - dr32) / (st * d12);
    const double dth3 = (dr32 * ct - dr12) / (st * d32);
    ub1[c] = dr13 * (2.0 * kub * (d13 - rub));
    u1[c] = -dth1 * dpot - ub1[c];
    u3[c] = -dth3 * dpot + ub1[c];
  }
  for (int c = 0; c < 3; c++) { f1[c] = u1[c]; f3[c] = u3[c]; f2[c] = -u1[c] - u3[c]; }
  e = kt * (theta - t0) * (theta - t0) + kub * (d13 - rub) * (d13 - rub);
}

__device__ inline void bonded_force_body(const PmBox& box, const PmBondedLaunch& a, int bx) {
  const int i = bx * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double f[3] = {0, 0, 0};
  if (a.do_bond) {
    for (int u = a.b_row[i]; u < a.b_row[i + 1]; u++) {
      const int k = a.b_idx[u];
      double f1[3], e, r12[3];
      bond_term(box, a, k, f1, e, r12);
      const double sg = (a.b_a[k] == i) ? 1.0 : -1.0;
      f[0] += sg * f1[0]; f[1] += sg * f1[1]; f[2] += sg * f1[2];
    }
  }
  if (a.do_angle) {
    for (int u = a.a_row[i]; u < a.a_row[i + 1]; u++) {
      const int k = a.a_idx[u];
      double f1[3], f2[3], f3[3], e, r12[3], r32[3];
      angle_term(box, a, k, f1, f2, f3, e, r12, r32);
      const double* g = (a.a_a[k] == i) ? f1 : (a.a_b[k] == i) ? f2 : f3;
      f[0] += g[0]; f[1] += g[1]; f[2] += g[2];
    }
  }
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
