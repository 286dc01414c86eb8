// pm_shake.hip -- SHAKE and RATTLE bond constraints (FP64) on gfx950.
//
// ModifierShake::doExecute (integrators/ModifierShake.cpp:21-103) and
// ModifierRattle::doExecute (integrators/ModifierRattle.cpp:20-85) with the
// errors of ModifierMetaShake::calcError (integrators/ModifierMetaShake.cpp:12-31)
// and ModifierMetaRattle::calcError (integrators/ModifierMetaRattle.cpp:12-31):
// while the mean error over all constraints exceeds epsilon, sweep over every
// constraint in list order (at most maxIter + 1 sweeps).
//
// The constraints of one connected group of atoms touch no other group's
// atoms, so a sweep over the whole list is, group by group, a sweep over each
// group's constraints in their list order: one thread per group repeats the
// reference's FP64 operations (no contraction) exactly.  Only the number of
// sweeps is global.  A probe kernel records every group's error after 0 ..
// maxIter sweeps (on a scratch copy of the positions or velocities); one
// block per sweep count sums the groups' errors in a fixed order into the mean
// error; the apply kernel's threads find the first sweep count whose mean is
// within epsilon and run that many sweeps on the real positions and
// velocities.
#include "pm_internal.h"

#include <math.h>

namespace {

__device__ inline void ld3(const double* p, int a, double* o) {
  o[0] = p[3 * a]; o[1] = p[3 * a + 1]; o[2] = p[3 * a + 2];
}
__device__ inline void st3(double* p, int a, const double* o) {
  p[3 * a] = o[0]; p[3 * a + 1] = o[1]; p[3 * a + 2] = o[2];
}
__device__ inline double dot3(const double* a, const double* b) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// ModifierMetaShake::calcError term of one constraint:
// |(x2 - x1)^2 - r0^2| / (2 r0^2)
__device__ inline double shake_err(const PmConstraints& c, const double* x, int k) {
#pragma clang fp contract(off)
  const int a1 = c.a1[k], a2 = c.a2[k];
  double p1[3], p2[3];
  ld3(x, a1, p1);
  ld3(x, a2, p2);
  const double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double r0sq = c.r0[k] * c.r0[k];
  return fabs(dot3(d, d) - r0sq) / (2.0 * r0sq);
}

// one SHAKE correction (ModifierShake.cpp:33-80): positions x, and with v
// non-null the velocities
__device__ inline void shake_one(const PmConstraints& c, double* x, double* v, const double* xl, int k) {
#pragma clang fp contract(off)
  const int a1 = c.a1[k], a2 = c.a2[k];
  const double rest = c.r0[k];
  double p1[3], p2[3], l1[3], l2[3];
  ld3(x, a1, p1);
  ld3(x, a2, p2);
  ld3(xl, a1, l1);
  ld3(xl, a2, l2);
  const double pab[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const double pabsq = dot3(pab, pab);
  const double rabsq = rest * rest;
  const double diffsq = rabsq - pabsq;
  const double rab[3] = {l1[0] - l2[0], l1[1] - l2[1], l1[2] - l2[2]};
  const double rpab = dot3(rab, pab);
  const double rM1 = c.rmass[a1], rM2 = c.rmass[a2];
  const double gab = diffsq / (2 * (rM1 + rM2) * rpab);
  double dp[3] = {rab[0] * gab, rab[1] * gab, rab[2] * gab};
  for (int d = 0; d < 3; d++) {
    p1[d] += dp[d] * rM1;
    p2[d] -= dp[d] * rM2;
  }
  st3(x, a1, p1);
  st3(x, a2, p2);
  if (v) {
    for (int d = 0; d < 3; d++) dp[d] /= c.dt;
    double v1[3], v2[3];
    ld3(v, a1, v1);
    ld3(v, a2, v2);
    for (int d = 0; d < 3; d++) {
      v1[d] += dp[d] * rM1;
      v2[d] -= dp[d] * rM2;
    }
    st3(v, a1, v1);
    st3(v, a2, v2);
  }
}

// ModifierMetaRattle::calcError term of one constraint: |(x1 - x2) . (v1 - v2)|
__device__ inline double rattle_err(const PmConstraints& c, const double* x, const double* v, int k) {
#pragma clang fp contract(off)
  const int a1 = c.a1[k], a2 = c.a2[k];
  double p1[3], p2[3], v1[3], v2[3];
  ld3(x, a1, p1);
  ld3(x, a2, p2);
  ld3(v, a1, v1);
  ld3(v, a2, v2);
  const double vab[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
  const double rab[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  return fabs(dot3(rab, vab));
}

// one RATTLE correction (ModifierRattle.cpp:31-66)
__device__ inline void rattle_one(const PmConstraints& c, const double* x, double* v, int k) {
#pragma clang fp contract(off)
  const int a1 = c.a1[k], a2 = c.a2[k];
  const double rM1 = c.rmass[a1], rM2 = c.rmass[a2];
  double p1[3], p2[3], v1[3], v2[3];
  ld3(x, a1, p1);
  ld3(x, a2, p2);
  ld3(v, a1, v1);
  ld3(v, a2, v2);
  const double rab[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const double rabsq = dot3(rab, rab);
  const double vab[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
  const double rvab = dot3(rab, vab);
  const double gab = -rvab / (c.dt * (rM1 + rM2) * rabsq);
  const double dp[3] = {rab[0] * gab, rab[1] * gab, rab[2] * gab};
  for (int d = 0; d < 3; d++) {
    v1[d] += dp[d] * c.dt * rM1;
    v2[d] -= dp[d] * c.dt * rM2;
  }
  st3(v, a1, v1);
  st3(v, a2, v2);
}

// probe: the group's error after 0 .. maxit sweeps, err[s * ngroup + g], on
// a scratch copy of its atoms' positions (SHAKE) or velocities (RATTLE)
template <bool RATTLE>
__global__ __launch_bounds__(256) void k_constraint_probe(PmConstraints c, const double* __restrict__ x,
                                                          const double* __restrict__ v,
                                                          const double* __restrict__ xl) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= c.ngroup) return;
  const int k0 = c.gstart[g], k1 = c.gstart[g + 1];
  double* s = c.scratch;
  const double* src = RATTLE ? v : x;
  for (int k = k0; k < k1; k++) {
    const int a1 = c.a1[k], a2 = c.a2[k];
    double t[3];
    ld3(src, a1, t);
    st3(s, a1, t);
    ld3(src, a2, t);
    st3(s, a2, t);
  }
  for (int it = 0;; it++) {
    double e = 0.0;
    for (int k = k0; k < k1; k++) e += RATTLE ? rattle_err(c, x, s, k) : shake_err(c, s, k);
    c.err[(long long)it * c.ngroup + g] = e;
    if (it == c.maxit) break;
    for (int k = k0; k < k1; k++) {
      if (RATTLE) rattle_one(c, x, s, k);
      else shake_one(c, s, nullptr, xl, k);
    }
  }
}

// mean[s] = sum over groups of err[s][g] / ncons, one block per s, fixed order
__global__ __launch_bounds__(256) void k_constraint_mean(PmConstraints c) {
  __shared__ double s_part[256];
  const int s = blockIdx.x;
  const double* e = c.err + (long long)s * c.ngroup;
  double t = 0.0;
  for (int g = threadIdx.x; g < c.ngroup; g += blockDim.x) t += e[g];
  s_part[threadIdx.x] = t;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s_part[threadIdx.x] += s_part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) c.mean[s] = s_part[0] / c.ncons;
}

// apply: the sweeps the reference runs (the first s with mean[s] <= eps, or
// maxit + 1), on x and v; SHAKE then records x as the next reference positions
template <bool RATTLE>
__global__ __launch_bounds__(256) void k_constraint_apply(PmConstraints c, double* __restrict__ x,
                                                          double* __restrict__ v, double* __restrict__ xl) {
  int sweeps = c.maxit + 1;
  for (int s = 0; s <= c.maxit; s++)
    if (!(c.mean[s] > c.eps)) { sweeps = s; break; }
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0 && c.sweeps) c.sweeps[0] = sweeps;
  if (g >= c.ngroup) return;
  const int k0 = c.gstart[g], k1 = c.gstart[g + 1];
  for (int it = 0; it < sweeps; it++)
    for (int k = k0; k < k1; k++) {
      if (RATTLE) rattle_one(c, x, v, k);
      else shake_one(c, x, v, xl, k);
    }
  if (!RATTLE)
    for (int k = k0; k < k1; k++) {
      const int a1 = c.a1[k], a2 = c.a2[k];
      double t[3];
      ld3(x, a1, t);
      st3(xl, a1, t);
      ld3(x, a2, t);
      st3(xl, a2, t);
    }
}

}  // namespace

bool pm_launch_constraints(hipStream_t s, const PmConstraints& c, int rattle, double* x, double* v, double* xl) {
  if (pm_recording()) {   // no batched form: replica batches fall back to chains
    pm_record(-1, 0, 0, 0, 0, nullptr, 0);
    return true;
  }
  if (c.ngroup <= 0) return true;
  const dim3 g((c.ngroup + 255) / 256), b(256);
  if (rattle) {
    hipLaunchKernelGGL(k_constraint_probe<true>, g, b, 0, s, c, x, v, xl);
    hipLaunchKernelGGL(k_constraint_mean, dim3(c.maxit + 1), b, 0, s, c);
    hipLaunchKernelGGL(k_constraint_apply<true>, g, b, 0, s, c, x, v, xl);
  } else {
    hipLaunchKernelGGL(k_constraint_probe<false>, g, b, 0, s, c, x, v, xl);
    hipLaunchKernelGGL(k_constraint_mean, dim3(c.maxit + 1), b, 0, s, c);
    hipLaunchKernelGGL(k_constraint_apply<false>, g, b, 0, s, c, x, v, xl);
  }
  return true;
}
