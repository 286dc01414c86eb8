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
// sweeps is global.  A probe kernel records the error after 0 .. maxIter
// sweeps (on copies of the positions or velocities), summed per block; every
// block of the apply kernel sums those parts in the same fixed order into the
// mean error, finds the first sweep count whose mean is within epsilon and
// runs that many sweeps on the real positions and velocities.  Groups of up
// to 4 atoms (waters, methyl groups) keep their atoms in registers.
#include "pm_internal.h"

#include <math.h>

#include <mutex>

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
#pragma unroll
  for (int d = 0; d < 3; d++) {
    p1[d] += dp[d] * rM1;
    p2[d] -= dp[d] * rM2;
  }
  st3(x, a1, p1);
  st3(x, a2, p2);
  if (v) {
#pragma unroll
    for (int d = 0; d < 3; d++) dp[d] /= c.dt;
    double v1[3], v2[3];
    ld3(v, a1, v1);
    ld3(v, a2, v2);
#pragma unroll
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
#pragma unroll
  for (int d = 0; d < 3; d++) {
    v1[d] += dp[d] * c.dt * rM1;
    v2[d] -= dp[d] * c.dt * rM2;
  }
  st3(v, a1, v1);
  st3(v, a2, v2);
}

// ---- groups of at most 4 atoms and 6 constraints: the atoms in registers, each constraint's
// atoms as local indices (la = local a1 | local a2 << 2); reads and writes go
// through selects over the unrolled atom slots, so nothing is indexed
// dynamically
struct Reg4 {
  double p[4][3];
};
__device__ inline void rget(const Reg4& r, int li, double* o) {
#pragma unroll
  for (int d = 0; d < 3; d++) o[d] = li == 0 ? r.p[0][d] : li == 1 ? r.p[1][d] : li == 2 ? r.p[2][d] : r.p[3][d];
}
__device__ inline void rput(Reg4& r, int li, const double* o) {
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int d = 0; d < 3; d++)
      if (li == i) r.p[i][d] = o[d];
}
__device__ inline void rload(Reg4& r, const double* src, int4 at) {
  const int a[4] = {at.x, at.y, at.z, at.w};
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int d = 0; d < 3; d++) r.p[i][d] = a[i] >= 0 ? src[3 * a[i] + d] : 0.0;
}
__device__ inline void rstore(const Reg4& r, double* dst, int4 at) {
  const int a[4] = {at.x, at.y, at.z, at.w};
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (a[i] >= 0)
#pragma unroll
      for (int d = 0; d < 3; d++) dst[3 * a[i] + d] = r.p[i][d];
}

// a small group's constraints (at most 6), loaded once: local atom indices,
// rest length, 1/mass of both atoms
struct Con {
  int la;
  double r0, m1, m2;
};
struct Cons6 {
  Con k[6];
  int n;
};
__device__ inline void load_cons(const PmConstraints& c, int k0, int k1, Cons6& q) {
  q.n = k1 - k0;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    const int k = k0 + (j < q.n ? j : 0);
    q.k[j].la = c.la[k];
    q.k[j].r0 = c.r0[k];
    q.k[j].m1 = c.rmass[c.a1[k]];
    q.k[j].m2 = c.rmass[c.a2[k]];
  }
}

// the same arithmetic as shake_err / shake_one / rattle_err / rattle_one on
// register-held atoms and constraints
__device__ inline double shake_err_r(const Con& q, const Reg4& X) {
#pragma clang fp contract(off)
  double p1[3], p2[3];
  rget(X, q.la & 3, p1);
  rget(X, q.la >> 2, p2);
  const double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double r0sq = q.r0 * q.r0;
  return fabs(dot3(d, d) - r0sq) / (2.0 * r0sq);
}
template <bool VEL>
__device__ inline void shake_one_r(const Con& q, double dt, Reg4& X, Reg4& V, const Reg4& XL) {
#pragma clang fp contract(off)
  const int l1 = q.la & 3, l2 = q.la >> 2;
  double p1[3], p2[3], q1[3], q2[3];
  rget(X, l1, p1);
  rget(X, l2, p2);
  rget(XL, l1, q1);
  rget(XL, l2, q2);
  const double pab[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const double pabsq = dot3(pab, pab);
  const double rabsq = q.r0 * q.r0;
  const double diffsq = rabsq - pabsq;
  const double rab[3] = {q1[0] - q2[0], q1[1] - q2[1], q1[2] - q2[2]};
  const double rpab = dot3(rab, pab);
  const double rM1 = q.m1, rM2 = q.m2;
  const double gab = diffsq / (2 * (rM1 + rM2) * rpab);
  double dp[3] = {rab[0] * gab, rab[1] * gab, rab[2] * gab};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    p1[d] += dp[d] * rM1;
    p2[d] -= dp[d] * rM2;
  }
  rput(X, l1, p1);
  rput(X, l2, p2);
  if (VEL) {
#pragma unroll
    for (int d = 0; d < 3; d++) dp[d] /= dt;
    double v1[3], v2[3];
    rget(V, l1, v1);
    rget(V, l2, v2);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      v1[d] += dp[d] * rM1;
      v2[d] -= dp[d] * rM2;
    }
    rput(V, l1, v1);
    rput(V, l2, v2);
  }
}
__device__ inline double rattle_err_r(const Con& q, const Reg4& X, const Reg4& V) {
#pragma clang fp contract(off)
  double p1[3], p2[3], v1[3], v2[3];
  rget(X, q.la & 3, p1);
  rget(X, q.la >> 2, p2);
  rget(V, q.la & 3, v1);
  rget(V, q.la >> 2, v2);
  const double vab[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
  const double rab[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  return fabs(dot3(rab, vab));
}
__device__ inline void rattle_one_r(const Con& q, double dt, const Reg4& X, Reg4& V) {
#pragma clang fp contract(off)
  const int l1 = q.la & 3, l2 = q.la >> 2;
  const double rM1 = q.m1, rM2 = q.m2;
  double p1[3], p2[3], v1[3], v2[3];
  rget(X, l1, p1);
  rget(X, l2, p2);
  rget(V, l1, v1);
  rget(V, l2, v2);
  const double rab[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const double rabsq = dot3(rab, rab);
  const double vab[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
  const double rvab = dot3(rab, vab);
  const double gab = -rvab / (dt * (rM1 + rM2) * rabsq);
  const double dp[3] = {rab[0] * gab, rab[1] * gab, rab[2] * gab};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    v1[d] += dp[d] * dt * rM1;
    v2[d] -= dp[d] * dt * rM2;
  }
  rput(V, l1, v1);
  rput(V, l2, v2);
}

// the register arithmetic with the local atom indices fixed at compile time
template <int L1, int L2, bool VEL>
__device__ inline void shake_one_c(const Con& q, double dt, Reg4& X, Reg4& V, const Reg4& XL) {
#pragma clang fp contract(off)
  const double* p1 = X.p[L1];
  const double* p2 = X.p[L2];
  const double pab[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const double pabsq = dot3(pab, pab);
  const double rabsq = q.r0 * q.r0;
  const double diffsq = rabsq - pabsq;
  const double rab[3] = {XL.p[L1][0] - XL.p[L2][0], XL.p[L1][1] - XL.p[L2][1], XL.p[L1][2] - XL.p[L2][2]};
  const double rpab = dot3(rab, pab);
  const double rM1 = q.m1, rM2 = q.m2;
  const double gab = diffsq / (2 * (rM1 + rM2) * rpab);
  double dp[3] = {rab[0] * gab, rab[1] * gab, rab[2] * gab};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    X.p[L1][d] += dp[d] * rM1;
    X.p[L2][d] -= dp[d] * rM2;
  }
  if (VEL) {
#pragma unroll
    for (int d = 0; d < 3; d++) dp[d] /= dt;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      V.p[L1][d] += dp[d] * rM1;
      V.p[L2][d] -= dp[d] * rM2;
    }
  }
}
template <int L1, int L2>
__device__ inline double shake_err_c(const Con& q, const Reg4& X) {
#pragma clang fp contract(off)
  const double d[3] = {X.p[L2][0] - X.p[L1][0], X.p[L2][1] - X.p[L1][1], X.p[L2][2] - X.p[L1][2]};
  const double r0sq = q.r0 * q.r0;
  return fabs(dot3(d, d) - r0sq) / (2.0 * r0sq);
}
template <int L1, int L2>
__device__ inline void rattle_one_c(const Con& q, double dt, const Reg4& X, Reg4& V) {
#pragma clang fp contract(off)
  const double rM1 = q.m1, rM2 = q.m2;
  const double rab[3] = {X.p[L1][0] - X.p[L2][0], X.p[L1][1] - X.p[L2][1], X.p[L1][2] - X.p[L2][2]};
  const double rabsq = dot3(rab, rab);
  const double vab[3] = {V.p[L1][0] - V.p[L2][0], V.p[L1][1] - V.p[L2][1], V.p[L1][2] - V.p[L2][2]};
  const double rvab = dot3(rab, vab);
  const double gab = -rvab / (dt * (rM1 + rM2) * rabsq);
  const double dp[3] = {rab[0] * gab, rab[1] * gab, rab[2] * gab};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    V.p[L1][d] += dp[d] * dt * rM1;
    V.p[L2][d] -= dp[d] * dt * rM2;
  }
}
template <int L1, int L2>
__device__ inline double rattle_err_c(const Reg4& X, const Reg4& V) {
#pragma clang fp contract(off)
  const double vab[3] = {V.p[L1][0] - V.p[L2][0], V.p[L1][1] - V.p[L2][1], V.p[L1][2] - V.p[L2][2]};
  const double rab[3] = {X.p[L1][0] - X.p[L2][0], X.p[L1][1] - X.p[L2][1], X.p[L1][2] - X.p[L2][2]};
  return fabs(dot3(rab, vab));
}

// one sweep / the error of a small group, constraints in list order; a
// water group (3 atoms, constraints (0,1), (0,2), (1,2) in that order) with
// its indices fixed at compile time
template <bool RATTLE, bool VEL>
__device__ inline void sweep_r(const Cons6& q, bool water, double dt, Reg4& X, Reg4& V, const Reg4& XL) {
  if (water) {
    if (RATTLE) {
      rattle_one_c<0, 1>(q.k[0], dt, X, V);
      rattle_one_c<0, 2>(q.k[1], dt, X, V);
      rattle_one_c<1, 2>(q.k[2], dt, X, V);
    } else {
      shake_one_c<0, 1, VEL>(q.k[0], dt, X, V, XL);
      shake_one_c<0, 2, VEL>(q.k[1], dt, X, V, XL);
      shake_one_c<1, 2, VEL>(q.k[2], dt, X, V, XL);
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < 6; j++)
    if (j < q.n) {
      if (RATTLE) rattle_one_r(q.k[j], dt, X, V);
      else shake_one_r<VEL>(q.k[j], dt, X, V, XL);
    }
}
template <bool RATTLE>
__device__ inline double error_r(const Cons6& q, bool water, const Reg4& X, const Reg4& V) {
  if (water) {
    double e = 0.0;
    e += RATTLE ? rattle_err_c<0, 1>(X, V) : shake_err_c<0, 1>(q.k[0], X);
    e += RATTLE ? rattle_err_c<0, 2>(X, V) : shake_err_c<0, 2>(q.k[1], X);
    e += RATTLE ? rattle_err_c<1, 2>(X, V) : shake_err_c<1, 2>(q.k[2], X);
    return e;
  }
  double e = 0.0;
#pragma unroll
  for (int j = 0; j < 6; j++)
    if (j < q.n) e += RATTLE ? rattle_err_r(q.k[j], X, V) : shake_err_r(q.k[j], X);
  return e;
}

// the wave's sum (lanes in a fixed order) into part[at] by lane 0
__device__ inline void wave_part(double e, double* part, long long at) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  if ((threadIdx.x & 63) == 0) part[at] = e;
}

// probe: every group's error after 0 .. maxit sweeps, summed per wave into
// part[s * waves + wave]; small groups in registers, larger ones on a
// scratch copy of their atoms' positions (SHAKE) or velocities (RATTLE)
template <bool RATTLE>
__global__ __launch_bounds__(256) void k_constraint_probe(PmConstraints c, const double* __restrict__ x,
                                                          const double* __restrict__ v,
                                                          const double* __restrict__ xl) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int nw = gridDim.x * (blockDim.x >> 6), gw = g >> 6;
  const bool valid = g < c.ngroup;
  const int k0 = valid ? c.gstart[g] : 0, k1 = valid ? c.gstart[g + 1] : 0;
  const int4 at = valid ? c.gatom[g] : make_int4(-1, -1, -1, -1);
  const bool small = valid && at.x >= 0;
  const bool water = small && at.w == PM_CONS_WATER;
  Reg4 X, V, XL;
  Cons6 q;
  if (small) {
    load_cons(c, k0, k1, q);
    rload(X, x, at);
    if (RATTLE) rload(V, v, at);
    else rload(XL, xl, at);
  } else if (valid) {
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
  }
  for (int it = 0;; it++) {
    double e = 0.0;
    if (small) {
      e = error_r<RATTLE>(q, water, X, V);
    } else if (valid) {
      for (int k = k0; k < k1; k++) e += RATTLE ? rattle_err(c, x, c.scratch, k) : shake_err(c, c.scratch, k);
    }
    wave_part(e, c.err, (long long)it * nw + gw);
    if (it == c.maxit) break;
    if (small) {
      sweep_r<RATTLE, false>(q, water, c.dt, X, V, XL);
    } else if (valid) {
      for (int k = k0; k < k1; k++) {
        if (RATTLE) rattle_one(c, x, c.scratch, k);
        else shake_one(c, c.scratch, nullptr, xl, k);
      }
    }
  }
}

// apply: every block sums the probe's wave parts in the same fixed order to
// the mean error after s sweeps (four sweep counts at a time, one per wave)
// and takes the reference's sweep count (the first s with mean <= eps, or
// maxit + 1); then that many sweeps on x and v, and SHAKE records x as the
// next reference positions
template <bool RATTLE>
__global__ __launch_bounds__(256) void k_constraint_apply(PmConstraints c, double* __restrict__ x,
                                                          double* __restrict__ v, double* __restrict__ xl) {
  __shared__ double s_mean[4];
  __shared__ int s_sweeps;
  const int nw = gridDim.x * (blockDim.x >> 6);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_sweeps = -1;
  for (int s0 = 0; s0 <= c.maxit; s0 += 4) {
    const int s = s0 + w;
    double t = 0.0;
    if (s <= c.maxit)
      for (int u = lane; u < nw; u += 64) t += c.err[(long long)s * nw + u];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (lane == 0) s_mean[w] = t / c.ncons;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int u = 0; u < 4 && s0 + u <= c.maxit; u++)
        if (!(s_mean[u] > c.eps)) { s_sweeps = s0 + u; break; }
    __syncthreads();
    if (s_sweeps >= 0) break;
  }
  if (s_sweeps < 0) s_sweeps = c.maxit + 1;   // every thread writes the same value
  __syncthreads();
  const int sweeps = s_sweeps;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0 && c.sweeps) c.sweeps[0] = sweeps;
  if (g >= c.ngroup) return;
  const int k0 = c.gstart[g], k1 = c.gstart[g + 1];
  const int4 at = c.gatom[g];
  if (at.x >= 0) {
    Reg4 X, V, XL;
    Cons6 q;
    load_cons(c, k0, k1, q);
    rload(X, x, at);
    rload(V, v, at);
    if (!RATTLE) rload(XL, xl, at);
    const bool water = at.w == PM_CONS_WATER;
    for (int it = 0; it < sweeps; it++) sweep_r<RATTLE, true>(q, water, c.dt, X, V, XL);
    if (!RATTLE) {
      rstore(X, x, at);
      rstore(X, xl, at);
    }
    rstore(V, v, at);
    return;
  }
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

// Work space of the allocation-free SHAKE entry point (pm_shake_noalloc): one
// copy per GPU for the life of the process.  Host code never passes `scratch`
// itself (in host code it names a host-side stand-in, not device memory):
// pm_scratch_init looks up each device's address once, with that device
// current, and every use goes through that address.
static __device__ __align__(256) unsigned char scratch[PM_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[PM_MAX_DEVICES] = {};
}  // namespace

int pm_scratch_init(int device) {
  if (device < 0 || device >= PM_MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* ptr = nullptr;
  // the caller has made `device` current
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
  g_scratch_ptr[device] = ptr;
  return 0;
}

void* pm_scratch_address(int device) {
  if (device < 0 || device >= PM_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

bool pm_launch_constraints(hipStream_t s, const PmConstraints& c, int rattle, double* x, double* v, double* xl) {
  if (pm_recording()) {   // no batched form: replica batches fall back to chains
    pm_record(-1, 0, 0, 0, 0, nullptr, 0);
    return true;
  }
  if (c.ngroup <= 0) return true;
  const dim3 g((c.ngroup + 255) / 256), b(256);
  if (rattle) {
    hipLaunchKernelGGL(k_constraint_probe<true>, g, b, 0, s, c, x, v, xl);
    hipLaunchKernelGGL(k_constraint_apply<true>, g, b, 0, s, c, x, v, xl);
  } else {
    hipLaunchKernelGGL(k_constraint_probe<false>, g, b, 0, s, c, x, v, xl);
    hipLaunchKernelGGL(k_constraint_apply<false>, g, b, 0, s, c, x, v, xl);
  }
  return true;
}
