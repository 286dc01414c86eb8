// pm_step.h -- per-atom pieces of the evaluation's start that the MD
// position-update kernels also run on the positions they have just written:
// the reference's cell assignment (binning) and the pair-list reuse test.
#pragma once

#include "pm_internal.h"

// PeriodicBoundaryConditions::minimalPosition (topology/PeriodicBoundaryConditions.h:120-160)
__device__ inline void min_pos(const PmBox& b, double cx, double cy, double cz, double& ox,
                               double& oy, double& oz) {
#pragma clang fp contract(off)
  double x = cx - b.origin[0], y = cy - b.origin[1], z = cz - b.origin[2];
  if (x * x + y * y + z * z <= b.d) { ox = x; oy = y; oz = z; return; }
  if (x < -b.h2[0]) x += b.h[0]; else if (x > b.h2[0]) x -= b.h[0];
  if (y < -b.h2[1]) y += b.h[1]; else if (y > b.h2[1]) y -= b.h[1];
  if (z < -b.h2[2]) z += b.h[2]; else if (z > b.h2[2]) z -= b.h[2];
  if (x * x + y * y + z * z <= b.d) { ox = x; oy = y; oz = z; return; }
  double tx = b.L[0] * __builtin_rint(b.er[0] * x);
  double ty = b.L[1] * __builtin_rint(b.er[1] * y);
  double tz = b.L[2] * __builtin_rint(b.er[2] * z);
  ox = x - tx; oy = y - ty; oz = z - tz;
}

// CubicCellManager::findCell (topology/CubicCellManager.h:61-64); coordinates
// outside [0, n) (the reference's x == +L/2 edge, or non-finite input) are
// clamped into the grid so that no index leaves its buffer.
__device__ inline int cell_index(const PmBox& b, double px, double py, double pz, double* mpz) {
#pragma clang fp contract(off)
  double mx, my, mz;
  min_pos(b, px, py, pz, mx, my, mz);
  double v[3] = {b.delta[0] + mx, b.delta[1] + my, b.delta[2] + mz};
  int c[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double f = __builtin_floor(v[k] * b.rreal[k]);
    int ck;
    if (!(f >= 0.0)) ck = 0;
    else if (f >= (double)b.n[k]) ck = b.n[k] - 1;
    else ck = (int)f;
    c[k] = ck;
  }
  if (mpz) *mpz = mz;
  return (c[0] * b.n[1] + c[1]) * b.n[2] + c[2];
}


// true when the atom at x has moved more than sqrt(lim2) A from its position
// at the last list build (b: 32-bit periodic fixed point, w = 1 when valid;
// the difference of two fixed-point coordinates is the minimal image).
// NaN positions count as moved.
__device__ inline bool pm_moved_past(const double* to_fix, const double* fix_to_A, const double* x, int4 b,
                                     float lim2) {
  if (b.w != 1) return true;
  const unsigned int fx = (unsigned int)__double2ll_rn(x[0] * to_fix[0]);
  const unsigned int fy = (unsigned int)__double2ll_rn(x[1] * to_fix[1]);
  const unsigned int fz = (unsigned int)__double2ll_rn(x[2] * to_fix[2]);
  const float dx = (float)(int)(fx - (unsigned int)b.x) * (float)fix_to_A[0];
  const float dy = (float)(int)(fy - (unsigned int)b.y) * (float)fix_to_A[1];
  const float dz = (float)(int)(fz - (unsigned int)b.z) * (float)fix_to_A[2];
  return !(dx * dx + dy * dy + dz * dz <= lim2);
}

// The binning of atom i at position x (buildCellLists.cpp:11-38 through
// cell_index): its cell, its sort key (minimal-image z) and an arrival rank
// within the cell from the per-cell count.  The wave's atoms of one cell take
// their ranks from one atomic (the group's lowest lane adds the group's size),
// so a wave issues one atomic per distinct cell instead of one per atom, all
// in one instruction.  Every active lane of the wave must call this together.
__device__ inline void pm_bin_atom(const PmBox& box, int i, const double* x, int* __restrict__ cell_of,
                                   float* __restrict__ key, int* __restrict__ rank, int* __restrict__ bin_count) {
  double mz;
  const int c = cell_index(box, x[0], x[1], x[2], &mz);
  cell_of[i] = c;
  // a NaN key (non-finite position) would break the strict (key, index)
  // order that makes every cell's slot ranks a permutation
  key[i] = mz == mz ? (float)mz : 0.0f;
  // groups of lanes with the same cell: no memory access in this loop
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(1);
  int leader_of = lane, off = 0, gsize = 0;
  while (todo) {
    const int leader = (int)__builtin_ctzll(todo);
    const int lc = __shfl(c, leader, 64);
    const unsigned long long same = __ballot(c == lc);
    if (c == lc) {
      leader_of = leader;
      off = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(same >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)same, 0));
    }
    if (lane == leader) gsize = __popcll(same);
    todo &= ~same;
  }
  // the leaders' atomics in one instruction, then every lane its group's base
  int base = 0;
  if (lane == leader_of) base = atomicAdd(&bin_count[c], gsize);
  base = __shfl(base, leader_of, 64);
  rank[i] = base + off;   // arrival rank; k_fill_slots orders each cell by (z, index)
}

