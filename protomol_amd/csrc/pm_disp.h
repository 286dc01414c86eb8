// pm_disp.h -- the pair-list reuse test on one atom, shared by the
// displacement-check kernel and the MD position-update kernels that run the
// test on the positions they have just written.
#pragma once

#include "pm_internal.h"

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
