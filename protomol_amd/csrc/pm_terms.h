// pm_terms.h -- FP64 device functions of the excluded-pair correction and the
// bonded terms, shared by their own kernels and by the PME gather kernel that
// adds them per atom.
#pragma once

#include "pm_internal.h"

#include <math.h>

namespace {

// PeriodicBoundaryConditions::minimalDifference (topology/PeriodicBoundaryConditions.h:76-117)
__device__ inline void min_diff(const PmBox& b, const double* c1, const double* c2, double* d, double& s) {
#pragma clang fp contract(off)
  double x = c2[0] - c1[0], y = c2[1] - c1[1], z = c2[2] - c1[2];
  s = x * x + y * y + z * z;
  if (s > b.d) {
    if (x < -b.h2[0]) x += b.h[0]; else if (x > b.h2[0]) x -= b.h[0];
    if (y < -b.h2[1]) y += b.h[1]; else if (y > b.h2[1]) y -= b.h[1];
    if (z < -b.h2[2]) z += b.h[2]; else if (z > b.h2[2]) z -= b.h[2];
    s = x * x + y * y + z * z;
    if (s > b.d) {
      x -= b.L[0] * __builtin_rint(b.er[0] * x);
      y -= b.L[1] * __builtin_rint(b.er[1] * y);
      z -= b.L[2] * __builtin_rint(b.er[2] * z);
      s = x * x + y * y + z * z;
    }
  }
  d[0] = x; d[1] = y; d[2] = z;
}

// NonbondedPMEwaldSystemForce::correctionTerm (forces/NonbondedPMEwaldSystemForce.h:942-981)
// for one exclusion pair a1 < a2: fij (applied -= to a1, += to a2) and energy.
__device__ inline void corr_pair(const PmBox& b, const double* pos, const double* q, int a1, int a2,
                                 int cls, double alpha, double csf, double* rij, double* fij, double& e) {
  double r2;
  min_diff(b, pos + 3 * a1, pos + 3 * a2, rij, r2);
  double qq = q[a1] * q[a2];
  if (cls == PM_EXCL_MODIFIED) qq *= 1 - csf;
  const double r = sqrt(r2);
  const double rr = 1 / r;
  const double ee = erf(alpha * r) * rr;
  e = -qq * ee;
  const double tapi = 2.0 * alpha / sqrt(3.14159265358979323846);
  const double s = (qq * (tapi * exp(-alpha * alpha * r2) - ee) * rr * rr);
  fij[0] = rij[0] * s; fij[1] = rij[1] * s; fij[2] = rij[2] * s;
}

__device__ inline double norm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// one bond: force on atom1 (atom2 gets the opposite), energy, r12
__device__ inline void bond_term(const PmBox& b, const PmBondedLaunch& a, int k, double* f1, double& e,
                                 double* r12) {
  const int a1 = a.b_a[k], a2 = a.b_b[k];
  const double r0 = a.b_r0[k], kb = a.b_k[k];
  double d2;
  min_diff(b, a.pos + 3 * a2, a.pos + 3 * a1, r12, d2);   // vector from atom 2 to atom 1
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
  double r13[3], d2;
  min_diff(b, a.pos + 3 * a2, a.pos + 3 * a1, r12, d2);
  min_diff(b, a.pos + 3 * a2, a.pos + 3 * a3, r32, d2);
  min_diff(b, a.pos + 3 * a3, a.pos + 3 * a1, r13, d2);
  const double d12 = norm3(r12), d32 = norm3(r32), d13 = norm3(r13);
  double cr[3] = {r12[1] * r32[2] - r12[2] * r32[1], r12[2] * r32[0] - r12[0] * r32[2],
                  r12[0] * r32[1] - r12[1] * r32[0]};
  const double theta = atan2(norm3(cr), r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2]);
  double st, ct;
  sincos(theta, &st, &ct);
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

// One torsion of kind k (0 dihedral, 1 improper), as
// MTorsionSystemForce::calcTorsion (forces/MTorsionSystemForce.h:34-153):
// the forces f1, f2, f3 (atom 1 gets f1, atom 2 f2 - f1, atom 3 f3 - f2,
// atom 4 -f3), the energy and the three bond vectors.
__device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline void torsion_term(const PmBox& bx, const PmBondedLaunch& a, int kind, int t, double* f1,
                                    double* f2, double* f3, double& e, double* r12, double* r23, double* r34) {
  const int4 at = a.t_atoms[kind][t];
  double d2, A[3], B[3], Cc[3];
  min_diff(bx, a.pos + 3 * at.y, a.pos + 3 * at.x, r12, d2);
  min_diff(bx, a.pos + 3 * at.z, a.pos + 3 * at.y, r23, d2);
  min_diff(bx, a.pos + 3 * at.w, a.pos + 3 * at.z, r34, d2);
  cross3(r12, r23, A);
  cross3(r23, r34, B);
  cross3(r23, A, Cc);
  const double ra = 1.0 / norm3(A), rb = 1.0 / norm3(B), rc = 1.0 / norm3(Cc);
  for (int c = 0; c < 3; c++) { A[c] *= ra; B[c] *= rb; Cc[c] *= rc; }
  const double cosp = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
  const double sinp = Cc[0] * B[0] + Cc[1] * B[1] + Cc[2] * B[2];
  const double phi = -atan2(sinp, cosp);
  const double pi = 3.14159265358979323846;
  double dpot = 0.0;
  e = 0.0;
  for (int u = a.t_term[kind][t]; u < a.t_term[kind][t + 1]; u++) {
    const int n = a.t_n[kind][u];
    const double k = a.t_k[kind][u], ph = a.t_phase[kind][u];
    if (n > 0) {
      double sn, cs;
      sincos(n * phi + ph, &sn, &cs);     // one argument reduction for both
      dpot -= n * k * sn;
      e += k * (1.0 + cs);
    } else {
      double diff = phi - ph;
      if (diff < -pi) diff += 2 * pi;
      else if (diff > pi) diff -= 2 * pi;
      dpot += 2.0 * k * diff;
      e += k * diff * diff;
    }
  }
  if (fabs(sinp) > 0.1) {
    double dA[3], dB[3];
    const double k1 = dpot / sinp;
    for (int c = 0; c < 3; c++) { dA[c] = (A[c] * cosp - B[c]) * ra; dB[c] = (B[c] * cosp - A[c]) * rb; }
    f1[0] = k1 * (r23[1] * dA[2] - r23[2] * dA[1]);
    f1[1] = k1 * (r23[2] * dA[0] - r23[0] * dA[2]);
    f1[2] = k1 * (r23[0] * dA[1] - r23[1] * dA[0]);
    f3[0] = k1 * (r23[2] * dB[1] - r23[1] * dB[2]);
    f3[1] = k1 * (r23[0] * dB[2] - r23[2] * dB[0]);
    f3[2] = k1 * (r23[1] * dB[0] - r23[0] * dB[1]);
    f2[0] = k1 * (r12[2] * dA[1] - r12[1] * dA[2] + r34[1] * dB[2] - r34[2] * dB[1]);
    f2[1] = k1 * (r12[0] * dA[2] - r12[2] * dA[0] + r34[2] * dB[0] - r34[0] * dB[2]);
    f2[2] = k1 * (r12[1] * dA[0] - r12[0] * dA[1] + r34[0] * dB[1] - r34[1] * dB[0]);
  } else {
    double dC[3], dB[3], x[3];
    const double k1 = -dpot / cosp;
    for (int c = 0; c < 3; c++) { dC[c] = (Cc[c] * sinp - B[c]) * rc; dB[c] = (B[c] * sinp - Cc[c]) * rb; }
    f1[0] = k1 * ((r23[1] * r23[1] + r23[2] * r23[2]) * dC[0] - r23[0] * r23[1] * dC[1] - r23[0] * r23[2] * dC[2]);
    f1[1] = k1 * ((r23[2] * r23[2] + r23[0] * r23[0]) * dC[1] - r23[1] * r23[2] * dC[2] - r23[1] * r23[0] * dC[0]);
    f1[2] = k1 * ((r23[0] * r23[0] + r23[1] * r23[1]) * dC[2] - r23[2] * r23[0] * dC[0] - r23[2] * r23[1] * dC[1]);
    cross3(dB, r23, x);
    for (int c = 0; c < 3; c++) f3[c] = x[c] * k1;
    f2[0] = k1 * (-(r23[1] * r12[1] + r23[2] * r12[2]) * dC[0] + (2.0 * r23[0] * r12[1] - r12[0] * r23[1]) * dC[1] +
                  (2.0 * r23[0] * r12[2] - r12[0] * r23[2]) * dC[2] + dB[2] * r34[1] - dB[1] * r34[2]);
    f2[1] = k1 * (-(r23[2] * r12[2] + r23[0] * r12[0]) * dC[1] + (2.0 * r23[1] * r12[2] - r12[1] * r23[2]) * dC[2] +
                  (2.0 * r23[1] * r12[0] - r12[1] * r23[0]) * dC[0] + dB[0] * r34[2] - dB[2] * r34[0]);
    f2[2] = k1 * (-(r23[0] * r12[0] + r23[1] * r12[1]) * dC[2] + (2.0 * r23[2] * r12[0] - r12[2] * r23[0]) * dC[0] +
                  (2.0 * r23[2] * r12[1] - r12[2] * r23[1]) * dC[1] + dB[1] * r34[0] - dB[0] * r34[1]);
  }
}

// Correction force on atom i: the sum over its exclusion partners (CSR,
// partner encoded as 2*atom + (1 if i is the pair's second atom)).
__device__ inline void corr_atom_force(const PmBox& box, int i, const double* __restrict__ pos,
                                       const double* __restrict__ q, const int* __restrict__ row,
                                       const int* __restrict__ partner, const int* __restrict__ cls, double alpha,
                                       double csf, double* f) {
  f[0] = f[1] = f[2] = 0.0;
  for (int u = row[i]; u < row[i + 1]; u++) {
    const int pe = partner[u];
    const int other = pe >> 1;
    const bool second = pe & 1;
    const int a1 = second ? other : i, a2 = second ? i : other;
    double rij[3], fij[3], e;
    corr_pair(box, pos, q, a1, a2, cls[u], alpha, csf, rij, fij, e);
    if (second) { f[0] += fij[0]; f[1] += fij[1]; f[2] += fij[2]; }
    else { f[0] -= fij[0]; f[1] -= fij[1]; f[2] -= fij[2]; }
  }
}

// Bond and angle forces on atom i: each term it belongs to, its own share.
__device__ inline void bonded_atom_force(const PmBox& box, const PmBondedLaunch& a, int i, double* f) {
  f[0] = f[1] = f[2] = 0.0;
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
  for (int kind = 0; kind < 2; kind++) {
    if (!a.do_tors[kind]) continue;
    for (int u = a.t_row[kind][i]; u < a.t_row[kind][i + 1]; u++) {
      const int t = a.t_idx[kind][u];
      double f1[3], f2[3], f3[3], e, r12[3], r23[3], r34[3];
      torsion_term(box, a, kind, t, f1, f2, f3, e, r12, r23, r34);
      const int4 at = a.t_atoms[kind][t];
      for (int c = 0; c < 3; c++)
        f[c] += (at.x == i) ? f1[c] : (at.y == i) ? f2[c] - f1[c] : (at.z == i) ? f3[c] - f2[c] : -f3[c];
    }
  }
}

}  // namespace
