// pm_md.hip -- FP64 kernels around the force kernels: the PME excluded-pair
// correction, energy reductions, fixed-point conversions and the leapfrog /
// Langevin position-velocity updates.
#include "pm_internal.h"
#include "pm_terms.h"
#include "pm_step.h"

#include <math.h>

namespace {

// per-atom: the correction force over the atom's exclusion partners
__global__ void k_corr_force(PmBox box, int n, const double* __restrict__ pos, const double* __restrict__ q,
                             const int* __restrict__ row, const int* __restrict__ partner,
                             const int* __restrict__ cls, double alpha, double csf,
                             long long* __restrict__ force_fix) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[3];
  corr_atom_force(box, i, pos, q, row, partner, cls, alpha, csf, f);
  force_fix[3 * i] += pm_force_fix(f[0], box.err);
  force_fix[3 * i + 1] += pm_force_fix(f[1], box.err);
  force_fix[3 * i + 2] += pm_force_fix(f[2], box.err);
}

// Adjoint of the correction term (FP64), per atom over its exclusion partners.
// mode 1: out[3i] += H v with, per pair, t = F dv + 2 F_u (d . dv) d added to
//   a1 and subtracted from a2 (d = x_a2 - x_a1, F = force/r of corr_pair,
//   F_u its derivative in u = r^2: the exact derivatives of -qq erf(ar)/r);
// mode 0: out[i] += dE/dq_i = -q_other erf(a r)/r (x (1 - csf) for MODIFIED)
//   and the point self term -2 a q_i / sqrt(pi) (NonbondedPMEwaldSystemForce.h:604-613).
__global__ void k_corr_adjoint(PmBox box, int n, const double* __restrict__ pos, const double* __restrict__ q,
                               const int* __restrict__ row, const int* __restrict__ partner,
                               const int* __restrict__ cls, double alpha, double csf, int mode,
                               const double* __restrict__ v, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double rpi = 1.0 / sqrt(3.14159265358979323846);
  double acc[3] = {0, 0, 0};
  for (int u = row[i]; u < row[i + 1]; u++) {
    const int pe = partner[u];
    const int other = pe >> 1;
    const bool second = pe & 1;
    const int a1 = second ? other : i, a2 = second ? i : other;
    double d[3], r2;
    min_diff(box, pos + 3 * a1, pos + 3 * a2, d, r2);
    double sc = 1.0;
    if (cls[u] == PM_EXCL_MODIFIED) sc = 1 - csf;
    const double r = sqrt(r2), rr = 1 / r;
    const double erfv = erf(alpha * r), ex = exp(-alpha * alpha * r2);
    if (mode == 0) {
      acc[0] += -sc * q[other] * erfv * rr;
      continue;
    }
    const double qq = q[a1] * q[a2] * sc;
    const double c = 2.0 * alpha * rpi;
    // E = -qq erf(ar)/r;  E_r and E_rr in r
    const double Er = -qq * (c * ex * rr - erfv * rr * rr);
    const double Err = qq * (2.0 * alpha * alpha * c * ex + 2.0 * c * ex * rr * rr - 2.0 * erfv * rr * rr * rr);
    const double F = -Er * rr;                         // force/r (fij = F d on a2)
    const double Fu = -(Err - Er * rr) / (2.0 * r2);   // dF/du
    const double dv[3] = {v[3 * a2] - v[3 * a1], v[3 * a2 + 1] - v[3 * a1 + 1], v[3 * a2 + 2] - v[3 * a1 + 2]};
    const double g = 2.0 * Fu * (d[0] * dv[0] + d[1] * dv[1] + d[2] * dv[2]);
    const double sgn = second ? -1.0 : 1.0;
    for (int k = 0; k < 3; k++) acc[k] += sgn * (F * dv[k] + g * d[k]);
  }
  if (mode == 0) out[i] += acc[0] - 2.0 * alpha * rpi * q[i];
  else { out[3 * i] += acc[0]; out[3 * i + 1] += acc[1]; out[3 * i + 2] += acc[2]; }
}

// out[k] = fixed[k] * 2^-40 + out[k]
__global__ void k_add_fix_to(int m, const long long* __restrict__ fixed, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] += (double)fixed[i] * PM_FORCE_UNFIX;
}

// per table entry: intra-molecular energy and virial, block partial sums
__global__ __launch_bounds__(256) void k_corr_energy(PmBox box, const double* __restrict__ pos,
                                                     const double* __restrict__ q, int ntab,
                                                     const int* __restrict__ ta, const int* __restrict__ tb,
                                                     const int* __restrict__ tc, double alpha, double csf,
                                                     int want_virial, double* __restrict__ epart) {
  __shared__ double s_red[4][7];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  if (t < ntab) {
    double rij[3], fij[3], e;
    corr_pair(box, pos, q, ta[t], tb[t], tc[t], alpha, csf, rij, fij, e);
    v[0] = e;
    if (want_virial) {
      // ScalarStructure::addVirial (base/ScalarStructure.cpp:41-63)
      v[1] = fij[0] * rij[0]; v[2] = fij[0] * rij[1]; v[3] = fij[0] * rij[2];
      v[4] = fij[1] * rij[1]; v[5] = fij[1] * rij[2]; v[6] = fij[2] * rij[2];
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
    double* ep = epart + (long long)blockIdx.x * PM_NE;
    ep[PM_E_COUL] = s[0];
    for (int u = 1; u < PM_E_VIR; u++) ep[u] = 0.0;
    ep[PM_E_VIR + 0] = s[1]; ep[PM_E_VIR + 1] = s[2]; ep[PM_E_VIR + 2] = s[3];
    ep[PM_E_VIR + 3] = s[2]; ep[PM_E_VIR + 4] = s[4]; ep[PM_E_VIR + 5] = s[5];
    ep[PM_E_VIR + 6] = s[3]; ep[PM_E_VIR + 7] = s[5]; ep[PM_E_VIR + 8] = s[6];
  }
}

// PME real-space term in boxes no wider than twice the cutoff
// (NonbondedPMEwaldSystemForce.h:658-768): every unordered pair i <= j (self
// pairs included) against every lattice vector of `lat` ((0,0,0) first, then
// the vectors of norm <= 2 rc by norm, PeriodicBoundaryConditions.cpp:80-111):
// the first image of a FULL-excluded pair or of a self pair is skipped, a
// MODIFIED pair takes coulombScalingFactor on its first image within the
// cutoff, images beyond the cutoff contribute nothing (Cutoff switching), the
// erfc by poly5 (mathutilities.cpp:36-40).  FP64 in both precisions: forces
// into the 64-bit fixed-point atom forces, energy and virial into one record
// per block.
#define PM_LAT_MAX 1024
__device__ inline double poly5_ref(double ar) {
  const double t = 1.0 / (1.0 + 0.3275911 * ar);
  return t * (0.254829592 + t * (-0.284496736 + t * (1.421413741 + t * (-1.453152027 + t * 1.061405429))));
}
__global__ __launch_bounds__(256) void k_ewald_lattice(PmBox box, int n, const double* __restrict__ pos,
                                                       const double* __restrict__ q, const int* __restrict__ xs_row,
                                                       const int* __restrict__ xs_part, const double* __restrict__ lat,
                                                       int nlat, double rc2, double alpha, double csf,
                                                       int want_virial, long long* __restrict__ force_fix,
                                                       double* __restrict__ epart) {
  __shared__ double s_lat[3 * PM_LAT_MAX];
  __shared__ double s_red[4][8];
  for (int t = threadIdx.x; t < 3 * nlat; t += blockDim.x) s_lat[t] = lat[t];
  __syncthreads();
  const double tapi = 2.0 * alpha / sqrt(3.14159265358979323846);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double ev[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const double qi = q[i];
    double fi[3] = {0, 0, 0};
    for (int j = i + threadIdx.x; j < n; j += blockDim.x) {
      double d[3], r2;
      min_diff(box, pos + 3 * i, pos + 3 * j, d, r2);
      int excl = PM_EXCL_NONE;
      if (i == j) excl = PM_EXCL_FULL;
      else
        for (int u = xs_row[i]; u < xs_row[i + 1]; u++)
          if ((xs_part[u] & 0xffffff) == j) { excl = xs_part[u] >> 24; break; }
      double fj[3] = {0, 0, 0};
      for (int t = 0; t < nlat; t++) {
        if (excl == PM_EXCL_FULL) { excl = PM_EXCL_NONE; continue; }
        const double rij[3] = {d[0] + s_lat[3 * t], d[1] + s_lat[3 * t + 1], d[2] + s_lat[3 * t + 2]};
        const double rs = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
        if (rs > rc2) continue;
        const double r = sqrt(rs);
        double qq = qi * q[j];
        if (excl == PM_EXCL_MODIFIED) qq *= csf;
        const double rr = 1.0 / r, ar = alpha * r;
        const double e = qq * exp(-ar * ar);
        const double energy = poly5_ref(ar) * e * rr;
        const double force = (energy + tapi * e) * rr * rr;
        const double fij[3] = {rij[0] * force, rij[1] * force, rij[2] * force};
        fi[0] -= fij[0]; fi[1] -= fij[1]; fi[2] -= fij[2];
        fj[0] += fij[0]; fj[1] += fij[1]; fj[2] += fij[2];
        ev[0] += energy;
        if (want_virial) {
          // ScalarStructure::addVirial (base/ScalarStructure.cpp:41-63)
          ev[1] += fij[0] * rij[0]; ev[2] += fij[0] * rij[1]; ev[3] += fij[0] * rij[2];
          ev[4] += fij[1] * rij[1]; ev[5] += fij[1] * rij[2]; ev[6] += fij[2] * rij[2];
        }
        excl = PM_EXCL_NONE;
      }
      for (int k = 0; k < 3; k++) {
        const long long v = pm_force_fix(fj[k], box.err);
        if (v != 0) atomicAdd((unsigned long long*)(force_fix + 3LL * j + k), (unsigned long long)v);
      }
    }
    // the i force: fixed point per thread, integer sum over the block
    for (int k = 0; k < 3; k++) {
      long long v = pm_force_fix(fi[k], box.err);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0 && v != 0) atomicAdd((unsigned long long*)(force_fix + 3LL * i + k), (unsigned long long)v);
    }
  }
  for (int u = 0; u < 7; u++) {
    double x = ev[u];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) s_red[wv][u] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0 && epart) {
    double s[7];
    for (int u = 0; u < 7; u++) s[u] = s_red[0][u] + s_red[1][u] + s_red[2][u] + s_red[3][u];
    double* ep = epart + (long long)blockIdx.x * PM_NE;
    ep[PM_E_COUL] = s[0];
    for (int u = 1; u < PM_E_VIR; u++) ep[u] = 0.0;
    ep[PM_E_VIR + 0] = s[1]; ep[PM_E_VIR + 1] = s[2]; ep[PM_E_VIR + 2] = s[3];
    ep[PM_E_VIR + 3] = s[2]; ep[PM_E_VIR + 4] = s[4]; ep[PM_E_VIR + 5] = s[5];
    ep[PM_E_VIR + 6] = s[3]; ep[PM_E_VIR + 7] = s[5]; ep[PM_E_VIR + 8] = s[6];
  }
}

// Deterministic two-stage sum of nparts energy records: stage 1 sums fixed
// contiguous chunks (one block each), stage 2 sums the chunk totals in order
// and adds the constant terms.
#define PM_RED_BLOCKS 256
__device__ inline void reduce_energy1_body(const double* __restrict__ parts, int nparts, int chunk,
                                                        double* __restrict__ partial, int bx) {
  __shared__ double s[256];
  const int c = threadIdx.x % PM_NE;           // component
  const int r0 = threadIdx.x / PM_NE;          // 16 records in flight per component
  const int lo = bx * chunk, hi = min(nparts, lo + chunk);
  double acc = 0;
  for (int p = lo + r0; p < hi; p += 256 / PM_NE) acc += parts[(long long)p * PM_NE + c];
  s[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < PM_NE) {
    double t = 0;
    for (int u = 0; u < 256 / PM_NE; u++) t += s[u * PM_NE + threadIdx.x];
    partial[bx * PM_NE + threadIdx.x] = t;
  }
}
__global__ __launch_bounds__(256) void k_reduce_energy1(const double* __restrict__ parts, int nparts, int chunk,
                                                        double* __restrict__ partial) { reduce_energy1_body(parts, nparts, chunk, partial, blockIdx.x); }
struct ReduceEnergy1Args { const double* parts; int nparts; int chunk; double* partial; };
__global__ __launch_bounds__(256) void kb_reduce_energy1(const ReduceEnergy1Args* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const ReduceEnergy1Args& q = t[r];
  reduce_energy1_body(q.parts, q.nparts, q.chunk, q.partial, blockIdx.x - pre[r]);
}


// stage 2: 16 record groups per component summed with independent loads,
// then combined in a fixed order (nblocks <= PM_RED_BLOCKS)
__device__ inline void reduce_energy2_body(const double* __restrict__ partial, int nblocks,
                                                        double const_coul, double const_vir_diag,
                                                        double* __restrict__ out, int overwrite, int bx) {
  __shared__ double s[256];
  const int c = threadIdx.x % PM_NE, g = threadIdx.x / PM_NE;
  double v[PM_RED_BLOCKS / 16];
#pragma unroll
  for (int k = 0; k < PM_RED_BLOCKS / 16; k++) {
    const int b = g + 16 * k;
    v[k] = b < nblocks ? partial[b * PM_NE + c] : 0.0;
  }
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < PM_RED_BLOCKS / 16; k++) acc += v[k];
  s[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < PM_NE) {
    double t = 0.0;
    for (int u = 0; u < 16; u++) t += s[u * PM_NE + threadIdx.x];
    if (c == PM_E_COUL) t += const_coul;
    if (c == PM_E_VIR + 0 || c == PM_E_VIR + 4 || c == PM_E_VIR + 8) t += const_vir_diag;
    out[c] = overwrite ? t : out[c] + t;
  }
}
__global__ __launch_bounds__(256) void k_reduce_energy2(const double* __restrict__ partial, int nblocks,
                                                        double const_coul, double const_vir_diag,
                                                        double* __restrict__ out, int overwrite) { reduce_energy2_body(partial, nblocks, const_coul, const_vir_diag, out, overwrite, blockIdx.x); }
struct ReduceEnergy2Args { const double* partial; int nblocks; double const_coul; double const_vir_diag; double* out; int overwrite; };
__global__ __launch_bounds__(256) void kb_reduce_energy2(const ReduceEnergy2Args* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const ReduceEnergy2Args& q = t[r];
  reduce_energy2_body(q.partial, q.nblocks, q.const_coul, q.const_vir_diag, q.out, q.overwrite, blockIdx.x - pre[r]);
}


__global__ void k_zero_ll(long long* p, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}
__global__ void k_zero_i(int* p, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}
// clears the fixed-point forces and the cell counts of one evaluation in one launch
// force arrays, and the cell counts of the binning when it runs (*rebuild != 0
// or rebuild null)
struct ZeroFcArgs { const int* rebuild; long long* f; long long nf; long long* fs; long long nfs; int* cc; int nc; };
__device__ inline void zero_force_cells_body(const ZeroFcArgs& a, int bx) {
  long long i = (long long)bx * blockDim.x + threadIdx.x;
  if (i < a.nf) a.f[i] = 0;
  if (i < a.nfs) a.fs[i] = 0;
  if (i < a.nc && (!a.rebuild || *a.rebuild != 0)) a.cc[i] = 0;
}
__global__ void k_zero_force_cells(ZeroFcArgs a) { zero_force_cells_body(a, blockIdx.x); }
__global__ void k_zero_d(double* p, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.0;
}
__global__ void k_fix_to_double(int n3, const long long* __restrict__ f, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) out[i] = (double)f[i] * PM_FORCE_UNFIX;
}
__global__ void k_add_fix(int n3, const long long* __restrict__ f, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) out[i] += (double)f[i] * PM_FORCE_UNFIX;
}

// LeapfrogIntegrator::doHalfKickdoDrift / doKickdoDrift
// (integrators/LeapfrogIntegrator.cpp:46-104) and StandardIntegrator::doHalfKick
// (StandardIntegrator.cpp:83-96).  mode 0: v += F*h*0.5/m then x += v*h;
// mode 1: v += F*h/m then x += v*h; mode 2: v += F*hh/m (hh = 0.5*dt*INV_TIMEFACTOR).
// the next evaluation's displacement test and binning of atom i at its new
// position x (PmStepWork)
__device__ inline void step_work(const PmStepWork& w, int i, const double* x) {
  if (w.flag && pm_moved_past(w.to_fix, w.fix_to_A, x, w.x_build[i], w.lim2)) w.flag[0] = 1;
  if (w.flag && w.x_prune && pm_moved_past(w.to_fix, w.fix_to_A, x, w.x_prune[i], w.lim2_in)) w.flag[1] = 1;
  if (w.bin_count) pm_bin_atom(w.box, i, x, w.cell_of, w.key, w.rank, w.bin_count);
}

struct KickArgs { int n; int mode; double h; const double* mass; long long* f; double* v; double* x;
                 PmStepWork w; int zero_f; };
__device__ inline void kick_drift_body(const KickArgs& a, int bx) {
#pragma clang fp contract(off)
  int i = bx * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double m = a.mass[i];
  double xn[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double F = (double)a.f[3 * i + k] * PM_FORCE_UNFIX;
    if (a.zero_f) a.f[3 * i + k] = 0;
    double vv = a.v[3 * i + k];
    if (a.mode == 0) vv += F * a.h * 0.5 / m;
    else vv += F * a.h / m;
    a.v[3 * i + k] = vv;
    xn[k] = a.x[3 * i + k];
    if (a.mode != 2) {
      xn[k] += vv * a.h;
      a.x[3 * i + k] = xn[k];
    }
  }
  step_work(a.w, i, xn);
}
__global__ void k_kick_drift(KickArgs a) { kick_drift_body(a, blockIdx.x); }

// counter-based normal deviates (SplitMix64 hash + Box-Muller)
__device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double gauss(unsigned long long seed, unsigned long long counter, unsigned long long idx) {
  unsigned long long h1 = mix64(seed ^ mix64(counter * 0x100000001B3ull + idx * 2));
  unsigned long long h2 = mix64(h1 ^ 0x5851F42D4C957F2Dull);
  double u1 = ((h1 >> 11) + 1) * (1.0 / 9007199254740993.0);   // (0, 1]
  double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);         // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// LangevinLeapfrogIntegrator::doHalfKick (integrators/LangevinLeapfrogIntegrator.cpp:58-79)
// The noise counter and the bath temperature come from device memory
// (PmMdDevice), so a captured step replays with fresh deviates and a bath
// temperature can be changed between replays.
struct LangevinArgs {
  int n; double dt; double g; const double* mass; long long* f; double* v; unsigned long long seed;
  const PmMdDevice* md; int offset; int zero_f;
};
__device__ inline void langevin_kick_body(int n, double dt, double g, const double* __restrict__ mass,
                                          long long* __restrict__ f, double* __restrict__ v,
                                          unsigned long long seed, const PmMdDevice* __restrict__ md, int offset,
                                          int zero_f, int bx) {
  int i = bx * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long counter = md->counter + (unsigned long long)offset;
  const double temperature = md->temperature;
  const double fdt = (1.0 - exp(-0.5 * g * dt)) / g;
  const double vdt = exp(-0.5 * g * dt);
  const double ndt = sqrt((1.0 - exp(-g * dt)) / (2.0 * g));
  const double fc = 2 * 0.001987191 * temperature * g;
  const double m = mass[i];
  const double s = sqrt(fc / m);
  for (int k = 0; k < 3; k++) {
    const double F = (double)f[3 * i + k] * PM_FORCE_UNFIX;
    if (zero_f) f[3 * i + k] = 0;
    const double xi = gauss(seed, counter, 3ull * i + k);
    v[3 * i + k] = v[3 * i + k] * vdt + F * fdt / m + xi * s * ndt;
  }
}
__global__ void k_langevin_kick(LangevinArgs a) {
  langevin_kick_body(a.n, a.dt, a.g, a.mass, a.f, a.v, a.seed, a.md, a.offset, a.zero_f, blockIdx.x);
}

// v *= factor (velocities carried to a new bath temperature)
__global__ void k_scale(int n3, double factor, double* __restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) v[i] *= factor;
}

// advances the noise counter by the two half kicks of a step
struct AdvanceArgs { PmMdDevice* md; };
__device__ inline void md_advance_body(PmMdDevice* md, int bx) {
  if (threadIdx.x == 0 && bx == 0) md->counter += 2;
}
__global__ void k_md_advance(AdvanceArgs a) { md_advance_body(a.md, blockIdx.x); }

struct DriftArgs { int n; double h; const double* v; double* x; PmStepWork w; };
__device__ inline void drift_body(const DriftArgs& a, int bx) {
#pragma clang fp contract(off)
  int i = bx * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double xn[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    xn[k] = a.x[3 * i + k] + a.h * a.v[3 * i + k];   // Vector3DBlock::intoWeightedAdd
    a.x[3 * i + k] = xn[k];
  }
  step_work(a.w, i, xn);
}
__global__ void k_drift(DriftArgs a) { drift_body(a, blockIdx.x); }

// batched launches over replicas (argument tables in device memory)
__global__ void kb_kick_drift(const KickArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const KickArgs& a = t[r];
  kick_drift_body(a, blockIdx.x - pre[r]);
}
__global__ void kb_langevin_kick(const LangevinArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const LangevinArgs& a = t[r];
  langevin_kick_body(a.n, a.dt, a.g, a.mass, a.f, a.v, a.seed, a.md, a.offset, a.zero_f, blockIdx.x - pre[r]);
}
__global__ void kb_drift(const DriftArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const DriftArgs& a = t[r];
  drift_body(a, blockIdx.x - pre[r]);
}
__global__ void kb_md_advance(const AdvanceArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  md_advance_body(t[r].md, blockIdx.x - pre[r]);
}
__global__ void kb_zero_force_cells(const ZeroFcArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  zero_force_cells_body(t[r], blockIdx.x - pre[r]);
}

inline dim3 grid_for(long long n, int b) { return dim3((unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1)); }

}  // namespace

void pm_launch_correction(hipStream_t s, const PmBox& box, int n, const double* pos, const double* q,
                          const int* row, const int* partner, const int* cls, double alpha, double csf,
                          int want_virial, long long* force_fix, int ntab, const int* tab_a,
                          const int* tab_b, const int* tab_c, double* epart, int nblk) {
  if (n > 0)
    hipLaunchKernelGGL(k_corr_force, grid_for(n, 256), dim3(256), 0, s, box, n, pos, q, row, partner, cls, alpha,
                       csf, force_fix);
  if (nblk > 0)
    hipLaunchKernelGGL(k_corr_energy, dim3(nblk), dim3(256), 0, s, box, pos, q, ntab, tab_a, tab_b, tab_c, alpha,
                       csf, want_virial, epart);
}

void pm_launch_corr_adjoint(hipStream_t s, const PmBox& box, int n, const double* pos, const double* q,
                            const int* row, const int* partner, const int* cls, double alpha, double csf, int mode,
                            const double* v, double* out) {
  if (n > 0)
    hipLaunchKernelGGL(k_corr_adjoint, grid_for(n, 256), dim3(256), 0, s, box, n, pos, q, row, partner, cls, alpha,
                       csf, mode, v, out);
}

void pm_launch_ewald_lattice(hipStream_t s, const PmBox& box, int n, const double* pos, const double* q,
                             const int* xs_row, const int* xs_part, const double* lat, int nlat, double rc2,
                             double alpha, double csf, int want_virial, long long* force_fix, double* epart,
                             int nblk) {
  if (n > 0 && nblk > 0)
    hipLaunchKernelGGL(k_ewald_lattice, dim3(nblk), dim3(256), 0, s, box, n, pos, q, xs_row, xs_part, lat, nlat, rc2,
                       alpha, csf, want_virial, force_fix, epart);
}

void pm_launch_add_fix_to(hipStream_t s, int m, const long long* fixed, double* out) {
  if (m > 0) hipLaunchKernelGGL(k_add_fix_to, grid_for(m, 256), dim3(256), 0, s, m, fixed, out);
}

int pm_reduce_scratch_records(int nparts) {
  (void)nparts;
  return PM_RED_BLOCKS;
}

void pm_launch_reduce_energy(hipStream_t s, const double* parts, int nparts, double constant_coul,
                             double constant_vir_diag, double* out, int overwrite, double* scratch) {
  int nb = nparts < PM_RED_BLOCKS ? (nparts > 0 ? nparts : 1) : PM_RED_BLOCKS;
  int chunk = (nparts + nb - 1) / nb;
  if (chunk < 1) chunk = 1;
  if (pm_recording()) {
    const ReduceEnergy1Args a1{parts, nparts, chunk, scratch};
    const ReduceEnergy2Args a2{scratch, nb, constant_coul, constant_vir_diag, out, overwrite};
    pm_record(PM_REC_REDUCE1, 0, (unsigned)nb, 256, 0, &a1, sizeof(a1));
    pm_record(PM_REC_REDUCE2, 0, 1, 256, 0, &a2, sizeof(a2));
    return;
  }
  hipLaunchKernelGGL(k_reduce_energy1, dim3(nb), dim3(256), 0, s, parts, nparts, chunk, scratch);
  hipLaunchKernelGGL(k_reduce_energy2, dim3(1), dim3(256), 0, s, scratch, nb, constant_coul, constant_vir_diag, out,
                     overwrite);
}

void pm_launch_zero_ll(hipStream_t s, long long* p, long long count) {
  if (count > 0) hipLaunchKernelGGL(k_zero_ll, grid_for(count, 256), dim3(256), 0, s, p, count);
}
void pm_launch_zero_i(hipStream_t s, int* p, long long count) {
  if (count > 0) hipLaunchKernelGGL(k_zero_i, grid_for(count, 256), dim3(256), 0, s, p, count);
}
void pm_launch_zero_force_cells(hipStream_t s, const int* rebuild, long long* f, long long nf, long long* fs,
                                long long nfs, int* cc, int nc) {
  long long m = nf > nc ? nf : nc;
  if (nfs > m) m = nfs;
  const ZeroFcArgs a{rebuild, f, nf, fs, nfs, cc, nc};
  if (pm_recording()) {
    pm_record(PM_REC_ZERO_FORCE_CELLS, 0, m > 0 ? grid_for(m, 256).x : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (m > 0) hipLaunchKernelGGL(k_zero_force_cells, grid_for(m, 256), dim3(256), 0, s, a);
}
void pm_launch_zero_d(hipStream_t s, double* p, long long count) {
  if (count > 0) hipLaunchKernelGGL(k_zero_d, grid_for(count, 256), dim3(256), 0, s, p, count);
}
void pm_launch_fix_to_double(hipStream_t s, int n3, const long long* f, double* out) {
  if (n3 > 0) hipLaunchKernelGGL(k_fix_to_double, grid_for(n3, 256), dim3(256), 0, s, n3, f, out);
}
void pm_launch_add_fix(hipStream_t s, int n3, const long long* f, double* out) {
  if (n3 > 0) hipLaunchKernelGGL(k_add_fix, grid_for(n3, 256), dim3(256), 0, s, n3, f, out);
}
void pm_launch_kick(hipStream_t s, int n, int mode, double h, const double* mass, long long* f,
                    double* v, double* x, const PmStepWork* w, int zero_f) {
  const KickArgs a{n, mode, h, mass, f, v, x, w ? *w : PmStepWork{}, zero_f};
  if (pm_recording()) {
    pm_record(PM_REC_KICK_DRIFT, 0, n > 0 ? grid_for(n, 256).x : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (n > 0) hipLaunchKernelGGL(k_kick_drift, grid_for(n, 256), dim3(256), 0, s, a);
}
void pm_launch_langevin_kick(hipStream_t s, int n, double dt_akma, double gamma_akma, const double* mass,
                             long long* f, double* v, unsigned long long seed, const PmMdDevice* md,
                             int offset, int zero_f) {
  const LangevinArgs a{n, dt_akma, gamma_akma, mass, f, v, seed, md, offset, zero_f};
  if (pm_recording()) {
    pm_record(PM_REC_LANGEVIN_KICK, 0, n > 0 ? grid_for(n, 256).x : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (n > 0) hipLaunchKernelGGL(k_langevin_kick, grid_for(n, 256), dim3(256), 0, s, a);
}
void pm_launch_scale(hipStream_t s, int n3, double factor, double* v) {
  if (n3 > 0) hipLaunchKernelGGL(k_scale, grid_for(n3, 256), dim3(256), 0, s, n3, factor, v);
}
void pm_launch_md_advance(hipStream_t s, PmMdDevice* md) {
  const AdvanceArgs a{md};
  if (pm_recording()) {
    pm_record(PM_REC_MD_ADVANCE, 0, 1, 64, 0, &a, sizeof(a));
    return;
  }
  hipLaunchKernelGGL(k_md_advance, dim3(1), dim3(64), 0, s, a);
}
void pm_launch_drift(hipStream_t s, int n, double h, const double* v, double* x, const PmStepWork* w) {
  const DriftArgs a{n, h, v, x, w ? *w : PmStepWork{}};
  if (pm_recording()) {
    pm_record(PM_REC_DRIFT, 0, n > 0 ? grid_for(n, 256).x : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (n > 0) hipLaunchKernelGGL(k_drift, grid_for(n, 256), dim3(256), 0, s, a);
}

bool pm_launch_batched_md(hipStream_t s, int kind, int variant, const void* t, const int* pre, int nrep,
                          unsigned total, unsigned threads, unsigned lds) {
  (void)variant; (void)lds;
  const dim3 g(total), b(threads);
  switch (kind) {
    case PM_REC_KICK_DRIFT: hipLaunchKernelGGL(kb_kick_drift, g, b, 0, s, (const KickArgs*)t, pre, nrep); return true;
    case PM_REC_LANGEVIN_KICK: hipLaunchKernelGGL(kb_langevin_kick, g, b, 0, s, (const LangevinArgs*)t, pre, nrep); return true;
    case PM_REC_DRIFT: hipLaunchKernelGGL(kb_drift, g, b, 0, s, (const DriftArgs*)t, pre, nrep); return true;
    case PM_REC_MD_ADVANCE: hipLaunchKernelGGL(kb_md_advance, g, b, 0, s, (const AdvanceArgs*)t, pre, nrep); return true;
    case PM_REC_ZERO_FORCE_CELLS: hipLaunchKernelGGL(kb_zero_force_cells, g, b, 0, s, (const ZeroFcArgs*)t, pre, nrep); return true;
    case PM_REC_REDUCE1: hipLaunchKernelGGL(kb_reduce_energy1, g, b, 0, s, (const ReduceEnergy1Args*)t, pre, nrep); return true;
    case PM_REC_REDUCE2: hipLaunchKernelGGL(kb_reduce_energy2, g, b, 0, s, (const ReduceEnergy2Args*)t, pre, nrep); return true;
    default: return false;
  }
}
