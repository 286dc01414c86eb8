// pm_bonded.hip -- FP64 bonded system forces on gfx950: bonds, angles,
// dihedrals and impropers.
//
// BondSystemForce::calcBond (forces/BondSystemForce.h:111-145),
// AngleSystemForce::calcAngle (forces/AngleSystemForce.h:113-200) and
// MTorsionSystemForce::calcTorsion (forces/MTorsionSystemForce.h:34-153) for
// DihedralSystemForce and ImproperSystemForce.  Forces
// are computed once per term and added to the atoms as 64-bit fixed-point
// atomics (deterministic: integer sums do not depend on their order); the
// PME gather's per-atom tail (pm_terms.h bonded_atom_force) is the other
// route.  Energies and virials are summed per term in fixed-order block
// reductions.
#include "pm_internal.h"
#include "pm_terms.h"

#include <math.h>

namespace {

// One thread per term (bond, angle, dihedral, improper; blocks of each kind
// in that order, kind of block bx from the prefix off[]): each term is
// evaluated once, its forces go to its atoms as 64-bit fixed-point atomic
// adds (integer addition is associative, so the sums do not depend on the
// order of the atomics), and with want_energy the block writes one energy /
// virial record.  a.n == 0 leaves the forces alone (the PME gather added
// them per atom).
struct BondedBlocks { int off[5]; };
__device__ inline void add_fix(long long* f, int i, const double* v, double sg, int* err) {
  for (int c = 0; c < 3; c++)
    atomicAdd(reinterpret_cast<unsigned long long*>(f + 3 * (long long)i + c),
              (unsigned long long)pm_force_fix(sg * v[c], err));
}
// KIND is a template parameter so that each kind's kernel holds only its
// own registers (the torsion's FP64 trigonometry would otherwise set the
// occupancy of all four)
// Forces are first summed per block in LDS over the block's atom window
// (the atoms of consecutive terms of a PSF lie close together) with 64-bit
// LDS atomics, then added to the global fixed-point forces once per atom and
// component; a block whose atoms span more than PM_BONDED_WINDOW atoms adds
// straight to global memory.  Both are integer sums: the result does not
// depend on the order.  Impropers are sparse (a protein has about one per six
// atoms), so a block of them spans more atoms: their window is twice as wide.
struct BondedShared {
  double red[4][7];
  int lo[4], hi[4];
};
template <int KIND>
__device__ inline void bonded_terms_body(PmBox box, PmBondedLaunch a, BondedBlocks bl, int bx, BondedShared& sh,
                                         unsigned long long* s_f) {
  constexpr int kind = KIND;
  constexpr int W = kind == 3 ? 2048 : 1024;   // atom window summed in LDS (s_f holds 3 W)
  double (*s_red)[7] = sh.red;
  int* s_lo = sh.lo;
  int* s_hi = sh.hi;
  constexpr int NA = kind == 0 ? 2 : kind == 1 ? 3 : 4;
  bx += bl.off[kind];           // the block's energy record among all kinds
  const int t = (bx - bl.off[kind]) * blockDim.x + threadIdx.x;
  int count;
  if constexpr (kind == 0) count = a.nb;
  else if constexpr (kind == 1) count = a.na;
  else count = a.nt[kind - 2];
  const bool have = t < count;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  int at[4] = {0, 0, 0, 0};
  double g[4][3];
  if (have) {
    double f1[3], f2[3] = {0, 0, 0}, f3[3] = {0, 0, 0}, e, r12[3], r23[3] = {0, 0, 0}, r34[3] = {0, 0, 0};
    if constexpr (kind == 0) {
      bond_term(box, a, t, f1, e, r12);
      at[0] = a.b_a[t]; at[1] = a.b_b[t];
      for (int c = 0; c < 3; c++) { g[0][c] = f1[c]; g[1][c] = -f1[c]; }
    } else if constexpr (kind == 1) {
      angle_term(box, a, t, f1, f2, f3, e, r12, r34);
      at[0] = a.a_a[t]; at[1] = a.a_b[t]; at[2] = a.a_c[t];
      for (int c = 0; c < 3; c++) { g[0][c] = f1[c]; g[1][c] = f2[c]; g[2][c] = f3[c]; }
      // the angle's virial is f1 r12 + f3 r32 (AngleSystemForce.h:180-198)
      f2[0] = f2[1] = f2[2] = 0.0;
    } else {
      torsion_term(box, a, kind - 2, t, f1, f2, f3, e, r12, r23, r34);
      const int4 q = a.t_atoms[kind - 2][t];
      at[0] = q.x; at[1] = q.y; at[2] = q.z; at[3] = q.w;
      for (int c = 0; c < 3; c++) {
        g[0][c] = f1[c]; g[1][c] = f2[c] - f1[c]; g[2][c] = f3[c] - f2[c]; g[3][c] = -f3[c];
      }
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
  if (a.n > 0) {   // uniform over the launch
    // the block's atom window
    int lo = 0x7fffffff, hi = -1;
    if (have)
      for (int j = 0; j < NA; j++) { lo = min(lo, at[j]); hi = max(hi, at[j]); }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo = min(lo, __shfl_xor(lo, off, 64));
      hi = max(hi, __shfl_xor(hi, off, 64));
    }
    if (lane == 0) { s_lo[wv] = lo; s_hi[wv] = hi; }
    __syncthreads();
    lo = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
    hi = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
    const int span = hi - lo + 1;
    const bool local = hi >= lo && span <= W;
    if (local) {
      for (int k = threadIdx.x; k < 3 * span; k += blockDim.x) s_f[k] = 0ull;
      __syncthreads();
      if (have)
        for (int j = 0; j < NA; j++)
          for (int c = 0; c < 3; c++)
            atomicAdd(&s_f[3 * (at[j] - lo) + c], (unsigned long long)pm_force_fix(g[j][c], box.err));
      __syncthreads();
      for (int k = threadIdx.x; k < 3 * span; k += blockDim.x) {
        const unsigned long long x = s_f[k];
        if (x) atomicAdd(reinterpret_cast<unsigned long long*>(a.force_fix + 3 * (long long)lo) + k, x);
      }
    } else if (have) {
      for (int j = 0; j < NA; j++) add_fix(a.force_fix, at[j], g[j], 1.0, box.err);
    }
  }
  if (!a.want_energy) return;   // uniform over the launch: no thread reaches the barrier
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
// Launch GROUP 0 runs the bond and angle blocks (one launch: their
// registers are alike), GROUP 2 the dihedrals, GROUP 3 the impropers.
template <int GROUP>
__device__ inline void bonded_group_body(PmBox box, const PmBondedLaunch& a, const BondedBlocks& bl, int b) {
  __shared__ BondedShared sh;
  __shared__ unsigned long long s_f[3 * (GROUP == 3 ? 2048 : 1024)];
  if constexpr (GROUP == 0) {
    const int nb0 = bl.off[1] - bl.off[0];
    if (b < nb0) bonded_terms_body<0>(box, a, bl, b, sh, s_f);
    else bonded_terms_body<1>(box, a, bl, b - nb0, sh, s_f);
  } else {
    bonded_terms_body<GROUP>(box, a, bl, b, sh, s_f);
  }
}
template <int GROUP>
__global__ __launch_bounds__(256) void k_bonded_terms(PmBox box, PmBondedLaunch a, BondedBlocks bl) {
  bonded_group_body<GROUP>(box, a, bl, blockIdx.x);
}
struct BondedTermArgs { PmBox box; PmBondedLaunch a; BondedBlocks bl; };
template <int GROUP>
__global__ __launch_bounds__(256) void kb_bonded_terms(const BondedTermArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const BondedTermArgs& q = t[r];
  bonded_group_body<GROUP>(q.box, q.a, q.bl, blockIdx.x - pre[r]);
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
  // nothing to do when neither forces (a.n > 0) nor energies are wanted
  const bool work = a.n > 0 || a.want_energy;
  const BondedTermArgs q{box, a, bl};
  const int groups[3] = {0, 2, 3};
  const int gblocks[3] = {nblk[0] + nblk[1], nblk[2], nblk[3]};
  for (int g = 0; g < 3; g++) {
    const unsigned blocks = work ? (unsigned)gblocks[g] : 0u;
    if (pm_recording()) {
      pm_record(PM_REC_BONDED_FORCE, groups[g], blocks, 256, 0, &q, sizeof(q));
      continue;
    }
    if (blocks == 0) continue;
    if (g == 0) hipLaunchKernelGGL(k_bonded_terms<0>, dim3(blocks), dim3(256), 0, s, box, a, bl);
    else if (g == 1) hipLaunchKernelGGL(k_bonded_terms<2>, dim3(blocks), dim3(256), 0, s, box, a, bl);
    else hipLaunchKernelGGL(k_bonded_terms<3>, dim3(blocks), dim3(256), 0, s, box, a, bl);
  }
}

bool pm_launch_batched_bonded(hipStream_t s, int kind, int variant, const void* t, const int* pre, int nrep,
                              unsigned total, unsigned threads, unsigned lds) {
  (void)lds;
  if (kind != PM_REC_BONDED_FORCE) return false;
  const BondedTermArgs* q = (const BondedTermArgs*)t;
  if (variant == 0) hipLaunchKernelGGL(kb_bonded_terms<0>, dim3(total), dim3(threads), 0, s, q, pre, nrep);
  else if (variant == 2) hipLaunchKernelGGL(kb_bonded_terms<2>, dim3(total), dim3(threads), 0, s, q, pre, nrep);
  else hipLaunchKernelGGL(kb_bonded_terms<3>, dim3(total), dim3(threads), 0, s, q, pre, nrep);
  return true;
}
