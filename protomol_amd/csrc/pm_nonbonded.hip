// pm_nonbonded.hip -- cell binning, slot layout, neighbour lists and the
// cutoff pair kernels (gfx950, wave64).
//
// Binning reproduces topology/buildCellLists.cpp:11-38 bit for bit (FP64,
// no contraction); the pair kernels evaluate OneAtomPair / OneAtomPairTwo
// (forces/OneAtomPair.h:127-188, forces/OneAtomPairTwo.h:131-199) once for
// every unordered pair within the switching cutoff and add both atoms' forces
// as 64-bit fixed-point integers (order-independent, deterministic run to run).
#include "pm_internal.h"
#include "pm_step.h"

#include <math.h>
#include <type_traits>

namespace {

// ------------------------------------------------------------------ binning
__device__ inline void bin_body(PmBox box, int n, const double* __restrict__ pos, int* __restrict__ cell_of,
                      int* __restrict__ cell_count, int* __restrict__ rank, float* __restrict__ key,
                      const int* __restrict__ rebuild, int bx) {
  int i = bx * blockDim.x + threadIdx.x;
  if (i >= n || (rebuild && *rebuild == 0)) return;
  pm_bin_atom(box, i, pos + 3 * i, cell_of, key, rank, cell_count);
}
__global__ void k_bin(PmBox box, int n, const double* __restrict__ pos, int* __restrict__ cell_of,
                      int* __restrict__ cell_count, int* __restrict__ rank, float* __restrict__ key,
                      const int* __restrict__ rebuild) { bin_body(box, n, pos, cell_of, cell_count, rank, key, rebuild, blockIdx.x); }
struct BinArgs { PmBox box; int n; const double* pos; int* cell_of; int* cell_count; int* rank; float* key; const int* rebuild; };
__global__ void kb_bin(const BinArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const BinArgs& q = t[r];
  bin_body(q.box, q.n, q.pos, q.cell_of, q.cell_count, q.rank, q.key, q.rebuild, blockIdx.x - pre[r]);
}


// Exclusive scan over a block of 1024 threads (16 waves): each thread's
// exclusive prefix of v; the block total in `total`.
__device__ inline int block_excl_scan(int v, int* s_w, int& total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(inc, off, 64);
    if (lane >= off) inc += u;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  int before = 0, tot = 0;
  for (int u = 0; u < 16; u++) {
    const int x = s_w[u];
    if (u < w) before += x;
    tot += x;
  }
  __syncthreads();   // s_w is reused by the next scan
  total = tot;
  return before + inc - v;
}

// Slot layout of one chunk of whole cell columns (nz consecutive cells share
// x and y; the chunk starts at a column's first cell) whose atom counts are
// in s_c[0, m).  The atoms of a column take consecutive slots, cell after
// cell, and the column's slots are padded to a multiple of m_cl after its
// last cell, so clusters of m_cl slots never straddle two columns and only
// the columns' ends carry padding.  Thread t owns cells 8t .. 8t+7: on
// return v[k] holds the cell's atom count, ra[k] / rs[k] its chunk-local
// exclusive atom / slot prefix and ps[k] its slots (atoms plus, at a
// column's last cell, the column's padding); ta / ts the chunk's totals.
// s_c is overwritten.
__device__ inline void scan_chunk(int* s_c, int* s_w, int m, int nz, int m_cl, int (&v)[8], int (&ra)[8],
                                  int (&ps)[8], int (&rs)[8], int& ta, int& ts) {
  const int t = threadIdx.x;
  int sa = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int idx = 8 * t + k;
    v[k] = idx < m ? s_c[idx] : 0;
    sa += v[k];
  }
  int acc = block_excl_scan(sa, s_w, ta);
  // the atom prefixes replace the counts (each thread rewrites only its own cells)
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int idx = 8 * t + k;
    ra[k] = acc;
    if (idx < m) s_c[idx] = acc;
    acc += v[k];
  }
  __syncthreads();
  int ss = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int idx = 8 * t + k;
    int p = v[k];
    if (idx < m && idx % nz == nz - 1) {
      const int col = ra[k] + v[k] - s_c[idx - nz + 1];   // the column's atoms
      p += (m_cl - col % m_cl) % m_cl;
    }
    ps[k] = p;
    ss += p;
  }
  acc = block_excl_scan(ss, s_w, ts);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    rs[k] = acc;
    acc += ps[k];
  }
}

// Whole columns per chunk of at most PM_SCAN_CHUNK cells (nz <= PM_SCAN_CHUNK).
#define PM_SCAN_CHUNK 8192
__host__ __device__ inline int scan_chunk_cells(int nz) { return PM_SCAN_CHUNK / nz * nz; }

// Cell starts (atoms), cell slots and every cluster's cell, in one workgroup
// that walks the cell grid a chunk of whole columns at a time, staged in LDS.
// The cell a cluster belongs to is the cell of its first slot.
__device__ inline void scan_cells_body(int ncells, int nz, int* __restrict__ cnt, int* __restrict__ cnt_out,
                                       int* __restrict__ start, int* __restrict__ slot,
                                       int* __restrict__ counts, int* __restrict__ cluster_cell,
                                       int ncl_max, int m_cl, const int* __restrict__ rebuild, int bx) {
  if (rebuild && *rebuild == 0) {
    // uniform over the block: before any barrier.  The binning counts start
    // from zero at the next binning (the MD position update bins every step)
    for (int c = threadIdx.x; c < ncells; c += blockDim.x) cnt[c] = 0;
    return;
  }
  __shared__ int s_c[PM_SCAN_CHUNK];
  __shared__ int s_w[16];
  const int t = threadIdx.x;
  const int ch = scan_chunk_cells(nz);
  int carry_a = 0, carry_s = 0;
  for (int base = 0; base < ncells; base += ch) {
    const int m = min(ch, ncells - base);
    {
      // eight independent loads in flight per thread (clamped, then masked)
      int tmp[8];
#pragma unroll
      for (int k = 0; k < 8; k++) tmp[k] = cnt[base + min(t + 1024 * k, m - 1)];
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (t + 1024 * k < m) {
          s_c[t + 1024 * k] = tmp[k];
          cnt_out[base + t + 1024 * k] = tmp[k];   // the layout's cell counts
          cnt[base + t + 1024 * k] = 0;            // binning counts cleared for the next binning
        }
    }
    __syncthreads();
    int v[8], ra[8], ps[8], rs[8], ta, ts;
    scan_chunk(s_c, s_w, m, nz, m_cl, v, ra, ps, rs, ta, ts);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int idx = 8 * t + k;
      if (idx < m) {
        const int c = base + idx, s0 = carry_s + rs[k];
        start[c] = carry_a + ra[k];
        slot[c] = s0;
        // clusters whose first slot is one of this cell's atoms
        for (int u = (s0 + m_cl - 1) / m_cl; u * m_cl < s0 + v[k] && u < ncl_max; u++) cluster_cell[u] = c;
      }
    }
    carry_a += ta;
    carry_s += ts;
    __syncthreads();   // s_c is refilled by the next chunk
  }
  if (t == 0) { counts[0] = carry_s; counts[1] = carry_s / m_cl; counts[2] = carry_a; }
}
__global__ __launch_bounds__(1024) void k_scan_cells(int ncells, int nz, int* __restrict__ cnt, int* __restrict__ cnt_out,
                                                     int* __restrict__ start, int* __restrict__ slot,
                                                     int* __restrict__ counts, int* __restrict__ cluster_cell,
                                                     int ncl_max, int m_cl, const int* __restrict__ rebuild) { scan_cells_body(ncells, nz, cnt, cnt_out, start, slot, counts, cluster_cell, ncl_max, m_cl, rebuild, blockIdx.x); }
struct ScanCellsArgs { int ncells; int nz; int* cnt; int* cnt_out; int* start; int* slot; int* counts; int* cluster_cell; int ncl_max; int m_cl; const int* rebuild; };
__global__ __launch_bounds__(1024) void kb_scan_cells(const ScanCellsArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const ScanCellsArgs& q = t[r];
  scan_cells_body(q.ncells, q.nz, q.cnt, q.cnt_out, q.start, q.slot, q.counts, q.cluster_cell, q.ncl_max, q.m_cl, q.rebuild, blockIdx.x - pre[r]);
}


// Large cell grids: the same layout split over one block per chunk.  Pass 1
// copies each chunk's counts to the layout, clears the binning counts and
// writes the chunk's atom and slot totals; pass 2 adds the totals of the
// chunks before it (a fixed-order sum) and writes the chunk's starts, slots,
// cluster cells and, in the last block, the totals.  Integer sums over the
// same chunks: the result is the single-block layout's.
__global__ __launch_bounds__(1024) void k_scan_chunks(int ncells, int nz, int* __restrict__ cnt, int* __restrict__ cnt_out,
                                                      int m_cl, int* __restrict__ aux, const int* __restrict__ rebuild) {
  const int ch = scan_chunk_cells(nz);
  const int base = blockIdx.x * ch;
  const int m = min(ch, ncells - base);
  const int t = threadIdx.x;
  if (rebuild && *rebuild == 0) {
    for (int k = t; k < m; k += blockDim.x) cnt[base + k] = 0;
    return;
  }
  __shared__ int s_c[PM_SCAN_CHUNK];
  __shared__ int s_w[16];
  for (int k = t; k < m; k += blockDim.x) {
    const int v = cnt[base + k];
    cnt_out[base + k] = v;
    cnt[base + k] = 0;
    s_c[k] = v;
  }
  __syncthreads();
  int v[8], ra[8], ps[8], rs[8], ta, ts;
  scan_chunk(s_c, s_w, m, nz, m_cl, v, ra, ps, rs, ta, ts);
  if (t == 0) {
    aux[2 * blockIdx.x] = ta;
    aux[2 * blockIdx.x + 1] = ts;
  }
}
__global__ __launch_bounds__(1024) void k_scan_finish(int ncells, int nz, const int* __restrict__ cnt_out,
                                                      int* __restrict__ start, int* __restrict__ slot,
                                                      int* __restrict__ counts, int* __restrict__ cluster_cell,
                                                      int ncl_max, int m_cl, const int* __restrict__ aux,
                                                      const int* __restrict__ rebuild) {
  if (rebuild && *rebuild == 0) return;
  __shared__ int s_c[PM_SCAN_CHUNK];
  __shared__ int s_w[16];
  __shared__ int s_carry[2];
  const int ch = scan_chunk_cells(nz);
  const int base = blockIdx.x * ch;
  const int m = min(ch, ncells - base);
  const int t = threadIdx.x;
  if (t == 0) {
    int ca = 0, cs = 0;
    for (int b = 0; b < (int)blockIdx.x; b++) { ca += aux[2 * b]; cs += aux[2 * b + 1]; }
    s_carry[0] = ca;
    s_carry[1] = cs;
  }
  for (int k = t; k < m; k += blockDim.x) s_c[k] = cnt_out[base + k];
  __syncthreads();
  const int carry_a = s_carry[0], carry_s = s_carry[1];
  int v[8], ra[8], ps[8], rs[8], ta, ts;
  scan_chunk(s_c, s_w, m, nz, m_cl, v, ra, ps, rs, ta, ts);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int idx = 8 * t + k;
    if (idx < m) {
      const int c = base + idx, s0 = carry_s + rs[k];
      start[c] = carry_a + ra[k];
      slot[c] = s0;
      for (int u = (s0 + m_cl - 1) / m_cl; u * m_cl < s0 + v[k] && u < ncl_max; u++) cluster_cell[u] = c;
    }
  }
  if (t == 0 && blockIdx.x == gridDim.x - 1) {
    const int tot_a = carry_a + ta, tot_s = carry_s + ts;
    counts[0] = tot_s; counts[1] = tot_s / m_cl; counts[2] = tot_a;
  }
}

__device__ inline void scatter_body(int n, const int* __restrict__ cell_of, const int* __restrict__ rank,
                          const int* __restrict__ cell_start, int* __restrict__ cell_atoms,
                          const int* __restrict__ rebuild, int bx) {
  int i = bx * blockDim.x + threadIdx.x;
  if (i >= n || (rebuild && *rebuild == 0)) return;
  cell_atoms[cell_start[cell_of[i]] + rank[i]] = i;
}
__global__ void k_scatter(int n, const int* __restrict__ cell_of, const int* __restrict__ rank,
                          const int* __restrict__ cell_start, int* __restrict__ cell_atoms,
                          const int* __restrict__ rebuild) { scatter_body(n, cell_of, rank, cell_start, cell_atoms, rebuild, blockIdx.x); }
struct ScatterArgs { int n; const int* cell_of; const int* rank; const int* cell_start; int* cell_atoms; const int* rebuild; };
__global__ void kb_scatter(const ScatterArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const ScatterArgs& q = t[r];
  scatter_body(q.n, q.cell_of, q.rank, q.cell_start, q.cell_atoms, q.rebuild, blockIdx.x - pre[r]);
}


__device__ inline unsigned int to_fix(double x, double s) {
  // periodic 32-bit fixed point: round(x * 2^32 / L) mod 2^32; positions
  // that have none (check_position flags them) map to 0
  double t = x * s;
  if (!(__builtin_fabs(t) < 9.0e18)) t = 0.0;
  long long v = __double2ll_rn(t);
  return (unsigned int)v;
}

__device__ inline unsigned long long to_fix64(double x, double s) {
  // round(x * 2^64 / L) mod 2^64 from round(x * 2^32 / L) split in two
  double t = x * s;                       // units of 2^-32 box
  if (!(__builtin_fabs(t) < 9.0e18)) t = 0.0;
  double hi = __builtin_floor(t);
  double lo = (t - hi) * 4294967296.0;    // fraction in units of 2^-64 box
  unsigned long long h = (unsigned long long)(long long)hi;
  unsigned long long l = (unsigned long long)__double2ll_rn(lo);
  return (h << 32) + l;
}

// A position whose periodic fixed point cannot be formed (not finite, or more
// than 2^31 box lengths from the origin) raises PM_FLAG_NONFINITE: such an
// atom would otherwise drop out of every pair silently.
__device__ inline void check_position(const PmBox& box, const double* x) {
  const bool ok = __builtin_fabs(x[0] * box.to_fix[0]) < 9.0e18 && __builtin_fabs(x[1] * box.to_fix[1]) < 9.0e18 &&
                  __builtin_fabs(x[2] * box.to_fix[2]) < 9.0e18;
  if (!ok) pm_raise(box.err, PM_FLAG_NONFINITE);
}

// One wave per cell: every atom's slot is its rank in (z, index) order within
// the cell (ties impossible: indices are distinct), so the layout does not
// depend on the arrival order of k_scatter.  Cells holding more than
// PM_SORT_MAX atoms (collapsed or non-finite coordinates) keep arrival order.
__device__ inline void fill_slots_body(PmBox box, int ncells, const double* __restrict__ pos,
                             const double* __restrict__ q, const int* __restrict__ type,
                             const int* __restrict__ cell_count, const int* __restrict__ cell_start,
                             const int* __restrict__ cell_slot, const int* __restrict__ cell_atoms,
                             const float* __restrict__ key, int4* __restrict__ slot_pos,
                             longlong4* __restrict__ slot_pos64, int* __restrict__ slot_meta,
                             int* __restrict__ atom_of_slot, int dummy_type, int dummy_slot, int m_cl,
                             const int* __restrict__ rebuild, int4* __restrict__ x_build,
                             int4* __restrict__ x_prune, int bx, int nb) {
  const int lane = threadIdx.x & 63;
  const int c = bx * 4 + (threadIdx.x >> 6);
  if (rebuild && *rebuild == 0) {
    // layout kept (list reuse): current positions and charges into the
    // slots, with the arithmetic of the build below, one thread per slot up
    // to the last cell's last atom (padding slots keep their place: they are
    // masked); when the inner lists are pruned in this evaluation, also the
    // positions they are pruned at
    const bool prune = x_prune && rebuild[1] != 0;
    const int send = ncells > 0 ? cell_slot[ncells - 1] + cell_count[ncells - 1] : 0;
    for (int sl = bx * blockDim.x + threadIdx.x; sl < send; sl += nb * blockDim.x) {
      const int ai = atom_of_slot[sl];
      if (ai < 0) continue;
      check_position(box, pos + 3 * ai);
      int4 p;
      p.x = (int)to_fix(pos[3 * ai], box.to_fix[0]);
      p.y = (int)to_fix(pos[3 * ai + 1], box.to_fix[1]);
      p.z = (int)to_fix(pos[3 * ai + 2], box.to_fix[2]);
      p.w = __float_as_int((float)q[ai]);
      slot_pos[sl] = p;
      if (prune) x_prune[ai] = make_int4(p.x, p.y, p.z, 1);
      if (slot_pos64) {
        longlong4 p8;
        p8.x = (long long)to_fix64(pos[3 * ai], box.to_fix[0]);
        p8.y = (long long)to_fix64(pos[3 * ai + 1], box.to_fix[1]);
        p8.z = (long long)to_fix64(pos[3 * ai + 2], box.to_fix[2]);
        p8.w = __double_as_longlong(q[ai]);
        slot_pos64[sl] = p8;
      }
    }
    return;
  }
  if (c == 0 && lane == 0) {
    // the shared padding slot of the pair lists
    slot_pos[dummy_slot] = make_int4(0, 0, 0, 0);
    if (slot_pos64) slot_pos64[dummy_slot] = make_longlong4(0, 0, 0, 0);
    slot_meta[dummy_slot] = dummy_type | (0x7fffff << 8);
    atom_of_slot[dummy_slot] = -1;
  }
  if (c >= ncells) return;
  const int k = cell_count[c];
  const int* a = cell_atoms + cell_start[c];
  const int s0 = cell_slot[c];
  const bool sort = k <= PM_SORT_MAX;
  for (int i = lane; i < k; i += 64) {
    const int ai = a[i];
    const float ki = key[ai];
    int rank = i;
    if (sort) {
      rank = 0;
      for (int j = 0; j < k; j++) {
        const int aj = a[j];
        const float kj = key[aj];
        rank += (kj < ki || (kj == ki && aj < ai)) ? 1 : 0;
      }
    }
    check_position(box, pos + 3 * ai);
    int4 p;
    p.x = (int)to_fix(pos[3 * ai], box.to_fix[0]);
    p.y = (int)to_fix(pos[3 * ai + 1], box.to_fix[1]);
    p.z = (int)to_fix(pos[3 * ai + 2], box.to_fix[2]);
    p.w = __float_as_int((float)q[ai]);
    slot_pos[s0 + rank] = p;
    if (slot_pos64) {
      // 64-bit periodic fixed point for the FP64 kernels
      longlong4 p8;
      p8.x = (long long)to_fix64(pos[3 * ai], box.to_fix[0]);
      p8.y = (long long)to_fix64(pos[3 * ai + 1], box.to_fix[1]);
      p8.z = (long long)to_fix64(pos[3 * ai + 2], box.to_fix[2]);
      p8.w = __double_as_longlong(q[ai]);
      slot_pos64[s0 + rank] = p8;
    }
    slot_meta[s0 + rank] = type[ai] | (ai << 8);
    atom_of_slot[s0 + rank] = ai;
    if (x_build) x_build[ai] = make_int4(p.x, p.y, p.z, 1);   // the position the layout was built at
    if (x_prune) x_prune[ai] = make_int4(p.x, p.y, p.z, 1);   // the inner lists are pruned at it too
  }
  // padding slots after a column's last cell, up to the next multiple of
  // m_cl: never used as j atoms, masked as i atoms, skipped by PME; placed
  // at the cell's lower corner
  if (c % box.n[2] != box.n[2] - 1) return;
  const int e = s0 + k, pe = (e + m_cl - 1) / m_cl * m_cl;
  const int u = e + lane;
  if (u < pe) {
    const int cz = c % box.n[2], cy = (c / box.n[2]) % box.n[1], cx = c / (box.n[1] * box.n[2]);
    int4 p;
    p.x = (int)(unsigned int)(long long)(box.cell0_fix[0] + cx * box.real_fix[0]);
    p.y = (int)(unsigned int)(long long)(box.cell0_fix[1] + cy * box.real_fix[1]);
    p.z = (int)(unsigned int)(long long)(box.cell0_fix[2] + cz * box.real_fix[2]);
    p.w = 0;
    slot_pos[u] = p;
    if (slot_pos64) {
      // the same point in 64-bit fixed point
      longlong4 p8;
      p8.x = (long long)((unsigned long long)(unsigned int)p.x << 32);
      p8.y = (long long)((unsigned long long)(unsigned int)p.y << 32);
      p8.z = (long long)((unsigned long long)(unsigned int)p.z << 32);
      p8.w = 0;
      slot_pos64[u] = p8;
    }
    slot_meta[u] = dummy_type | (0x7fffff << 8);
    atom_of_slot[u] = -1;
  }
}
__global__ __launch_bounds__(256) void k_fill_slots(PmBox box, int ncells, const double* __restrict__ pos,
                             const double* __restrict__ q, const int* __restrict__ type,
                             const int* __restrict__ cell_count, const int* __restrict__ cell_start,
                             const int* __restrict__ cell_slot, const int* __restrict__ cell_atoms,
                             const float* __restrict__ key, int4* __restrict__ slot_pos,
                             longlong4* __restrict__ slot_pos64, int* __restrict__ slot_meta,
                             int* __restrict__ atom_of_slot, int dummy_type, int dummy_slot, int m_cl,
                             const int* __restrict__ rebuild, int4* __restrict__ x_build,
                             int4* __restrict__ x_prune) { fill_slots_body(box, ncells, pos, q, type, cell_count, cell_start, cell_slot, cell_atoms, key, slot_pos, slot_pos64, slot_meta, atom_of_slot, dummy_type, dummy_slot, m_cl, rebuild, x_build, x_prune, blockIdx.x, gridDim.x); }
struct FillSlotsArgs { PmBox box; int ncells; const double* pos; const double* q; const int* type; const int* cell_count; const int* cell_start; const int* cell_slot; const int* cell_atoms; const float* key; int4* slot_pos; longlong4* slot_pos64; int* slot_meta; int* atom_of_slot; int dummy_type; int dummy_slot; int m_cl; const int* rebuild; int4* x_build; int4* x_prune; };
__global__ __launch_bounds__(256) void kb_fill_slots(const FillSlotsArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const FillSlotsArgs& q = t[r];
  fill_slots_body(q.box, q.ncells, q.pos, q.q, q.type, q.cell_count, q.cell_start, q.cell_slot, q.cell_atoms, q.key, q.slot_pos, q.slot_pos64, q.slot_meta, q.atom_of_slot, q.dummy_type, q.dummy_slot, q.m_cl, q.rebuild, q.x_build, q.x_prune, blockIdx.x - pre[r], pre[r + 1] - pre[r]);
}


// List reuse: rebuild (flag = 1) when an atom has moved more than half the
// list skin from where the layout was built (or the layout is invalid: w != 1).
// Every thread that finds one writes the same value, so no atomics are needed.
__device__ inline void check_disp_body(PmBox box, int n, const double* __restrict__ pos, const int4* __restrict__ x_build,
                             float lim2, int* __restrict__ flag, const int4* __restrict__ x_prune, float lim2_in,
                             int bx) {
  const int i = bx * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (pm_moved_past(box.to_fix, box.fix_to_A, pos + 3 * i, x_build[i], lim2)) flag[0] = 1;
  if (x_prune && pm_moved_past(box.to_fix, box.fix_to_A, pos + 3 * i, x_prune[i], lim2_in)) flag[1] = 1;
}
__global__ void k_check_disp(PmBox box, int n, const double* __restrict__ pos, const int4* __restrict__ x_build,
                             float lim2, int* __restrict__ flag, const int4* __restrict__ x_prune, float lim2_in) {
  check_disp_body(box, n, pos, x_build, lim2, flag, x_prune, lim2_in, blockIdx.x);
}
struct CheckDispArgs { PmBox box; int n; const double* pos; const int4* x_build; float lim2; int* flag; const int4* x_prune; float lim2_in; };
__global__ void kb_check_disp(const CheckDispArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const CheckDispArgs& q = t[r];
  check_disp_body(q.box, q.n, q.pos, q.x_build, q.lim2, q.flag, q.x_prune, q.lim2_in, blockIdx.x - pre[r]);
}


__global__ void k_set_flag(int* flag, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = v;
}

__global__ void k_invalidate_build(int n, int4* __restrict__ x_build) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x_build[i] = make_int4(0, 0, 0, 0);
}

// periodic distance (fixed-point units) from a point at offset t = x - lo
// (mod 2^32) to the interval [lo, lo + w]
__device__ inline float cell_gap(unsigned int t, unsigned int w) {
  if (t <= w) return 0.f;
  const unsigned int up = t - w, down = 0u - t;
  return (float)(up < down ? up : down);
}

// ------------------------------------------------------------ wave helpers
__device__ inline int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

template <int M>
__device__ inline float wave_min_m(float v) {
#pragma unroll
  for (int off = 1; off < M; off <<= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
template <int M>
__device__ inline float wave_max_m(float v) {
#pragma unroll
  for (int off = 1; off < M; off <<= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
template <int M>
__device__ inline int wave_imin_m(int v) {
#pragma unroll
  for (int off = 1; off < M; off <<= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}
template <int M>
__device__ inline int wave_imax_m(int v) {
#pragma unroll
  for (int off = 1; off < M; off <<= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline int wrap_cell(int c, int d, int n) {
  int v = (c + d) % n;
  return v < 0 ? v + n : v;
}

// ------------------------------------------------------------- pair lists
// One wave per i-cluster (m consecutive slots of one cell column): the
// candidates are the atoms, from the cluster's first slot on, of the cells of
// the stencil's columns that meet the list cutoff around the cluster; those
// that pass the list test are appended to the cluster's list.  The pair
// kernel keeps a pair only when the j slot follows the i slot, so every
// unordered pair within the cutoff is listed exactly once.  Bit 31 marks an
// entry that some cluster atom excludes or modifies; its masks (FULL bit ii,
// MODIFIED bit 16 + ii) go to jmask.  BB: 1 the bounding-box test, 0 the
// exact per-atom test, 2 the one a.bbox_filter names (batched launches, whose
// replicas may differ)
template <int M, int BB>
__device__ inline void list_body(PmBox box, PmNbLaunch a, float cut2_list, int bx) {
  const bool bbox = BB == 2 ? a.bbox_filter != 0 : BB == 1;
  if (a.rebuild && *a.rebuild == 0) return;   // lists kept from the last build
  extern __shared__ int s_dyn[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nst = a.nstencil;
  // up to two slot ranges per stencil column
  int* s_start = s_dyn + wave * (6 * nst + 1 + 2 * a.xs_tcap);
  int* s_cnt = s_start + 2 * nst;
  int* s_pref = s_cnt + 2 * nst;     // 2 nst + 1 entries
  int* s_tri = s_pref + 2 * nst + 1; // xs_tcap entries: partner | ii << 24 | class << 28
  int* s_srt = s_tri + a.xs_tcap; // the same, sorted by partner
  const int cl = bx * 4 + wave;
  const int ncl = a.counts[1];
  const bool active = cl < ncl;  // every wave reaches every barrier below
  const int cell = active ? a.cluster_cell[cl] : 0;
  const int n0 = box.n[0], n1 = box.n[1], n2 = box.n[2];
  const int cz = cell % n2, cy = (cell / n2) % n1, cx = cell / (n1 * n2);

  // i-cluster: anchor, bounding box and exclusion partners
  const int4 anchor = active ? a.slot_pos[cl * M] : make_int4(0, 0, 0, 0);
  const float fx = (float)box.fix_to_A[0], fy = (float)box.fix_to_A[1], fz = (float)box.fix_to_A[2];
  float lox = 1e30f, loy = 1e30f, loz = 1e30f, hix = -1e30f, hiy = -1e30f, hiz = -1e30f;
  int xbeg = 0, deg = 0;
  if (active && lane < M) {
    int4 p = a.slot_pos[cl * M + lane];
    int m = a.slot_meta[cl * M + lane];
    int o = (m >> 8) & 0x7fffff;
    if (o != 0x7fffff) {
      float dx = (float)(int)(p.x - anchor.x) * fx;
      float dy = (float)(int)(p.y - anchor.y) * fy;
      float dz = (float)(int)(p.z - anchor.z) * fz;
      lox = hix = dx; loy = hiy = dy; loz = hiz = dz;
      xbeg = a.xs_row[o];
      deg = a.xs_row[o + 1] - xbeg;
    }
  }
  lox = wave_min_m<M>(lox); loy = wave_min_m<M>(loy); loz = wave_min_m<M>(loz);
  hix = wave_max_m<M>(hix); hiy = wave_max_m<M>(hiy); hiz = wave_max_m<M>(hiz);
  lox = __shfl(lox, 0, 64); loy = __shfl(loy, 0, 64); loz = __shfl(loz, 0, 64);
  hix = __shfl(hix, 0, 64); hiy = __shfl(hiy, 0, 64); hiz = __shfl(hiz, 0, 64);
  // partners of the cluster's atoms into LDS (deg <= xs_tcap / M per atom)
  const int tinc = wave_incl_scan(deg, lane);
  const int ntri = __shfl(tinc, 63, 64);
  int pmin = 0x7fffffff, pmax = -1;
  for (int u = 0; u < deg; u++) {
    const int w = a.xs_part[xbeg + u];
    const int o = w & 0xffffff;
    s_tri[tinc - deg + u] = o | (lane << 24) | ((w >> 24) << 28);
    pmin = min(pmin, o);
    pmax = max(pmax, o);
  }
  pmin = __shfl(wave_imin_m<M>(pmin), 0, 64);
  pmax = __shfl(wave_imax_m<M>(pmax), 0, 64);

  // exact test: the cluster's positions, and which of its slots hold atoms
  // (padding slots, placed at a cell corner, must not keep an entry)
  int qx[M], qy[M], qz[M];
  unsigned int qreal = 0;
  if (active && !bbox) {
#pragma unroll
    for (int ii = 0; ii < M; ii++) {
      const int4 q = a.slot_pos[cl * M + ii];
      if (((a.slot_meta[cl * M + ii] >> 8) & 0x7fffff) != 0x7fffff) qreal |= 1u << ii;
      qx[ii] = q.x;
      qy[ii] = q.y;
      qz[ii] = q.z;
    }
  }
  // the cells of a column lie at j * hz - oa (j = cell - cz) relative to the
  // anchor, which the binning put in cell cz
  const float hz = (float)(box.real_fix[2] * box.fix_to_A[2]);
  const float rhz = 1.f / hz;
  const float oa = (float)(((double)((unsigned int)anchor.z - (unsigned int)(long long)box.cell0_fix[2]) -
                            cz * box.real_fix[2]) * box.fix_to_A[2]);
  // the bounding-box test holds for clusters no taller than the cell
  // diagonal (the box is then large enough for minimal images relative to
  // the anchor); a taller cluster (a sparse column) keeps every candidate
  const float rx = (float)(box.real_fix[0] * box.fix_to_A[0]), ry = (float)(box.real_fix[1] * box.fix_to_A[1]);
  const bool loose = bbox && hiz - loz > sqrtf(rx * rx + ry * ry + hz * hz);
  const int first = cl * M;
  // stencil columns (one per lane): per column the z interval, relative to
  // the anchor, within the list cutoff of a cluster atom (exact test; of the
  // bounding box in the bounding-box test), its cells as one or two
  // contiguous slot ranges (two when it wraps round the box), clipped to the
  // slots from this cluster's first on (half list); the non-empty ranges are
  // compacted in stencil order
  int nact = 0;
  for (int s0 = 0; s0 < nst; s0 += 64) {
    const int s = s0 + lane;
    int st0 = 0, ct0 = 0, st1 = 0, ct1 = 0;
    if (active && s < nst) {
      const int4 d = a.stencil[s];
      const int wx = wrap_cell(cx, d.x, n0), wy = wrap_cell(cy, d.y, n1);
      const unsigned int bx = (unsigned int)(long long)(box.cell0_fix[0] + wx * box.real_fix[0]);
      const unsigned int by = (unsigned int)(long long)(box.cell0_fix[1] + wy * box.real_fix[1]);
      float zl = 1e30f, zh = -1e30f;
      if (bbox) {
        // gap between the column's rectangle (padded by 1e-3 A) and the cluster's bounding box
        const float clx = (float)(int)(bx - (unsigned int)anchor.x) * fx - 1e-3f;
        const float cly = (float)(int)(by - (unsigned int)anchor.y) * fy - 1e-3f;
        const float chx = clx + rx + 2e-3f, chy = cly + ry + 2e-3f;
        const float gx = fmaxf(0.f, fmaxf(clx - hix, lox - chx));
        const float gy = fmaxf(0.f, fmaxf(cly - hiy, loy - chy));
        const float r2 = cut2_list - gx * gx - gy * gy;
        if (r2 > 0.f) {
          const float r = sqrtf(r2);
          zl = loz - r;
          zh = hiz + r;
        }
      } else {
        // per axis, the periodic distance from each of the cluster's atoms
        // to the column's interval (padded by 1e-3 A) bounds the pair
        // kernel's minimal-image difference from below
        const unsigned int wdx = (unsigned int)(long long)box.real_fix[0] + 1u;
        const unsigned int wdy = (unsigned int)(long long)box.real_fix[1] + 1u;
#pragma unroll
        for (int ii = 0; ii < M; ii++) {
          const float gx = fmaxf(cell_gap((unsigned int)qx[ii] - bx, wdx) * fx - 1e-3f, 0.f);
          const float gy = fmaxf(cell_gap((unsigned int)qy[ii] - by, wdy) * fy - 1e-3f, 0.f);
          const float r2 = cut2_list - gx * gx - gy * gy;
          if (((qreal >> ii) & 1u) && r2 > 0.f) {
            const float r = sqrtf(r2);
            const float dz = (float)(int)((unsigned int)qz[ii] - (unsigned int)anchor.z) * fz;
            zl = fminf(zl, dz - r);
            zh = fmaxf(zh, dz + r);
          }
        }
      }
      if (zl <= zh) {
        // cells j0 .. j1 (relative to cz) meet [zl, zh], each padded by 1e-3 A
        const int j0 = (int)floorf((zl - 1e-3f + oa) * rhz), j1 = (int)floorf((zh + 1e-3f + oa) * rhz);
        int za = 0, zb = n2 - 1, zd = -1;   // cell ranges [za, zb] and [0, zd]
        if (j1 - j0 + 1 < n2) {
          za = wrap_cell(cz, j0, n2);
          zb = wrap_cell(cz, j1, n2);
          if (za > zb) { zd = zb; zb = n2 - 1; }
        }
        const int col = (wx * n1 + wy) * n2;
        int e = a.cell_slot[col + zb] + a.cell_count[col + zb];
        st0 = max(a.cell_slot[col + za], first);
        ct0 = max(e - st0, 0);
        if (zd >= 0) {
          e = a.cell_slot[col + zd] + a.cell_count[col + zd];
          st1 = max(a.cell_slot[col], first);
          ct1 = max(e - st1, 0);
        }
      }
    }
    const unsigned long long bal0 = __ballot(ct0 > 0), bal1 = __ballot(ct1 > 0);
    const int b0 = __builtin_amdgcn_mbcnt_hi((unsigned)(bal0 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal0, 0));
    if (ct0 > 0) { s_start[nact + b0] = st0; s_cnt[nact + b0] = ct0; }
    nact += __popcll(bal0);
    const int b1 = __builtin_amdgcn_mbcnt_hi((unsigned)(bal1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal1, 0));
    if (ct1 > 0) { s_start[nact + b1] = st1; s_cnt[nact + b1] = ct1; }
    nact += __popcll(bal1);
  }
  __syncthreads();
  int carry = 0;
  for (int s0 = 0; s0 < nact; s0 += 64) {
    const int s = s0 + lane;
    const int v = (active && s < nact) ? s_cnt[s] : 0;
    const int inc = wave_incl_scan(v, lane);
    if (active && s < nact) s_pref[s] = carry + inc - v;
    carry += __shfl(inc, 63, 64);
  }
  if (active && lane == 0) s_pref[nact] = carry;
  // exclusion partners sorted by partner index (rank sort, ties by position)
  for (int t = lane; t < ntri; t += 64) {
    const int w = s_tri[t], key = w & 0xffffff;
    int rank = 0;
    for (int u = 0; u < ntri; u++) {
      const int ku = s_tri[u] & 0xffffff;
      rank += (ku < key || (ku == key && u < t)) ? 1 : 0;
    }
    s_srt[rank] = w;
  }
  __syncthreads();
  if (!active) return;
  const int total = carry;
  int nsteps = 0;
  while ((1 << nsteps) < ntri + 1) nsteps++;

  int* out = a.jlist + (long long)cl * a.lcap;
  unsigned int* outm = a.jmask + (long long)cl * a.lcap;
  int count = 0;
  // candidates beyond PM_CAND_FACTOR * lcap mean a degenerate cell layout:
  // stop scanning and report it as a list overflow
  const int scan = min(total, PM_CAND_FACTOR * a.lcap);
  if (scan < total && lane == 0) atomicOr(a.overflow, PM_FLAG_CAPACITY);
  int scur = 0;   // stencil entry holding candidate k0 (uniform)
  for (int k0 = 0; k0 < scan; k0 += 64) {
    const int k = k0 + lane;
    // owner entry of every lane: walk the boundaries inside this batch
    while (scur + 1 < nact && s_pref[scur + 1] <= k0) scur++;
    int s = scur;
    for (int t = scur + 1; t < nact && s_pref[t] <= k0 + 63; t++)
      if (k >= s_pref[t]) s = t;
    bool pass = false;
    int sj = 0, o = 0;
    if (k < scan) {
      sj = s_start[s] + (k - s_pref[s]);
      const int4 p = a.slot_pos[sj];
      pass = true;
      if (bbox) {
        const float dx = (float)(int)(p.x - anchor.x) * fx;
        const float dy = (float)(int)(p.y - anchor.y) * fy;
        const float dz = (float)(int)(p.z - anchor.z) * fz;
        const float gx = fmaxf(0.f, fmaxf(lox - dx, dx - hix));
        const float gy = fmaxf(0.f, fmaxf(loy - dy, dy - hiy));
        const float gz = fmaxf(0.f, fmaxf(loz - dz, dz - hiz));
        pass = loose || gx * gx + gy * gy + gz * gz < cut2_list;
      } else {
        // box too small for the bounding-box test: keep the entry when one
        // of the cluster's atoms lies within the list cutoff, by the pair
        // kernel's own minimal-image difference (the cluster's positions
        // are wave-uniform loads)
        pass = false;
#pragma unroll
        for (int ii = 0; ii < M; ii++) {
          const float dx = (float)(int)(p.x - qx[ii]) * fx;
          const float dy = (float)(int)(p.y - qy[ii]) * fy;
          const float dz = (float)(int)(p.z - qz[ii]) * fz;
          pass = pass || (((qreal >> ii) & 1u) && dx * dx + dy * dy + dz * dz < cut2_list);
        }
      }
      o = (a.slot_meta[sj] >> 8) & 0x7fffff;
    }
    // exclusion masks of the entry against the cluster's atoms: binary search
    // of the sorted partners, then every entry with this partner
    int xm = 0;
    const bool pre = pass && o >= pmin && o <= pmax;
    if (__ballot(pre) != 0ull) {
      int lo = 0, hi = ntri;
      for (int it = 0; it < nsteps; it++) {
        const int mid = (lo + hi) >> 1;
        if (lo < hi) {
          if ((s_srt[mid] & 0xffffff) < o) lo = mid + 1; else hi = mid;
        }
      }
      if (pre) {
        for (int t = lo; t < ntri; t++) {
          const int w = s_srt[t];
          if ((w & 0xffffff) != o) break;
          xm |= ((w >> 28) == PM_EXCL_FULL ? 1 : 65536) << ((w >> 24) & 15);
        }
      }
    }
    const unsigned long long bal = __ballot(pass);
    const int before = __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
    const int idx = count + before;
    if (pass && idx < a.lcap) {
      out[idx] = sj | (xm != 0 ? (int)0x80000000 : 0);
      if (xm != 0) outm[idx] = (unsigned int)xm;
    }
    count += __popcll(bal);
  }
  if (lane == 0) {
    if (count > a.lcap) { atomicOr(a.overflow, PM_FLAG_CAPACITY); count = a.lcap; }
    a.jcount[cl] = count;
  }
}
template <int M, int BB>
__global__ __launch_bounds__(256) void k_list(PmBox box, PmNbLaunch a, float cut2_list) { list_body<M, BB>(box, a, cut2_list, blockIdx.x); }
struct ListArgs { PmBox box; PmNbLaunch a; float cut2_list; };
template <int M>
__global__ __launch_bounds__(256) void kb_list(const ListArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const ListArgs& q = t[r];
  list_body<M, 2>(q.box, q.a, q.cut2_list, blockIdx.x - pre[r]);
}


// list statistics: total entries and issued lanes (pair evaluations, masked
// ones included) over all clusters: a pair kernel iteration takes jw entries
// and issues 64 * mult lanes (rows of M lanes: jw = 64 / M, mult 1; a lane
// per entry: jw = 64, mult M)
__global__ __launch_bounds__(256) void k_list_stats(const int* __restrict__ jcount, const int* __restrict__ counts,
                                                    unsigned long long* __restrict__ out, int jw, int mult) {
  __shared__ unsigned long long s0[256], s1[256];
  const int ncl = counts[1];
  unsigned long long e = 0, l = 0;
  for (int c0 = 0; c0 < ncl; c0 += 256 * 8) {
    int v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int c = c0 + 256 * k + threadIdx.x;
      v[k] = c < ncl ? jcount[c] : 0;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      e += v[k];
      l += (unsigned long long)((v[k] + jw - 1) / jw) * 64 * mult;
    }
  }
  s0[threadIdx.x] = e;
  s1[threadIdx.x] = l;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) { s0[threadIdx.x] += s0[threadIdx.x + off]; s1[threadIdx.x] += s1[threadIdx.x + off]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = s0[0]; out[1] = s1[0]; }
}

// slot forces of the pair kernels into the per-atom fixed-point forces
__device__ inline void fold_slots_body(const int* __restrict__ counts, const int* __restrict__ atom_of_slot,
                             long long* __restrict__ force_slot, long long* __restrict__ force_fix,
                             int* __restrict__ flag, unsigned long long* __restrict__ builds, int clear, int bx) {
  if (flag && bx == 0 && threadIdx.x == 0) {
    // end of an evaluation that reuses lists (every reader of the flags ran
    // before): count a build or a prune of the inner lists, clear the flags
    if (flag[0]) builds[0] += 1ull;
    else if (flag[1]) builds[1] += 1ull;
    flag[0] = 0;
    flag[1] = 0;
  }
  const int sl = bx * blockDim.x + threadIdx.x;
  if (sl >= counts[0]) return;
  const int at = atom_of_slot[sl];
  if (at < 0) return;
  force_fix[3LL * at] += force_slot[3LL * sl];
  force_fix[3LL * at + 1] += force_slot[3LL * sl + 1];
  force_fix[3LL * at + 2] += force_slot[3LL * sl + 2];
  if (clear) {
    // the next evaluation's pair kernels start from zero (only real slots
    // ever receive pair forces)
    force_slot[3LL * sl] = 0;
    force_slot[3LL * sl + 1] = 0;
    force_slot[3LL * sl + 2] = 0;
  }
}
__global__ void k_fold_slots(const int* __restrict__ counts, const int* __restrict__ atom_of_slot,
                             long long* __restrict__ force_slot, long long* __restrict__ force_fix,
                             int* __restrict__ flag, unsigned long long* __restrict__ builds, int clear) {
  fold_slots_body(counts, atom_of_slot, force_slot, force_fix, flag, builds, clear, blockIdx.x);
}
struct FoldSlotsArgs { const int* counts; const int* atom_of_slot; long long* force_slot; long long* force_fix; int* flag; unsigned long long* builds; int clear; };
__global__ void kb_fold_slots(const FoldSlotsArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const FoldSlotsArgs& q = t[r];
  fold_slots_body(q.counts, q.atom_of_slot, q.force_slot, q.force_fix, q.flag, q.builds, q.clear, blockIdx.x - pre[r]);
}


// ----------------------------------------------------- inner-list pruning
// One wave per i-cluster, on evaluations that build the lists or find an
// atom moved past half the inner skin (either flag; otherwise it returns at
// once): the entries of the full list of which some lane is a pair the pair
// kernel keeps (real i atom, j slot after the i slot) within the pair cutoff +
// the inner skin at the current positions, in list order, with their mask
// words, into the inner list.  The pair kernel then walks only these until
// the next prune: a pair within the cutoff later was within cutoff + inner
// skin here while every atom stays within half the inner skin of its position
// now.  Distances from the 32-bit fixed-point slots in both precisions (the
// cutoff is padded as the list cutoff is).
template <int M>
__device__ inline void prune_body(const PmBox& box, const PmNbLaunch& a, int bx) {
  if (a.rebuild[0] == 0 && a.rebuild[1] == 0) return;
  constexpr int JW = 64 / M;
  const int lane = threadIdx.x & 63;
  const int cl = bx * 4 + (threadIdx.x >> 6);
  if (cl >= a.counts[1]) return;
  const int ii = lane % M, js = lane / M;
  const int si = cl * M + ii;
  const bool ireal = ((a.slot_meta[si] >> 8) & 0x7fffff) != 0x7fffff;
  const int4 pi = a.slot_pos[si];
  const float sx = (float)box.fix_to_A[0], sy = (float)box.fix_to_A[1], sz = (float)box.fix_to_A[2];
  const long long lbase = (long long)cl * a.lcap;
  const int* __restrict__ lst = a.jlist + lbase;
  const int cnt = a.jcount[cl];
  int icnt = 0;   // wave-uniform
  for (int base = 0; base < cnt; base += JW) {
    const int k = base + js;
    const bool valid = k < cnt;
    const int e = valid ? lst[k] : a.dummy_slot;
    const int sj = e & 0x7fffffff;
    const int4 pj = a.slot_pos[sj];
    const float dx = (float)(int)(pj.x - pi.x) * sx;
    const float dy = (float)(int)(pj.y - pi.y) * sy;
    const float dz = (float)(int)(pj.z - pi.z) * sz;
    const bool kp = valid && ireal && sj > si && dx * dx + dy * dy + dz * dz <= a.cut2_prune;
    unsigned long long t = __ballot(kp);
    // one bit per row of M lanes, at the row's first lane (shifts of at most
    // M - 1 in total never carry a bit into the next row's first lane)
    if constexpr (M == 8) { t |= t >> 4; t |= t >> 2; t |= t >> 1; t &= 0x0101010101010101ull; }
    else { t |= t >> 2; t |= t >> 1; t &= 0x1111111111111111ull; }
    const int rowbit = js * M;
    if (ii == 0 && ((t >> rowbit) & 1ull)) {
      const int pos = icnt + __popcll(t & ((1ull << rowbit) - 1ull));
      a.ilist[lbase + pos] = e;
      if (e < 0) a.imask[lbase + pos] = a.jmask[lbase + k];
    }
    icnt += __popcll(t);
  }
  if (lane == 0) a.icount[cl] = icnt;
}
template <int M>
__global__ __launch_bounds__(256) void k_prune(PmBox box, PmNbLaunch a) { prune_body<M>(box, a, blockIdx.x); }
struct PruneArgs { PmBox box; PmNbLaunch a; };
template <int M>
__global__ __launch_bounds__(256) void kb_prune(const PruneArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  prune_body<M>(t[r].box, t[r].a, blockIdx.x - pre[r]);
}

// ------------------------------------------------------------ pair kernel
template <typename T> struct Prec;
template <> struct Prec<float> {
  static __device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
  static __device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ inline float expneg(float x) {  // exp(-x), x >= 0
    return __builtin_amdgcn_exp2f(-x * 1.4426950408889634f);
  }
  static __device__ inline float sqrt_(float x) { return __builtin_sqrtf(x); }
};
template <> struct Prec<double> {
  static __device__ inline double rsq(double x) { return 1.0 / sqrt(x); }
  static __device__ inline double rcp(double x) { return 1.0 / x; }
  static __device__ inline double expneg(double x) { return exp(-x); }
  static __device__ inline double sqrt_(double x) { return sqrt(x); }
};

// LJ A and B of a pair: the type table (LennardJonesParameterTable) or, when
// per-atom parameters are set, the CHARMM combination of the two atoms'
// (R, sqrt(eps)) as frontend/buildTopology.cpp:643-664 fills the table; the
// 1-4 values for MODIFIED pairs.
template <typename T, typename LJ4>
__device__ inline void lj_ab(const PmNbLaunch& a, const LJ4* s_lj, int ti, int mj, int nt1, int oi, int excl,
                             T& A, T& B) {
  const bool per_atom = sizeof(T) == 4 ? a.atom_lj_f != nullptr : a.atom_lj_d != nullptr;
  if (per_atom) {
    // padding and dummy slots carry the atom index 0x7fffff; their lanes are
    // masked, so they read atom 0's parameters instead of past the array
    const int oj0 = (mj >> 8) & 0x7fffff;
    const int ia = oi == 0x7fffff ? 0 : oi, ja = oj0 == 0x7fffff ? 0 : oj0;
    LJ4 pi_, pj_;
    if constexpr (sizeof(T) == 4) { pi_ = a.atom_lj_f[ia]; pj_ = a.atom_lj_f[ja]; }
    else { pi_ = a.atom_lj_d[ia]; pj_ = a.atom_lj_d[ja]; }
    const bool m = excl == PM_EXCL_MODIFIED;
    const T rs = m ? (T)pi_.z + (T)pj_.z : (T)pi_.x + (T)pj_.x;
    const T ee = m ? (T)pi_.w * (T)pj_.w : (T)pi_.y * (T)pj_.y;
    const T r2 = rs * rs, r6 = r2 * r2 * r2;
    A = r6 * r6 * ee;
    B = T(2) * r6 * ee;
  } else {
    const LJ4 ab = s_lj[ti * nt1 + (mj & 0xff)];
    A = excl == PM_EXCL_MODIFIED ? (T)ab.z : (T)ab.x;
    B = excl == PM_EXCL_MODIFIED ? (T)ab.w : (T)ab.y;
  }
}

// C1 / C2 / Cutoff switching functions (topology/C1SwitchingFunction.h:64-75,
// C2SwitchingFunction.h:64-82, CutoffSwitchingFunction.h:64-68).
template <typename T, int SWF>
__device__ inline void switching(T r2, T dist, T cut2, T on2, T sw1, T sw2, T sw3, T c15_1, T c05_3,
                                 T c15_3, T& v, T& d) {
  if constexpr (SWF == PM_SWF_C2) {
    T c2 = cut2 - r2;
    T c4 = c2 * (sw2 + T(2) * r2);
    bool in = r2 >= on2;
    v = in ? sw1 * (c2 * c4) : T(1);
    d = in ? sw3 * (c2 * c2 - c4) : T(0);
    if (r2 > cut2) { v = T(0); d = T(0); }
  } else if constexpr (SWF == PM_SWF_C1) {
    v = T(1) - dist * (c15_1 - r2 * c05_3);
    d = -c15_1 * Prec<T>::rcp(dist) + dist * c15_3;
    if (r2 > cut2) { v = T(0); d = T(0); }
  } else {
    v = r2 > cut2 ? T(0) : T(1);
    d = T(0);
  }
}

// force = force * S - energy * dS, energy = energy * S; the Cutoff function
// (S = 0 or 1, dS = 0) reduces to a select with the same result
template <typename T, int SWF>
__device__ inline void apply_switch(T r2, T dist, T cut2, T on2, T sw1, T sw2, T sw3, T c15_1, T c05_3, T c15_3,
                                    T& f, T& e) {
  if constexpr (SWF == PM_SWF_CUTOFF) {
    if (r2 > cut2) { f = T(0); e = T(0); }
  } else {
    T v, d;
    switching<T, SWF>(r2, dist, cut2, on2, sw1, sw2, sw3, c15_1, c05_3, c15_3, v, d);
    f = f * v - e * d;
    e = e * v;
  }
}


// Sum over the M lanes of a row (lanes M*r .. M*r+M-1) with DPP moves: quad
// xor 1, quad xor 2 (M = 4), then half-row mirror (M = 8).  Every lane
// of the row ends with the same value, added in the same order on every run.
template <int CTRL>
__device__ inline float dpp_f(float v) {
  // old = 0 (the identity of the sum): the move folds into v_add_f32_dpp
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ inline double dpp_f(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int M, typename T>
__device__ inline T row_sum(T v) {
  v += dpp_f<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);    // quad_perm [2,3,0,1]
  if constexpr (M >= 8) v += dpp_f<0x141>(v);   // row_half_mirror: sums of 8 lanes
  return v;
}

// 64-bit fixed point (2^40 per kcal/mol/A) of one pair-force component:
// integer part of v * 2^8 and a 32-bit fraction, so no FP64 is needed in the
// FP32 kernel (|v| < 2^23 kcal/mol/A).
__device__ inline long long force_fix_of(float v) {
  const float w = v * 256.0f;
  const float fl = __builtin_floorf(w);
  const int hi = (int)fl;
  const float fr = __builtin_fminf(w - fl, 0.99999994f);
  const unsigned int lo = (unsigned int)(fr * 4294967296.0f);
  return (long long)(((unsigned long long)(unsigned int)hi << 32) | lo);
}
__device__ inline long long force_fix_of(double v) { return __double2ll_rn(v * PM_FORCE_FIX); }

// Half-list pair kernel: one wave per i-cluster of M atoms, lane = jslot * M +
// ii, so a wave evaluates 64 / M list entries against the cluster per iteration.
// A pair is kept when the j slot follows the i slot (every unordered pair is
// listed once, by the cluster of its lower slot).  The i force accumulates
// per lane (FP32 partial sums of PM_IFLUSH trips added into FP64 totals in the
// mixed kernel, FP64 in the FP64 kernel) and is reduced over the j slots at
// the end; the j force of each entry is summed over its row of M lanes with
// DPP and added with one 64-bit integer atomic per component.  Forces
// accumulate per slot, so the atomics of one instruction fall on a few
// contiguous cache lines; integer addition is associative, so the result does
// not depend on their order.
//
// Arithmetic per pair (OneAtomPair.h:127-188, OneAtomPairTwo.h:131-199):
// LJ E = A r^-12 - B r^-6, F/r = (12 A r^-12 - 6 B r^-6) / r^2; Coulomb
// E = qq / r, F/r = E / r^2; Ewald real space E = qq poly5(a r) e^(-a^2 r^2) / r,
// F/r = (E + 2a/sqrt(pi) qq e^(-a^2 r^2)) / r^2; switched F = F S - E dS,
// E = E S.  Every term of a pass switches at the pass's cutoff (the host
// splits two-term forces whose switched terms end at different cutoffs), so
// no term is clipped inside the pair cutoff here.
#define PM_IFLUSH 4
// PA: per-atom LJ parameters (pm_set_lj_per_atom) instead of the type table,
// a kernel of its own so that its registers do not set the table kernel's
// occupancy.
// The block's copy of the LJ type table in LDS (every thread of the block
// must call this before any wave evaluates a cluster).
template <typename T, int LJSWF>
__device__ inline void pair_load_table(const PmNbLaunch& a) {
  extern __shared__ unsigned char s_lj_raw[];
  typedef typename std::conditional<sizeof(T) == 4, float4, double4>::type LJ4;
  LJ4* s_lj = reinterpret_cast<LJ4*>(s_lj_raw);
  const int nt1 = a.ntypes1;
  if constexpr (LJSWF != PM_SWF_NONE) {
    const LJ4* src;
    if constexpr (sizeof(T) == 4) src = a.lj_f; else src = a.lj_d;
    for (int t = threadIdx.x; t < nt1 * nt1; t += blockDim.x) s_lj[t] = src[t];
    __syncthreads();
  }
}

// The pairs of i-cluster cl, evaluated by one wave.
template <typename T, int LJSWF, int CK, int CSWF, bool WE, int M, bool PA>
__device__ inline void pair_cluster(const PmBox& box, const PmNbLaunch& a, const PmPairParams& pp, const int cl) {
  extern __shared__ unsigned char s_lj_raw[];
  typedef typename std::conditional<sizeof(T) == 4, float4, double4>::type LJ4;
  const LJ4* s_lj = reinterpret_cast<const LJ4*>(s_lj_raw);
  const int nt1 = a.ntypes1;
  const int lane = threadIdx.x & 63;
  const int ncl = a.counts[1];
  if (cl >= ncl) {
    if (WE && cl < a.ncl_max && lane < PM_NE) a.epart[(long long)cl * PM_NE + lane] = 0.0;
    return;
  }
  constexpr int JW = 64 / M;
  // The mixed-precision force kernel needs 72 VGPRs, which lets seven of
  // its waves share a SIMD; reserving registers up to v79 holds it to six,
  // so that the PME kernels running beside it on the second stream find two
  // free wave slots per SIMD instead of one (98k water: the step's PME chain
  // and this kernel then end together, 257.6 -> 246.0 us per MD step).
  if constexpr (sizeof(T) == 4 && !WE) asm volatile("" ::: "v79");
  const int ii = lane % M, js = lane / M;
  const int si = cl * M + ii;
  const int mi = a.slot_meta[si];
  const int ti = mi & 0xff, oi = (mi >> 8) & 0x7fffff;
  const bool ireal = oi != 0x7fffff;
  // positions: 32-bit periodic fixed point (mixed) or 64-bit (FP64); the
  // wrapped integer difference is the minimal-image vector
  typedef typename std::conditional<sizeof(T) == 4, int4, longlong4>::type P4;
  const P4* __restrict__ spos;
  if constexpr (sizeof(T) == 4) spos = a.slot_pos; else spos = a.slot_pos64;
  const P4 pi = spos[si];
  T qi;
  if constexpr (sizeof(T) == 4) qi = __int_as_float(pi.w); else qi = __longlong_as_double(pi.w);
  const T sx = sizeof(T) == 4 ? (T)box.fix_to_A[0] : (T)(box.fix_to_A[0] * 2.3283064365386963e-10);
  const T sy = sizeof(T) == 4 ? (T)box.fix_to_A[1] : (T)(box.fix_to_A[1] * 2.3283064365386963e-10);
  const T sz = sizeof(T) == 4 ? (T)box.fix_to_A[2] : (T)(box.fix_to_A[2] * 2.3283064365386963e-10);

  // parameters in working precision
  const T cut2 = (T)pp.cut2;
  const T lj_cut2 = (T)pp.lj_cut2, lj_on2 = (T)pp.lj_on2, lj_sw1 = (T)pp.lj_sw1, lj_sw2 = (T)pp.lj_sw2,
          lj_sw3 = (T)pp.lj_sw3, lj_c15_1 = (T)pp.lj_c15_1, lj_c05_3 = (T)pp.lj_c05_3,
          lj_c15_3 = (T)pp.lj_c15_3;
  const T c_cut2 = (T)pp.c_cut2, c_on2 = (T)pp.c_on2, c_sw1 = (T)pp.c_sw1, c_sw2 = (T)pp.c_sw2,
          c_sw3 = (T)pp.c_sw3, c_c15_1 = (T)pp.c_c15_1, c_c05_3 = (T)pp.c_c05_3, c_c15_3 = (T)pp.c_c15_3;
  const T alpha = (T)pp.alpha, tapi = (T)pp.two_alpha_pi, csf = (T)pp.csf;
  // e^(-a^2 r^2) = 2^(r^2 * ex_k), and p a for A&S 7.1.26's t = 1/(1 + p a r)
  const T ex_k = (T)(-(pp.alpha * pp.alpha) * 1.4426950408889634);
  const T p_alpha = (T)(0.3275911 * pp.alpha);
  const bool modify = pp.modify != 0;

  double fxa = 0, fya = 0, fza = 0;       // i force totals (FP64)
  T fxs = 0, fys = 0, fzs = 0;            // partial sums since the last flush
  T elj = 0, ec = 0;
  T vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0;
  // the inner (pruned) lists when kept (k_prune brought them up to date
  // before this kernel), else the full lists
  const long long lbase = (long long)cl * a.lcap;
  const int* __restrict__ lst = (a.inner ? a.ilist : a.jlist) + lbase;
  const unsigned int* __restrict__ msk = (a.inner ? a.imask : a.jmask) + lbase;
  const int cnt = a.inner ? a.icount[cl] : a.jcount[cl];
  const int last = cnt > 0 ? cnt - 1 : 0;
  const bool vir = WE && a.want_virial != 0;
  long long* __restrict__ ff = a.force_slot;
  const int comp = ii < 3 ? ii : ii - 3;   // force component a lane adds (lanes ii < 6)

  // Two iterations per trip with double-buffered prefetch: while row A
  // computes, the j atom of row B and the entry of the next trip's row A are
  // in flight, and the other way round, so no register copies carry values
  // across the loop; the trip's two rows share one atomic instruction.  Loads
  // past the list's end re-read its last entry (masked), so no load branches.
  int eA = cnt > 0 ? lst[min(js, last)] : a.dummy_slot;
  int eB = cnt > 0 ? lst[min(js + JW, last)] : a.dummy_slot;
  P4 pA = spos[eA & 0x7fffffff];
  int mA = a.slot_meta[eA & 0x7fffffff];
  P4 pB;
  int mB;

  // one iteration: the pairs of entry k against the cluster; returns this
  // lane's component of the entry's j force (row sum) and the j slot
  auto row = [&](const int k, const int e, const P4& pj, const int mj, int& sj_out) -> T {
    const bool valid = k < cnt;
    const int sj = e & 0x7fffffff;
    const bool flagged = e < 0;
    sj_out = sj;
    T dx, dy, dz;
    if constexpr (sizeof(T) == 4) {
      dx = (T)(int)(pj.x - pi.x) * sx;
      dy = (T)(int)(pj.y - pi.y) * sy;
      dz = (T)(int)(pj.z - pi.z) * sz;
    } else {
      dx = (T)(long long)((unsigned long long)pj.x - (unsigned long long)pi.x) * sx;
      dy = (T)(long long)((unsigned long long)pj.y - (unsigned long long)pi.y) * sy;
      dz = (T)(long long)((unsigned long long)pj.z - (unsigned long long)pi.z) * sz;
    }
    const T r2 = dx * dx + dy * dy + dz * dz;
    bool mask = valid && ireal && sj > si && (r2 <= cut2);
    // exclusions: only entries the list flagged hold an excluded pair
    bool modified = false, any_mod = false;
    if (__ballot(flagged) != 0ull) {
      if (flagged && mask) {
        const int xm = (int)msk[k];
        if ((xm >> ii) & 1) mask = false;
        else modified = ((xm >> (16 + ii)) & 1) != 0;
      }
      any_mod = __ballot(modified) != 0ull;
    }
    T qj;
    if constexpr (sizeof(T) == 4) qj = __int_as_float(pj.w); else qj = __longlong_as_double(pj.w);
    T rr2, rinv;
    if constexpr (sizeof(T) == 4) { rinv = Prec<T>::rsq(r2); rr2 = rinv * rinv; }
    else { rr2 = T(1) / r2; rinv = sqrt(rr2); }
    T F, E1 = 0, E2 = 0;
    if constexpr (LJSWF != PM_SWF_NONE) {
      T A, B;
      if constexpr (PA) {
        lj_ab<T>(a, s_lj, ti, mj, nt1, oi, modified ? PM_EXCL_MODIFIED : PM_EXCL_NONE, A, B);
      } else {
        const LJ4 ab = s_lj[ti * nt1 + (mj & 0xff)];
        A = (T)ab.x;
        B = (T)ab.y;
        if (any_mod && modified) { A = (T)ab.z; B = (T)ab.w; }   // A14, B14 (LennardJonesForce.h:55-60)
      }
      const T r6 = rr2 * rr2 * rr2;
      const T t12 = A * (r6 * r6);
      const T t6 = B * r6;
      T e1 = t12 - t6;
      T f1 = (T(12) * t12 - T(6) * t6) * rr2;
      if constexpr (LJSWF == PM_SWF_C2) {
        // C2SwitchingFunction.h:64-82: S = sw1 c2^2 (c^2 + 2r^2 - 3s^2) for s^2 <= r^2,
        // dS = sw3 (c2^2 - c2 (c^2 + 2r^2 - 3s^2)), with c2 = c^2 - r^2
        const T c2 = lj_cut2 - r2;
        const T c4 = c2 * (lj_sw2 + T(2) * r2);
        const bool in = r2 >= lj_on2;
        const T v = in ? lj_sw1 * (c2 * c4) : T(1);
        const T d = in ? lj_sw3 * (c2 * c2 - c4) : T(0);
        f1 = f1 * v - e1 * d;
        e1 = e1 * v;
      } else if constexpr (LJSWF == PM_SWF_C1) {
        if (modify)
          apply_switch<T, LJSWF>(r2, r2 * rinv, lj_cut2, lj_on2, lj_sw1, lj_sw2, lj_sw3, lj_c15_1, lj_c05_3,
                                 lj_c15_3, f1, e1);
      }
      F = f1;
      E1 = e1;
    }
    if constexpr (CK != PM_COUL_NONE) {
      T e2, f2;
      T qq = qi * qj;
      if (any_mod && modified) qq *= csf;   // coulombScalingFactor (CoulombForce.h:49-51)
      if constexpr (CK == PM_COUL_PLAIN) {
        e2 = qq * rinv;
        f2 = e2 * rr2;
      } else {
        // CoulombEwaldRealForce.h:45-73 with poly5 (A&S 7.1.26) for erfc
        T t;
        if constexpr (sizeof(T) == 4) t = Prec<T>::rcp(T(1) + p_alpha * (r2 * rinv));
        else t = Prec<T>::rcp(T(1) + T(0.3275911) * (alpha * (r2 * rinv)));
        const T poly = t * (T(0.254829592) + t * (T(-0.284496736) + t * (T(1.421413741) +
                                                 t * (T(-1.453152027) + t * T(1.061405429)))));
        T ex;
        if constexpr (sizeof(T) == 4) ex = qq * __builtin_amdgcn_exp2f(r2 * ex_k);
        else { const T ar = alpha * (r2 * rinv); ex = qq * Prec<T>::expneg(ar * ar); }
        e2 = poly * ex * rinv;
        f2 = (e2 + tapi * ex) * rr2;
      }
      if constexpr (CSWF == PM_SWF_C2) {
        const T c2 = c_cut2 - r2;
        const T c4 = c2 * (c_sw2 + T(2) * r2);
        const bool in = r2 >= c_on2;
        const T v = in ? c_sw1 * (c2 * c4) : T(1);
        const T d = in ? c_sw3 * (c2 * c2 - c4) : T(0);
        f2 = f2 * v - e2 * d;
        e2 = e2 * v;
      } else if constexpr (CSWF == PM_SWF_C1) {
        if (modify)
          apply_switch<T, CSWF>(r2, r2 * rinv, c_cut2, c_on2, c_sw1, c_sw2, c_sw3, c_c15_1, c_c05_3, c_c15_3, f2, e2);
      }
      if constexpr (LJSWF != PM_SWF_NONE) F += f2; else F = f2;
      E2 = e2;
    }
    F = mask ? F : T(0);
    // fij = F * (xj - xi): -fij on i, +fij on j (OneAtomPair.h:178-181)
    const T fx_ = F * dx, fy_ = F * dy, fz_ = F * dz;
    fxs -= fx_;
    fys -= fy_;
    fzs -= fz_;
    if constexpr (WE) {
      elj += mask ? E1 : T(0);
      ec += mask ? E2 : T(0);
      if (vir) {
        vxx += fx_ * dx; vxy += fx_ * dy; vxz += fx_ * dz;
        vyy += fy_ * dy; vyz += fy_ * dz; vzz += fz_ * dz;
      }
    }
    const T jx = row_sum<M>(fx_), jy = row_sum<M>(fy_), jz = row_sum<M>(fz_);
    return comp == 0 ? jx : (comp == 1 ? jy : jz);
  };

  bool range_bad = false;   // a converted force beyond the fixed-point range (or not finite)
  int trip = 0;
  for (int base = 0; base < cnt; base += 2 * JW) {
    const int kA = base + js, kB = kA + JW;
    // row A; prefetch the j atom of row B and the entry of the next row A
    const int eA_now = eA;
    pB = spos[eB & 0x7fffffff];
    mB = a.slot_meta[eB & 0x7fffffff];
    eA = lst[min(kA + 2 * JW, last)];
    int sA;
    const T vA = row(kA, eA_now, pA, mA, sA);
    // row B; prefetch the j atom of the next row A and the entry of the next row B
    const int eB_now = eB;
    pA = spos[eA & 0x7fffffff];
    mA = a.slot_meta[eA & 0x7fffffff];
    eB = lst[min(kB + 2 * JW, last)];
    int sB;
    const T vB = row(kB, eB_now, pB, mB, sB);
    if constexpr (M >= 8) {
      // lanes ii = 0, 1, 2 add the x, y, z component of row A, lanes 3, 4, 5 those of row B
      const T v = ii < 3 ? vA : vB;
      const int sl = ii < 3 ? sA : sB;
      if (ii < 6 && v != T(0)) {
        range_bad |= !(__builtin_fabs(v) < T(PM_FORCE_LIMIT));
        atomicAdd((unsigned long long*)(ff + 3LL * sl + comp), (unsigned long long)force_fix_of(v));
      }
    } else {
      // rows of four lanes: lanes ii = 0, 1, 2 add row A's components, then row B's
      if (ii < 3 && vA != T(0)) {
        range_bad |= !(__builtin_fabs(vA) < T(PM_FORCE_LIMIT));
        atomicAdd((unsigned long long*)(ff + 3LL * sA + comp), (unsigned long long)force_fix_of(vA));
      }
      if (ii < 3 && vB != T(0)) {
        range_bad |= !(__builtin_fabs(vB) < T(PM_FORCE_LIMIT));
        atomicAdd((unsigned long long*)(ff + 3LL * sB + comp), (unsigned long long)force_fix_of(vB));
      }
    }
    if (sizeof(T) == 4 && ++trip == PM_IFLUSH) {
      trip = 0;
      fxa += (double)fxs; fya += (double)fys; fza += (double)fzs;
      fxs = T(0); fys = T(0); fzs = T(0);
    }
  }
  fxa += (double)fxs; fya += (double)fys; fza += (double)fzs;
  // reduce the i force over the j slots (lanes ii, ii+8, ...); lanes
  // js = 0, 1, 2 then add its x, y, z component
#pragma unroll
  for (int off = M; off < 64; off <<= 1) {
    fxa += __shfl_xor(fxa, off, 64);
    fya += __shfl_xor(fya, off, 64);
    fza += __shfl_xor(fza, off, 64);
  }
  if (js < 3 && ireal) {
    const double iv = js == 0 ? fxa : (js == 1 ? fya : fza);
    range_bad |= !(__builtin_fabs(iv) < PM_FORCE_LIMIT);
    atomicAdd((unsigned long long*)(ff + 3LL * si + js), (unsigned long long)__double2ll_rn(iv * PM_FORCE_FIX));
  }
  if (__ballot(range_bad) != 0ull && lane == 0) pm_raise(box.err, PM_FLAG_FORCE_RANGE);
  if constexpr (WE) {
    // energies and virial of the cluster's pairs (each pair counted once)
    double r[8] = {ireal ? (double)ec : 0.0, ireal ? (double)elj : 0.0,
                   ireal ? (double)vxx : 0.0, ireal ? (double)vxy : 0.0, ireal ? (double)vxz : 0.0,
                   ireal ? (double)vyy : 0.0, ireal ? (double)vyz : 0.0, ireal ? (double)vzz : 0.0};
    const int nr = vir ? 8 : 2;
    for (int u = 0; u < nr; u++) {
      double v = r[u];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
      r[u] = v;
    }
    if (lane == 0) {
      double* ep = a.epart + (long long)cl * PM_NE;
      ep[PM_E_COUL] = r[0];
      ep[PM_E_LJ] = r[1];
      for (int u = 2; u < PM_E_VIR; u++) ep[u] = 0.0;
      if (vir) {
        // fij . diff with fij = F*d; the reference adds virial of +fij (ScalarStructure::addVirial)
        ep[PM_E_VIR + 0] = r[2]; ep[PM_E_VIR + 1] = r[3]; ep[PM_E_VIR + 2] = r[4];
        ep[PM_E_VIR + 3] = r[3]; ep[PM_E_VIR + 4] = r[5]; ep[PM_E_VIR + 5] = r[6];
        ep[PM_E_VIR + 6] = r[4]; ep[PM_E_VIR + 7] = r[6]; ep[PM_E_VIR + 8] = r[7];
      } else {
        for (int u = 0; u < 9; u++) ep[PM_E_VIR + u] = 0.0;
      }
    }
  }
}
// One wave per i-cluster, four clusters per block.
template <typename T, int LJSWF, int CK, int CSWF, bool WE, int M, bool PA>
__global__ __launch_bounds__(256) void k_pair(PmBox box, PmNbLaunch a, PmPairParams pp) {
  pair_load_table<T, LJSWF>(a);
  pair_cluster<T, LJSWF, CK, CSWF, WE, M, PA>(box, a, pp, blockIdx.x * 4 + (threadIdx.x >> 6));
}

// orders one wave's LDS writes before its lanes read each other's values
// (the wave's LDS operations complete in order; this keeps the compiler
// from moving the accesses across)
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave-uniform copies (scalar registers) of values every lane holds alike
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ inline long long uni(long long v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo);
}
__device__ inline double uni(double v) { return __longlong_as_double(uni(__double_as_longlong(v))); }

// PM_JL_LDS (pm_internal.h) bytes per wave: an iteration's j forces (3 x 64
// values and 64 slots), later the i-force reduction, PM_JL_NV(T) values of
// the 64 lanes per pass.
#define PM_JL_NV(T) (PM_JL_LDS / 64 / (int)sizeof(T))

// Half-list pair kernel with one lane per list entry: one wave per i-cluster
// of M atoms, whose positions, charges and types sit in scalar registers; a
// lane takes one entry (64 per iteration) and evaluates it against the M i
// atoms in turn.  The j force of an entry is then whole in its lane and is
// added with one 64-bit integer atomic per component, with no cross-lane
// sums; the i forces accumulate per lane in FP32 (one term per iteration and
// i atom; FP64 in the FP64 kernel) and are summed over the lanes in FP64 at
// the end, through LDS, in lane order.  Same pair arithmetic and the same
// half-list rule as pair_cluster; only the order in which FP32 partial sums
// are formed differs.
template <typename T, int LJSWF, int CK, int CSWF, bool WE, int M, bool PA>
__device__ inline void pair_cluster_jl(const PmBox& box, const PmNbLaunch& a, const PmPairParams& pp, const int cl) {
  extern __shared__ unsigned char s_lj_raw[];
  typedef typename std::conditional<sizeof(T) == 4, float4, double4>::type LJ4;
  const LJ4* s_lj = reinterpret_cast<const LJ4*>(s_lj_raw);
  const int nt1 = a.ntypes1;
  const int lane = threadIdx.x & 63;
  const int ncl = a.counts[1];
  if (cl >= ncl) {
    if (WE && cl < a.ncl_max && lane < PM_NE) a.epart[(long long)cl * PM_NE + lane] = 0.0;
    return;
  }
  typedef typename std::conditional<sizeof(T) == 4, int4, longlong4>::type P4;
  typedef typename std::conditional<sizeof(T) == 4, int, long long>::type PI;
  const P4* __restrict__ spos;
  if constexpr (sizeof(T) == 4) spos = a.slot_pos; else spos = a.slot_pos64;
  const int c0 = cl * M;
  // the cluster's atoms, wave-uniform
  PI ipx[M], ipy[M], ipz[M];
  T iq[M];
  int ity[M], ioi[M];
  unsigned int ireal = 0;
#pragma unroll
  for (int ii = 0; ii < M; ii++) {
    const P4 p = spos[c0 + ii];
    const int m = uni(a.slot_meta[c0 + ii]);
    ipx[ii] = uni(p.x);
    ipy[ii] = uni(p.y);
    ipz[ii] = uni(p.z);
    if constexpr (sizeof(T) == 4) iq[ii] = __int_as_float(uni(p.w)); else iq[ii] = __longlong_as_double(uni(p.w));
    ity[ii] = m & 0xff;
    ioi[ii] = (m >> 8) & 0x7fffff;
    if (ioi[ii] != 0x7fffff) ireal |= 1u << ii;
  }
  const T sx = sizeof(T) == 4 ? (T)box.fix_to_A[0] : (T)(box.fix_to_A[0] * 2.3283064365386963e-10);
  const T sy = sizeof(T) == 4 ? (T)box.fix_to_A[1] : (T)(box.fix_to_A[1] * 2.3283064365386963e-10);
  const T sz = sizeof(T) == 4 ? (T)box.fix_to_A[2] : (T)(box.fix_to_A[2] * 2.3283064365386963e-10);

  const T cut2 = (T)pp.cut2;
  const T lj_cut2 = (T)pp.lj_cut2, lj_on2 = (T)pp.lj_on2, lj_sw1 = (T)pp.lj_sw1, lj_sw2 = (T)pp.lj_sw2,
          lj_sw3 = (T)pp.lj_sw3, lj_c15_1 = (T)pp.lj_c15_1, lj_c05_3 = (T)pp.lj_c05_3,
          lj_c15_3 = (T)pp.lj_c15_3;
  const T c_cut2 = (T)pp.c_cut2, c_on2 = (T)pp.c_on2, c_sw1 = (T)pp.c_sw1, c_sw2 = (T)pp.c_sw2,
          c_sw3 = (T)pp.c_sw3, c_c15_1 = (T)pp.c_c15_1, c_c05_3 = (T)pp.c_c05_3, c_c15_3 = (T)pp.c_c15_3;
  const T alpha = (T)pp.alpha, tapi = (T)pp.two_alpha_pi, csf = (T)pp.csf;
  const T ex_k = (T)(-(pp.alpha * pp.alpha) * 1.4426950408889634);
  const T p_alpha = (T)(0.3275911 * pp.alpha);
  const bool modify = pp.modify != 0;

  T fix[M], fiy[M], fiz[M];   // i forces of this lane's entries
#pragma unroll
  for (int ii = 0; ii < M; ii++) { fix[ii] = T(0); fiy[ii] = T(0); fiz[ii] = T(0); }
  T elj = 0, ec = 0;
  T vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0;
  const long long lbase = (long long)cl * a.lcap;
  const int* __restrict__ lst = (a.inner ? a.ilist : a.jlist) + lbase;
  const unsigned int* __restrict__ msk = (a.inner ? a.imask : a.jmask) + lbase;
  const int cnt = a.inner ? a.icount[cl] : a.jcount[cl];
  const int last = cnt > 0 ? cnt - 1 : 0;
  const bool vir = WE && a.want_virial != 0;
  long long* __restrict__ ff = a.force_slot;
  bool range_bad = false;
  // this wave's LDS area (after the LJ table): the j forces of an iteration
  // (3 x 64 values and 64 slots), then the i-force reduction
  constexpr int NV = PM_JL_NV(T);
  const size_t tbytes = LJSWF != PM_SWF_NONE ? ((size_t)nt1 * nt1 * sizeof(LJ4) + 15) / 16 * 16 : 0;
  T* s_red = reinterpret_cast<T*>(s_lj_raw + tbytes) + (threadIdx.x >> 6) * 64 * NV;
  T* s_jf = s_red;
  int* s_js = reinterpret_cast<int*>(s_red + 3 * 64);

  // the entry and j atom of this iteration, and the entry of the next, in
  // flight while the previous iteration computes; loads past the list's end
  // re-read its last entry (masked)
  int e = cnt > 0 ? lst[min(lane, last)] : a.dummy_slot;
  P4 pj = spos[e & 0x7fffffff];
  int mj = a.slot_meta[e & 0x7fffffff];
  int en = cnt > 0 ? lst[min(lane + 64, last)] : a.dummy_slot;
  for (int base = 0; base < cnt; base += 64) {
    const P4 pn = spos[en & 0x7fffffff];
    const int mn = a.slot_meta[en & 0x7fffffff];
    const int enn = lst[min(base + 128 + lane, last)];
    const int k = base + lane;
    const bool valid = k < cnt;
    const int sj = e & 0x7fffffff;
    // exclusions: only entries the list flagged hold an excluded pair
    const bool flagged = valid && e < 0;
    unsigned int xm = 0;
    const bool anyflag = __ballot(flagged) != 0ull;
    if (anyflag && flagged) xm = msk[k];
    const bool any_mod = anyflag && __ballot((xm >> 16) != 0u) != 0ull;
    T qj;
    if constexpr (sizeof(T) == 4) qj = __int_as_float(pj.w); else qj = __longlong_as_double(pj.w);
    const int tj = mj & 0xff;
    T fjx = 0, fjy = 0, fjz = 0;
    // the pair of this lane's entry with cluster atom ii: MOD false for the
    // ordinary pairs (excluded and modified ones masked; xm is 0 for entries
    // the list did not flag), true for the modified (1-4) pairs, with the
    // 1-4 LJ parameters and the Coulomb scaling
    auto pair_ii = [&](const int ii, auto mod) {
      constexpr bool MOD = decltype(mod)::value;
      T dx, dy, dz;
      if constexpr (sizeof(T) == 4) {
        dx = (T)(int)(pj.x - ipx[ii]) * sx;
        dy = (T)(int)(pj.y - ipy[ii]) * sy;
        dz = (T)(int)(pj.z - ipz[ii]) * sz;
      } else {
        dx = (T)(long long)((unsigned long long)pj.x - (unsigned long long)ipx[ii]) * sx;
        dy = (T)(long long)((unsigned long long)pj.y - (unsigned long long)ipy[ii]) * sy;
        dz = (T)(long long)((unsigned long long)pj.z - (unsigned long long)ipz[ii]) * sz;
      }
      const T r2 = dx * dx + dy * dy + dz * dz;
      const bool mask = valid && sj > c0 + ii && (r2 <= cut2) &&
                        (MOD ? (xm & (65536u << ii)) != 0u : (xm & (65537u << ii)) == 0u);
      if (__ballot(mask) == 0ull) return;   // no lane's pair with atom ii counts
      T rr2, rinv;
      if constexpr (sizeof(T) == 4) { rinv = Prec<T>::rsq(r2); rr2 = rinv * rinv; }
      else { rr2 = T(1) / r2; rinv = sqrt(rr2); }
      T F, E1 = 0, E2 = 0;
      if constexpr (LJSWF != PM_SWF_NONE) {
        T A, B;
        if constexpr (PA) {
          lj_ab<T>(a, s_lj, ity[ii], mj, nt1, ioi[ii], MOD ? PM_EXCL_MODIFIED : PM_EXCL_NONE, A, B);
        } else {
          const LJ4 ab = s_lj[ity[ii] * nt1 + tj];
          // A14, B14 for modified pairs (LennardJonesForce.h:55-60)
          A = MOD ? (T)ab.z : (T)ab.x;
          B = MOD ? (T)ab.w : (T)ab.y;
        }
        const T r6 = rr2 * rr2 * rr2;
        const T t12 = A * (r6 * r6);
        const T t6 = B * r6;
        T e1 = t12 - t6;
        T f1 = (T(12) * t12 - T(6) * t6) * rr2;
        if constexpr (LJSWF == PM_SWF_C2) {
          const T c2 = lj_cut2 - r2;
          const T c4 = c2 * (lj_sw2 + T(2) * r2);
          const bool in = r2 >= lj_on2;
          const T v = in ? lj_sw1 * (c2 * c4) : T(1);
          const T d = in ? lj_sw3 * (c2 * c2 - c4) : T(0);
          f1 = f1 * v - e1 * d;
          e1 = e1 * v;
        } else if constexpr (LJSWF == PM_SWF_C1) {
          if (modify)
            apply_switch<T, LJSWF>(r2, r2 * rinv, lj_cut2, lj_on2, lj_sw1, lj_sw2, lj_sw3, lj_c15_1, lj_c05_3,
                                   lj_c15_3, f1, e1);
        }
        F = f1;
        E1 = e1;
      }
      if constexpr (CK != PM_COUL_NONE) {
        T e2, f2;
        T qq = iq[ii] * qj;
        if constexpr (MOD) qq *= csf;   // coulombScalingFactor (CoulombForce.h:49-51)
        if constexpr (CK == PM_COUL_PLAIN) {
          e2 = qq * rinv;
          f2 = e2 * rr2;
        } else {
          T t;
          if constexpr (sizeof(T) == 4) t = Prec<T>::rcp(T(1) + p_alpha * (r2 * rinv));
          else t = Prec<T>::rcp(T(1) + T(0.3275911) * (alpha * (r2 * rinv)));
          const T poly = t * (T(0.254829592) + t * (T(-0.284496736) + t * (T(1.421413741) +
                                                   t * (T(-1.453152027) + t * T(1.061405429)))));
          T ex;
          if constexpr (sizeof(T) == 4) ex = qq * __builtin_amdgcn_exp2f(r2 * ex_k);
          else { const T ar = alpha * (r2 * rinv); ex = qq * Prec<T>::expneg(ar * ar); }
          e2 = poly * ex * rinv;
          f2 = (e2 + tapi * ex) * rr2;
        }
        if constexpr (CSWF == PM_SWF_C2) {
          const T c2 = c_cut2 - r2;
          const T c4 = c2 * (c_sw2 + T(2) * r2);
          const bool in = r2 >= c_on2;
          const T v = in ? c_sw1 * (c2 * c4) : T(1);
          const T d = in ? c_sw3 * (c2 * c2 - c4) : T(0);
          f2 = f2 * v - e2 * d;
          e2 = e2 * v;
        } else if constexpr (CSWF == PM_SWF_C1) {
          if (modify)
            apply_switch<T, CSWF>(r2, r2 * rinv, c_cut2, c_on2, c_sw1, c_sw2, c_sw3, c_c15_1, c_c05_3, c_c15_3, f2, e2);
        }
        if constexpr (LJSWF != PM_SWF_NONE) F += f2; else F = f2;
        E2 = e2;
      }
      F = mask ? F : T(0);
      // fij = F * (xj - xi): -fij on i, +fij on j (OneAtomPair.h:178-181)
      const T fx_ = F * dx, fy_ = F * dy, fz_ = F * dz;
      fix[ii] -= fx_;
      fiy[ii] -= fy_;
      fiz[ii] -= fz_;
      fjx += fx_;
      fjy += fy_;
      fjz += fz_;
      if constexpr (WE) {
        elj += mask ? E1 : T(0);
        ec += mask ? E2 : T(0);
        if (vir) {
          vxx += fx_ * dx; vxy += fx_ * dy; vxz += fx_ * dz;
          vyy += fy_ * dy; vyz += fy_ * dz; vzz += fz_ * dz;
        }
      }
    };
#pragma unroll
    for (int ii = 0; ii < M; ii++)
      if ((ireal >> ii) & 1u) pair_ii(ii, std::false_type{});   // padding slots skipped (uniform)
    if (any_mod) {
#pragma unroll
      for (int ii = 0; ii < M; ii++)
        if ((ireal >> ii) & 1u) pair_ii(ii, std::true_type{});
    }
    // j forces: the 64 entries' x, y, z regrouped through LDS so that lane
    // v of round r adds component v % 3 of entry v / 3 (v = 64 r + lane):
    // each atomic instruction covers about 21 entries' components, which lie
    // in a few cache lines when the entries' slots are consecutive (as along
    // a column range), instead of 64 entries' single components
    s_jf[3 * lane + 0] = fjx;
    s_jf[3 * lane + 1] = fjy;
    s_jf[3 * lane + 2] = fjz;
    s_js[lane] = sj;
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const int v = 64 * r + lane, ent = v / 3, comp = v - 3 * ent;
      const T val = s_jf[v];
      range_bad |= !(__builtin_fabs(val) < T(PM_FORCE_LIMIT));
      if (val != T(0))
        atomicAdd((unsigned long long*)(ff + 3LL * s_js[ent] + comp), (unsigned long long)force_fix_of(val));
    }
    wave_lds_sync();
    e = en;
    pj = pn;
    mj = mn;
    en = enn;
  }
  // i forces: per pass, every lane's partial sums of NV values into LDS;
  // lane L adds value L % NV over the G = 64 / NV lane segments' share
  // L / NV in FP64, then the G segment sums of each value are added over the
  // lanes (the same order on every run).  The cluster's slots are
  // consecutive, so value c is component c % 3 of slot c0 + c / 3.
  {
    constexpr int G = 64 / NV, SEG = 64 / G;
    static_assert(NV * sizeof(T) * 64 == PM_JL_LDS && 3 * 64 * sizeof(T) + 64 * sizeof(int) <= PM_JL_LDS, "LDS area");
#pragma unroll
    for (int pass = 0; pass < (3 * M + NV - 1) / NV; pass++) {
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const int c = pass * NV + v, ii = c / 3, comp = c % 3;
        if (c < 3 * M) s_red[v * 64 + lane] = comp == 0 ? fix[ii] : (comp == 1 ? fiy[ii] : fiz[ii]);
      }
      wave_lds_sync();
      const int v = lane % NV, g = lane / NV;
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < SEG; t++) acc += (double)s_red[v * 64 + g * SEG + t];
#pragma unroll
      for (int off = NV; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
      const int c = pass * NV + v;
      if (lane < NV && c < 3 * M && ((ireal >> (c / 3)) & 1u)) {
        range_bad |= !(__builtin_fabs(acc) < PM_FORCE_LIMIT);
        atomicAdd((unsigned long long*)(ff + 3LL * c0 + c), (unsigned long long)__double2ll_rn(acc * PM_FORCE_FIX));
      }
      wave_lds_sync();
    }
  }
  if (__ballot(range_bad) != 0ull && lane == 0) pm_raise(box.err, PM_FLAG_FORCE_RANGE);
  if constexpr (WE) {
    // energies and virial of the cluster's pairs (each pair counted once)
    double r[8] = {(double)ec, (double)elj, (double)vxx, (double)vxy, (double)vxz,
                   (double)vyy, (double)vyz, (double)vzz};
    const int nr = vir ? 8 : 2;
    for (int u = 0; u < nr; u++) {
      double v = r[u];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
      r[u] = v;
    }
    if (lane == 0) {
      double* ep = a.epart + (long long)cl * PM_NE;
      ep[PM_E_COUL] = r[0];
      ep[PM_E_LJ] = r[1];
      for (int u = 2; u < PM_E_VIR; u++) ep[u] = 0.0;
      if (vir) {
        ep[PM_E_VIR + 0] = r[2]; ep[PM_E_VIR + 1] = r[3]; ep[PM_E_VIR + 2] = r[4];
        ep[PM_E_VIR + 3] = r[3]; ep[PM_E_VIR + 4] = r[5]; ep[PM_E_VIR + 5] = r[6];
        ep[PM_E_VIR + 6] = r[4]; ep[PM_E_VIR + 7] = r[6]; ep[PM_E_VIR + 8] = r[7];
      } else {
        for (int u = 0; u < 9; u++) ep[PM_E_VIR + u] = 0.0;
      }
    }
  }
}
template <typename T, int LJSWF, int CK, int CSWF, bool WE, int M, bool PA>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 6 : 1) void k_pair_jl(PmBox box, PmNbLaunch a, PmPairParams pp) {
  pair_load_table<T, LJSWF>(a);
  pair_cluster_jl<T, LJSWF, CK, CSWF, WE, M, PA>(box, a, pp, uni((int)(blockIdx.x * 4 + (threadIdx.x >> 6))));
}
struct PairArgs { PmBox box; PmNbLaunch a; PmPairParams pp; };
// batched launches: JL selects the kernel layout (one lane per entry or a row per entry)
template <typename T, int LJSWF, int CK, int CSWF, bool WE, int M, bool JL>
__global__ __launch_bounds__(256) void kb_pair(const PairArgs* __restrict__ t, const int* __restrict__ pre, int nrep) {
  const int r = pm_batch_replica(pre, nrep, blockIdx.x);
  const PairArgs& q = t[r];
  pair_load_table<T, LJSWF>(q.a);
  const int cl = (blockIdx.x - pre[r]) * 4 + (threadIdx.x >> 6);
  if constexpr (JL) pair_cluster_jl<T, LJSWF, CK, CSWF, WE, M, false>(q.box, q.a, q.pp, uni(cl));
  else pair_cluster<T, LJSWF, CK, CSWF, WE, M, false>(q.box, q.a, q.pp, cl);
}


// ------------------------------------------------------- adjoint pair kernel
// Force/r F(u) of one switched term and its derivative F_u = dF/du, from the
// term's energy E, force/r F0 and F0_u; S, D = 2 dS/du and S_uu from the
// switching function (OneAtomPair.h:170-175: force = force * S - energy * D).
template <typename T>
__device__ inline void switched_fu(T E, T F0, T F0u, T S, T D, T Suu, T& F, T& Fu) {
  F = F0 * S - E * D;
  // E_u = -F0/2 for the exact-derivative potentials (LJ, Coulomb)
  Fu = D * F0 + S * F0u - T(2) * E * Suu;
}

// MODE 1: Hessian-vector product of the pair energy, H v, with
//   (H v)_i += F dv + 2 F_u (d . dv) d,  (H v)_j -= the same,
//   d = x_j - x_i, dv = v_j - v_i, u = |d|^2, F = force/r of the pair as the
//   force kernel computes it (so that H is minus the derivative of exactly
//   those forces), F_u = dF/du.
// MODE 2: parameter gradients per atom: (dE/dq, dE/dR, dE/de) with R the
//   LJ radius and e = sqrt(eps) of per-atom parameters (dR, de zero with a
//   type table); MODIFIED pairs contribute to dE/dq only.
// The pairs come from the same cluster lists as the forces (each listed once);
// results accumulate per slot as 64-bit fixed point.
template <typename T, int LJSWF, int CK, int CSWF, int MODE>
__global__ __launch_bounds__(256) void k_adjoint(PmBox box, PmNbLaunch a, PmPairParams pp,
                                                 const double4* __restrict__ slot_v, long long* __restrict__ out) {
  constexpr int M = 8, JW = 8;
  extern __shared__ unsigned char s_lj_raw[];
  typedef typename std::conditional<sizeof(T) == 4, float4, double4>::type LJ4;
  LJ4* s_lj = reinterpret_cast<LJ4*>(s_lj_raw);
  const int nt1 = a.ntypes1;
  if constexpr (LJSWF != PM_SWF_NONE) {
    const LJ4* src;
    if constexpr (sizeof(T) == 4) src = a.lj_f; else src = a.lj_d;
    for (int t = threadIdx.x; t < nt1 * nt1; t += blockDim.x) s_lj[t] = src[t];
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cl = blockIdx.x * 4 + wave;
  if (cl >= a.counts[1]) return;
  const int ii = lane % M, js = lane / M;
  const int si = cl * M + ii;
  const int mi = a.slot_meta[si];
  const int ti = mi & 0xff, oi = (mi >> 8) & 0x7fffff;
  const bool ireal = oi != 0x7fffff;
  typedef typename std::conditional<sizeof(T) == 4, int4, longlong4>::type P4;
  const P4* __restrict__ spos;
  if constexpr (sizeof(T) == 4) spos = a.slot_pos; else spos = a.slot_pos64;
  const P4 pi = spos[si];
  T qi;
  if constexpr (sizeof(T) == 4) qi = __int_as_float(pi.w); else qi = __longlong_as_double(pi.w);
  const T sx = sizeof(T) == 4 ? (T)box.fix_to_A[0] : (T)(box.fix_to_A[0] * 2.3283064365386963e-10);
  const T sy = sizeof(T) == 4 ? (T)box.fix_to_A[1] : (T)(box.fix_to_A[1] * 2.3283064365386963e-10);
  const T sz = sizeof(T) == 4 ? (T)box.fix_to_A[2] : (T)(box.fix_to_A[2] * 2.3283064365386963e-10);
  const T cut2 = (T)pp.cut2;
  const T lj_cut2 = (T)pp.lj_cut2, lj_on2 = (T)pp.lj_on2, lj_sw1 = (T)pp.lj_sw1, lj_sw2 = (T)pp.lj_sw2,
          lj_sw3 = (T)pp.lj_sw3, lj_c15_1 = (T)pp.lj_c15_1, lj_c05_3 = (T)pp.lj_c05_3,
          lj_c15_3 = (T)pp.lj_c15_3;
  const T c_cut2 = (T)pp.c_cut2, c_on2 = (T)pp.c_on2, c_sw1 = (T)pp.c_sw1, c_sw2 = (T)pp.c_sw2,
          c_sw3 = (T)pp.c_sw3, c_c15_1 = (T)pp.c_c15_1, c_c05_3 = (T)pp.c_c05_3, c_c15_3 = (T)pp.c_c15_3;
  const T alpha = (T)pp.alpha, tapi = (T)pp.two_alpha_pi, csf = (T)pp.csf;
  const bool modify = pp.modify != 0;
  const bool per_atom = sizeof(T) == 4 ? a.atom_lj_f != nullptr : a.atom_lj_d != nullptr;
  const double4 vid = slot_v[si];
  const T vix = (T)vid.x, viy = (T)vid.y, viz = (T)vid.z;
  T Ri = 0, ei = 0;
  if (MODE == 2 && per_atom && ireal) {
    if constexpr (sizeof(T) == 4) { const float4 w = a.atom_lj_f[oi]; Ri = w.x; ei = w.y; }
    else { const double4 w = a.atom_lj_d[oi]; Ri = w.x; ei = w.y; }
  }
  double ax = 0, ay = 0, az = 0;
  const int* __restrict__ lst = a.jlist + (long long)cl * a.lcap;
  const int cnt = a.jcount[cl];
  const int comp = ii;   // lanes ii = 0, 1, 2 add the x, y, z component of a row

  for (int base = 0; base < cnt; base += JW) {
    const int k = base + js;
    const bool valid = k < cnt;
    const int e = valid ? lst[k] : a.dummy_slot;
    const bool flagged = e < 0;
    const int sj = e & 0x7fffffff;
    const P4 pj = spos[sj];
    const int mj = a.slot_meta[sj];
    const int oj = (mj >> 8) & 0x7fffff;
    T dx, dy, dz;
    if constexpr (sizeof(T) == 4) {
      dx = (T)(int)(pj.x - pi.x) * sx;
      dy = (T)(int)(pj.y - pi.y) * sy;
      dz = (T)(int)(pj.z - pi.z) * sz;
    } else {
      dx = (T)(long long)((unsigned long long)pj.x - (unsigned long long)pi.x) * sx;
      dy = (T)(long long)((unsigned long long)pj.y - (unsigned long long)pi.y) * sy;
      dz = (T)(long long)((unsigned long long)pj.z - (unsigned long long)pi.z) * sz;
    }
    const T r2 = dx * dx + dy * dy + dz * dz;
    bool mask = valid && ireal && sj > si && (r2 <= cut2);
    int excl = PM_EXCL_NONE;
    if (__ballot(flagged && mask) != 0ull) {
      if (flagged && mask) {
        const int xm = a.jmask[(long long)cl * a.lcap + k];
        if ((xm >> ii) & 1) { excl = PM_EXCL_FULL; mask = false; }
        else if ((xm >> (16 + ii)) & 1) excl = PM_EXCL_MODIFIED;
      }
    }
    T qj;
    if constexpr (sizeof(T) == 4) qj = __int_as_float(pj.w); else qj = __longlong_as_double(pj.w);
    const T r2s = mask ? r2 : T(1);   // keeps masked lanes finite
    T rinv, rr2;
    if constexpr (sizeof(T) == 4) { rinv = Prec<T>::rsq(r2s); rr2 = rinv * rinv; }
    else { rr2 = T(1) / r2s; rinv = sqrt(rr2); }
    const T dist = r2s * rinv;
    T c0 = 0, c1 = 0, c2 = 0;   // this lane's contribution to the i atom (j gets its mirror)
    if constexpr (MODE == 1) {
      T F = 0, Fu = 0;
      if constexpr (LJSWF != PM_SWF_NONE) {
        T A, B;
        lj_ab<T>(a, s_lj, ti, mj, nt1, oi, excl, A, B);
        const T r6 = rr2 * rr2 * rr2;
        const T E1 = A * r6 * r6 - B * r6;
        const T F1 = (T(12) * A * r6 * r6 - T(6) * B * r6) * rr2;
        const T F1u = (T(-84) * A * r6 * r6 + T(24) * B * r6) * rr2 * rr2;
        if (modify && LJSWF != PM_SWF_CUTOFF) {
          T S, D;
          switching<T, LJSWF == PM_SWF_CUTOFF ? PM_SWF_CUTOFF : LJSWF>(
              r2s, dist, lj_cut2, lj_on2, lj_sw1, lj_sw2, lj_sw3, lj_c15_1, lj_c05_3, lj_c15_3, S, D);
          T Suu = T(0);
          if constexpr (LJSWF == PM_SWF_C2) Suu = r2s >= lj_on2 ? T(2) * lj_sw1 * (lj_sw2 + T(2) * r2s - T(4) * (lj_cut2 - r2s)) : T(0);
          else if constexpr (LJSWF == PM_SWF_C1) Suu = (lj_c15_1 * rr2 + lj_c15_3) / (T(4) * dist);
          T Fs, Fsu;
          switched_fu<T>(E1, F1, F1u, S, D, Suu, Fs, Fsu);
          F += Fs; Fu += Fsu;
        } else if (modify && r2s > lj_cut2) {
        } else {
          F += F1; Fu += F1u;
        }
      }
      if constexpr (CK != PM_COUL_NONE) {
        T qq = qi * qj;
        if (excl == PM_EXCL_MODIFIED) qq *= csf;
        T E2, F2, F2u;
        if constexpr (CK == PM_COUL_PLAIN) {
          E2 = qq * rinv;
          F2 = E2 * rr2;
          F2u = T(-1.5) * F2 * rr2;
        } else {
          const T ar = alpha * dist;
          const T t = T(1) / (T(1) + T(0.3275911) * ar);
          const T P = t * (T(0.254829592) + t * (T(-0.284496736) + t * (T(1.421413741) +
                                              t * (T(-1.453152027) + t * T(1.061405429)))));
          const T Pp = T(0.254829592) + t * (T(2) * T(-0.284496736) + t * (T(3) * T(1.421413741) +
                                              t * (T(4) * T(-1.453152027) + t * T(5) * T(1.061405429))));
          const T X = exp(-ar * ar);
          E2 = qq * P * X * rinv;
          const T h = -T(0.3275911) * alpha * t * t * Pp * rinv - T(2) * alpha * alpha * P - P * rr2;
          const T E2u = T(0.5) * qq * X * h * rinv;
          F2 = (E2 + tapi * qq * X) * rr2;
          F2u = (E2u - alpha * alpha * tapi * qq * X) * rr2 - F2 * rr2;
        }
        if (modify && CSWF != PM_SWF_CUTOFF) {
          T S, D;
          switching<T, CSWF>(r2s, dist, c_cut2, c_on2, c_sw1, c_sw2, c_sw3, c_c15_1, c_c05_3, c_c15_3, S, D);
          T Suu = T(0);
          if constexpr (CSWF == PM_SWF_C2) Suu = r2s >= c_on2 ? T(2) * c_sw1 * (c_sw2 + T(2) * r2s - T(4) * (c_cut2 - r2s)) : T(0);
          else if constexpr (CSWF == PM_SWF_C1) Suu = (c_c15_1 * rr2 + c_c15_3) / (T(4) * dist);
          T Fs, Fsu;
          switched_fu<T>(E2, F2, F2u, S, D, Suu, Fs, Fsu);
          F += Fs; Fu += Fsu;
        } else if (modify && r2s > c_cut2) {
        } else {
          F += F2; Fu += F2u;
        }
      }
      F = mask ? F : T(0);
      Fu = mask ? Fu : T(0);
      const double4 vjd = slot_v[sj];
      const T dvx = (T)vjd.x - vix, dvy = (T)vjd.y - viy, dvz = (T)vjd.z - viz;
      const T g = T(2) * Fu * (dx * dvx + dy * dvy + dz * dvz);
      c0 = F * dvx + g * dx;
      c1 = F * dvy + g * dy;
      c2 = F * dvz + g * dz;
      ax += (double)c0; ay += (double)c1; az += (double)c2;
      // j receives -t
      c0 = -c0; c1 = -c1; c2 = -c2;
    } else {
      // dE/dq: the Coulomb energy per unit charge product
      T eu_i = 0, eu_j = 0, dRs = 0, dee = 0;
      if constexpr (CK != PM_COUL_NONE) {
        T eu;
        if constexpr (CK == PM_COUL_PLAIN) eu = rinv;
        else {
          const T ar = alpha * dist;
          const T t = T(1) / (T(1) + T(0.3275911) * ar);
          const T P = t * (T(0.254829592) + t * (T(-0.284496736) + t * (T(1.421413741) +
                                              t * (T(-1.453152027) + t * T(1.061405429)))));
          eu = P * exp(-ar * ar) * rinv;
        }
        if (excl == PM_EXCL_MODIFIED) eu *= csf;
        if (modify) {
          if (CSWF == PM_SWF_CUTOFF) { if (r2s > c_cut2) eu = T(0); }
          else {
            T S, D;
            switching<T, CSWF>(r2s, dist, c_cut2, c_on2, c_sw1, c_sw2, c_sw3, c_c15_1, c_c05_3, c_c15_3, S, D);
            eu *= S;
          }
        }
        eu_i = qj * eu;
        eu_j = qi * eu;
      }
      if constexpr (LJSWF != PM_SWF_NONE) {
        if (per_atom && excl != PM_EXCL_MODIFIED) {
          T Rj, ej;
          const int ja = oj == 0x7fffff ? 0 : oj;   // padding / dummy slot: masked lane
          if constexpr (sizeof(T) == 4) { const float4 w = a.atom_lj_f[ja]; Rj = w.x; ej = w.y; }
          else { const double4 w = a.atom_lj_d[ja]; Rj = w.x; ej = w.y; }
          const T rs = Ri + Rj, ee = ei * ej;
          const T rs2 = rs * rs, rs5 = rs2 * rs2 * rs, rs6 = rs5 * rs;
          const T r6 = rr2 * rr2 * rr2;
          T S = T(1);
          if (modify) {
            if (LJSWF == PM_SWF_CUTOFF) { if (r2s > lj_cut2) S = T(0); }
            else {
              T D;
              switching<T, LJSWF == PM_SWF_CUTOFF ? PM_SWF_CUTOFF : LJSWF>(
                  r2s, dist, lj_cut2, lj_on2, lj_sw1, lj_sw2, lj_sw3, lj_c15_1, lj_c05_3, lj_c15_3, S, D);
            }
          }
          dRs = S * T(12) * ee * (rs5 * rs6 * r6 * r6 - rs5 * r6);
          dee = S * (rs6 * rs6 * r6 * r6 - T(2) * rs6 * r6);
          c1 = dRs;
          c2 = dee * ej;
        }
      }
      if (!mask) { eu_i = eu_j = T(0); c1 = c2 = T(0); dee = T(0); dRs = T(0); }
      ax += (double)eu_i; ay += (double)c1; az += (double)c2;
      c0 = eu_j;
      c1 = dRs;
      c2 = dee * ei;
    }
    // j-side: sum over the row, lanes ii = 0, 1, 2 add components 0, 1, 2
    const T j0 = row_sum<M>(c0), j1 = row_sum<M>(c1), j2 = row_sum<M>(c2);
    const T jc = comp == 0 ? j0 : (comp == 1 ? j1 : j2);
    if (ii < 3 && jc != T(0))
      atomicAdd((unsigned long long*)(out + 3LL * sj + comp), (unsigned long long)force_fix_of(jc));
  }
#pragma unroll
  for (int off = M; off < 64; off <<= 1) {
    ax += __shfl_xor(ax, off, 64);
    ay += __shfl_xor(ay, off, 64);
    az += __shfl_xor(az, off, 64);
  }
  if (js < 3 && ireal) {
    const double iv = js == 0 ? ax : (js == 1 ? ay : az);
    atomicAdd((unsigned long long*)(out + 3LL * si + js), (unsigned long long)__double2ll_rn(iv * PM_FORCE_FIX));
  }
}

// per-slot copy of a per-atom vector (zero for padding slots)
__global__ void k_slot_vec(const int* __restrict__ counts, const int* __restrict__ atom_of_slot,
                           const double* __restrict__ v, double4* __restrict__ sv, int nslot_max) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nslot_max) return;   // the padding slot nslot_max included
  const int at = s < counts[0] ? atom_of_slot[s] : -1;
  sv[s] = at >= 0 ? make_double4(v[3 * at], v[3 * at + 1], v[3 * at + 2], 0.0) : make_double4(0, 0, 0, 0);
}

// ---------------------------------------------------- binning diagnostics
// PeriodicBoundaryConditions::minimalDifference (.h:76-117), orthogonal branch.
__device__ inline double min_diff_r2(const PmBox& b, const double* c1, const double* c2) {
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
      s = x * x + y * y + z * z;
    }
  }
  return s;
}

__device__ inline bool diag_pair_in(const PmBox& b, const double* pos, int i, int j, double cut2,
                                    const int* molecule, const int* ex_row, const int* ex_partner,
                                    const int* ex_cls) {
  double r2 = min_diff_r2(b, pos + 3 * i, pos + 3 * j);
  if (r2 > cut2) return false;
  if (molecule[i] == molecule[j]) {
    int lo = min(i, j), hi = max(i, j), r = PM_EXCL_NONE;
    for (int u = ex_row[lo]; u < ex_row[lo + 1]; u++)
      if (ex_partner[u] == hi) r = ex_cls[u];
    if (r == PM_EXCL_FULL) return false;
  }
  return true;
}

// One thread per cell: candidate and in-cutoff pairs of the cell pairs whose
// first cell it is, in the enumeration of
// topology/CellListEnumerator_periodicBoundaries.h:143-179.
__global__ void k_diag(PmBox box, const double* __restrict__ pos, const int* __restrict__ cell_start,
                       const int* __restrict__ cell_count, const int* __restrict__ cell_atoms,
                       const int4* __restrict__ half, int nhalf, const int* __restrict__ molecule,
                       const int* __restrict__ ex_row, const int* __restrict__ ex_partner,
                       const int* __restrict__ ex_cls, double cut2, int ncells,
                       long long* __restrict__ cand, long long* __restrict__ inrc) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  long long nc = 0, ni = 0;
  int k = cell_count[c];
  if (k > PM_DIAG_MAX) {
    // collapsed or non-finite coordinates: report instead of an unbounded loop
    cand[c] = -1;
    inrc[c] = -1;
    return;
  }
  if (k > 0) {
    const int* A = cell_atoms + cell_start[c];
    for (int u = 0; u < k; u++)
      for (int w = u + 1; w < k; w++) {
        nc++;
        ni += diag_pair_in(box, pos, A[u], A[w], cut2, molecule, ex_row, ex_partner, ex_cls);
      }
    int cz = c % box.n[2], cy = (c / box.n[2]) % box.n[1], cx = c / (box.n[1] * box.n[2]);
    for (int t = 0; t < nhalf; t++) {
      int4 d = half[t];
      int nb = (wrap_cell(cx, d.x, box.n[0]) * box.n[1] + wrap_cell(cy, d.y, box.n[1])) * box.n[2] +
               wrap_cell(cz, d.z, box.n[2]);
      if (nb < c) continue;
      int kb = cell_count[nb];
      if (kb == 0 || kb > PM_DIAG_MAX) continue;
      const int* B = cell_atoms + cell_start[nb];
      for (int u = 0; u < k; u++)
        for (int w = 0; w < kb; w++) {
          nc++;
          ni += diag_pair_in(box, pos, A[u], B[w], cut2, molecule, ex_row, ex_partner, ex_cls);
        }
    }
  }
  cand[c] = nc;
  inrc[c] = ni;
}

}  // namespace

// ----------------------------------------------------------------- launchers
// pair-kernel variant of a record: precision, potentials and switching,
// energy flag and cluster size
static int pair_variant(bool dbl, int lj, int ck, int cs, bool we, int m, bool jlanes) {
  return (dbl ? 1 : 0) | (lj << 1) | (ck << 3) | (cs << 5) | ((we ? 1 : 0) << 7) | ((m == 4 ? 1 : 0) << 8) |
         ((jlanes ? 1 : 0) << 9);
}
void pm_launch_bin(hipStream_t s, const PmBox& box, int n, const double* pos, int* cell_of,
                   int* cell_count, int* rank, float* key, const int* rebuild) {
  if (pm_recording()) {
    const BinArgs a{box, n, pos, cell_of, cell_count, rank, key, rebuild};
    pm_record(PM_REC_BIN, 0, n > 0 ? (n + 255) / 256 : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (n > 0)
    hipLaunchKernelGGL(k_bin, dim3((n + 255) / 256), dim3(256), 0, s, box, n, pos, cell_of, cell_count, rank, key,
                       rebuild);
}

void pm_launch_scan_cells(hipStream_t s, int ncells, int nz, int* bin_count, int* cell_count, int* cell_start,
                          int* cell_slot, int* counts, int* cluster_cell, int ncl_max, int m, const int* rebuild,
                          int* aux) {
  const int ch = scan_chunk_cells(nz);
  const int nchunks = (ncells + ch - 1) / ch;
  if (!pm_recording() && aux && nchunks > 2) {
    hipLaunchKernelGGL(k_scan_chunks, dim3(nchunks), dim3(1024), 0, s, ncells, nz, bin_count, cell_count, m, aux,
                       rebuild);
    hipLaunchKernelGGL(k_scan_finish, dim3(nchunks), dim3(1024), 0, s, ncells, nz, cell_count, cell_start,
                       cell_slot, counts, cluster_cell, ncl_max, m, aux, rebuild);
    return;
  }
  if (pm_recording()) {
    const ScanCellsArgs a{ncells, nz, bin_count, cell_count, cell_start, cell_slot, counts, cluster_cell, ncl_max, m,
                          rebuild};
    pm_record(PM_REC_SCAN_CELLS, 0, 1, 1024, 0, &a, sizeof(a));
    return;
  }
  hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, ncells, nz, bin_count, cell_count, cell_start,
                     cell_slot, counts, cluster_cell, ncl_max, m, rebuild);
}

void pm_launch_scatter(hipStream_t s, int n, const int* cell_of, const int* rank, const int* cell_start,
                       int* cell_atoms, const int* rebuild) {
  if (pm_recording()) {
    const ScatterArgs a{n, cell_of, rank, cell_start, cell_atoms, rebuild};
    pm_record(PM_REC_SCATTER, 0, n > 0 ? (n + 255) / 256 : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (n > 0)
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, s, n, cell_of, rank, cell_start,
                       cell_atoms, rebuild);
}

void pm_launch_fill_slots(hipStream_t s, const PmBox& box, int ncells, const double* pos,
                          const double* q, const int* type, const int* cell_count,
                          const int* cell_start, const int* cell_slot, int* cell_atoms,
                          const float* key, int4* slot_pos, longlong4* slot_pos64, int* slot_meta,
                          int* atom_of_slot, int dummy_type, int dummy_slot, int m, const int* rebuild,
                          int4* x_build, int4* x_prune) {
  int nb = (ncells + 3) / 4;
  if (pm_recording()) {
    const FillSlotsArgs a{box, ncells, pos, q, type, cell_count, cell_start, cell_slot, cell_atoms, key, slot_pos,
                          slot_pos64, slot_meta, atom_of_slot, dummy_type, dummy_slot, m, rebuild, x_build, x_prune};
    pm_record(PM_REC_FILL_SLOTS, 0, nb > 0 ? nb : 1, 256, 0, &a, sizeof(a));
    return;
  }
  hipLaunchKernelGGL(k_fill_slots, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, box, ncells, pos, q, type,
                     cell_count, cell_start, cell_slot, cell_atoms, key, slot_pos, slot_pos64, slot_meta,
                     atom_of_slot, dummy_type, dummy_slot, m, rebuild, x_build, x_prune);
}

void pm_launch_check_disp(hipStream_t s, const PmBox& box, int n, const double* pos, const int4* x_build,
                          float lim2, int* flag, const int4* x_prune, float lim2_in) {
  if (pm_recording()) {
    const CheckDispArgs a{box, n, pos, x_build, lim2, flag, x_prune, lim2_in};
    pm_record(PM_REC_CHECK_DISP, 0, n > 0 ? (n + 255) / 256 : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (n > 0)
    hipLaunchKernelGGL(k_check_disp, dim3((n + 255) / 256), dim3(256), 0, s, box, n, pos, x_build, lim2, flag,
                       x_prune, lim2_in);
}

void pm_launch_set_flag(hipStream_t s, int* flag, int v) {
  if (pm_recording()) { pm_record(-1, 0, 0, 0, 0, nullptr, 0); return; }   // not batched
  hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, flag, v);
}

void pm_launch_invalidate_build(hipStream_t s, int n, int4* x_build) {
  if (pm_recording()) { pm_record(-1, 0, 0, 0, 0, nullptr, 0); return; }   // not batched
  if (n > 0) hipLaunchKernelGGL(k_invalidate_build, dim3((n + 255) / 256), dim3(256), 0, s, n, x_build);
}

void pm_launch_fold_slots(hipStream_t s, int nslot_max, const int* counts, const int* atom_of_slot,
                          long long* force_slot, long long* force_fix, int* flag,
                          unsigned long long* builds, int clear) {
  if (pm_recording()) {
    const FoldSlotsArgs a{counts, atom_of_slot, force_slot, force_fix, flag, builds, clear};
    pm_record(PM_REC_FOLD_SLOTS, 0, nslot_max > 0 ? (nslot_max + 255) / 256 : 0, 256, 0, &a, sizeof(a));
    return;
  }
  if (nslot_max > 0)
    hipLaunchKernelGGL(k_fold_slots, dim3((nslot_max + 255) / 256), dim3(256), 0, s, counts, atom_of_slot,
                       force_slot, force_fix, flag, builds, clear);
}

void pm_launch_prune(hipStream_t s, const PmBox& box, const PmNbLaunch& a) {
  const dim3 g((a.ncl_max + 3) / 4), b(256);
  if (pm_recording()) {
    const PruneArgs q{box, a};
    pm_record(PM_REC_PRUNE, a.m, g.x, 256, 0, &q, sizeof(q));
    return;
  }
  if (a.m == 4) hipLaunchKernelGGL(k_prune<4>, g, b, 0, s, box, a);
  else hipLaunchKernelGGL(k_prune<8>, g, b, 0, s, box, a);
}

void pm_launch_list_stats(hipStream_t s, const int* jcount, const int* counts, unsigned long long* out, int m,
                          int jlanes) {
  hipLaunchKernelGGL(k_list_stats, dim3(1), dim3(256), 0, s, jcount, counts, out, jlanes ? 64 : 64 / m,
                     jlanes ? m : 1);
}

void pm_launch_list(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p) {
  double cut = sqrt(p.cut2) + a.skin;
  float c = (float)(cut * (1.0 + 1e-5) + 1e-4);
  size_t lds = 4 * (6 * a.nstencil + 1 + 2 * a.xs_tcap) * sizeof(int);
  if (pm_recording()) {
    const ListArgs q{box, a, c * c};
    pm_record(PM_REC_LIST, a.m, (a.ncl_max + 3) / 4, 256, (unsigned)lds, &q, sizeof(q));
    return;
  }
  const dim3 g((a.ncl_max + 3) / 4), b(256);
  if (a.m == 4) {
    if (a.bbox_filter) hipLaunchKernelGGL((k_list<4, 1>), g, b, lds, s, box, a, c * c);
    else hipLaunchKernelGGL((k_list<4, 0>), g, b, lds, s, box, a, c * c);
  } else {
    if (a.bbox_filter) hipLaunchKernelGGL((k_list<8, 1>), g, b, lds, s, box, a, c * c);
    else hipLaunchKernelGGL((k_list<8, 0>), g, b, lds, s, box, a, c * c);
  }
}

template <typename T, int LJ, int CK, int CS>
static void launch_pair_t(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p) {
  size_t lds = (LJ != PM_SWF_NONE) ? (size_t)a.ntypes1 * a.ntypes1 * (sizeof(T) == 4 ? 16 : 32) : 0;
  // one block per four clusters (batched launches: per replica)
  const dim3 gb((a.ncl_max + 3) / 4), b(256);
  const dim3 g = gb;
  const bool pa = LJ != PM_SWF_NONE && (sizeof(T) == 4 ? a.atom_lj_f != nullptr : a.atom_lj_d != nullptr);
  // one lane per entry: the LJ table, then each wave's i-force reduction area
  const size_t ljl = (lds + 15) / 16 * 16 + (size_t)4 * PM_JL_LDS;
  if (pm_recording()) {
    // per-atom LJ parameters have no batched kernel: the record is marked
    // unsupported and the batch runs its replicas as chains
    const PairArgs q{box, a, p};
    pm_record(pa ? -1 : PM_REC_PAIR,
              pair_variant(sizeof(T) == 8, LJ, CK, CS, a.want_energy != 0, a.m, a.jlanes != 0), gb.x, 256,
              (unsigned)(a.jlanes ? ljl : lds), &q, sizeof(q));
    return;
  }
  if (a.jlanes) {
    if constexpr (LJ != PM_SWF_NONE) {
      if (pa) {
        if (a.want_energy) hipLaunchKernelGGL((k_pair_jl<T, LJ, CK, CS, true, 8, true>), g, b, ljl, s, box, a, p);
        else hipLaunchKernelGGL((k_pair_jl<T, LJ, CK, CS, false, 8, true>), g, b, ljl, s, box, a, p);
        return;
      }
    }
    if (a.m == 4) {
      if (a.want_energy) hipLaunchKernelGGL((k_pair_jl<T, LJ, CK, CS, true, 4, false>), g, b, ljl, s, box, a, p);
      else hipLaunchKernelGGL((k_pair_jl<T, LJ, CK, CS, false, 4, false>), g, b, ljl, s, box, a, p);
    } else {
      if (a.want_energy) hipLaunchKernelGGL((k_pair_jl<T, LJ, CK, CS, true, 8, false>), g, b, ljl, s, box, a, p);
      else hipLaunchKernelGGL((k_pair_jl<T, LJ, CK, CS, false, 8, false>), g, b, ljl, s, box, a, p);
    }
    return;
  }
  if constexpr (LJ != PM_SWF_NONE) {
    if (pa) {
      // clusters of 8 only (the layout uses 8 whenever per-atom parameters are set)
      if (a.want_energy) hipLaunchKernelGGL((k_pair<T, LJ, CK, CS, true, 8, true>), g, b, lds, s, box, a, p);
      else hipLaunchKernelGGL((k_pair<T, LJ, CK, CS, false, 8, true>), g, b, lds, s, box, a, p);
      return;
    }
  }
  if (a.m == 4) {
    if (a.want_energy) hipLaunchKernelGGL((k_pair<T, LJ, CK, CS, true, 4, false>), g, b, lds, s, box, a, p);
    else hipLaunchKernelGGL((k_pair<T, LJ, CK, CS, false, 4, false>), g, b, lds, s, box, a, p);
  } else {
    if (a.want_energy) hipLaunchKernelGGL((k_pair<T, LJ, CK, CS, true, 8, false>), g, b, lds, s, box, a, p);
    else hipLaunchKernelGGL((k_pair<T, LJ, CK, CS, false, 8, false>), g, b, lds, s, box, a, p);
  }
}

template <typename T>
static void launch_pair_p(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p) {
  const int lj = p.lj_swf, ck = p.coul_kind, cs = p.coul_swf;
#define PM_CASE(L, C, S) \
  if (lj == L && ck == C && (C == PM_COUL_NONE || cs == S)) { launch_pair_t<T, L, C, S>(s, box, a, p); return; }
  // single-term forces (registerForceExemplarsCutoff.cpp:39-63)
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_C1)
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_C2)
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_NONE, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C1, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_CUTOFF, PM_COUL_NONE, PM_SWF_CUTOFF)
  // OneAtomPairTwo ids (:75-96) and fused LJ + Coulomb force pairs
  PM_CASE(PM_SWF_C1, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_C1)
  PM_CASE(PM_SWF_C2, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_CUTOFF, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_C2)
  PM_CASE(PM_SWF_C1, PM_COUL_PLAIN, PM_SWF_C1)
#undef PM_CASE
}

template <typename T, int LJ, int CK, int CS>
static void launch_adjoint_t(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p, int mode,
                             const double4* sv, long long* out) {
  size_t lds = (LJ != PM_SWF_NONE) ? (size_t)a.ntypes1 * a.ntypes1 * (sizeof(T) == 4 ? 16 : 32) : 0;
  const dim3 g((a.ncl_max + 3) / 4), b(256);
  if (mode == 1) hipLaunchKernelGGL((k_adjoint<T, LJ, CK, CS, 1>), g, b, lds, s, box, a, p, sv, out);
  else hipLaunchKernelGGL((k_adjoint<T, LJ, CK, CS, 2>), g, b, lds, s, box, a, p, sv, out);
}

template <typename T>
static void launch_adjoint_p(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p, int mode,
                             const double4* sv, long long* out) {
  const int lj = p.lj_swf, ck = p.coul_kind, cs = p.coul_swf;
#define PM_CASE(L, C, S) \
  if (lj == L && ck == C && (C == PM_COUL_NONE || cs == S)) { launch_adjoint_t<T, L, C, S>(s, box, a, p, mode, sv, out); return; }
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_C1)
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_C2)
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_NONE, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C1, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_CUTOFF, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C1, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_C1)
  PM_CASE(PM_SWF_C2, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_CUTOFF, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_C2)
  PM_CASE(PM_SWF_C1, PM_COUL_PLAIN, PM_SWF_C1)
#undef PM_CASE
}

void pm_launch_adjoint(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p, int precision,
                       int mode, const double4* slot_v, long long* out) {
  if (precision == 1) launch_adjoint_p<double>(s, box, a, p, mode, slot_v, out);
  else launch_adjoint_p<float>(s, box, a, p, mode, slot_v, out);
}

void pm_launch_slot_vec(hipStream_t s, const int* counts, const int* atom_of_slot, const double* v, double4* sv,
                        int nslot_max) {
  hipLaunchKernelGGL(k_slot_vec, dim3((nslot_max + 1 + 255) / 256), dim3(256), 0, s, counts, atom_of_slot, v, sv,
                     nslot_max);
}

void pm_launch_pair(hipStream_t s, const PmBox& box, const PmNbLaunch& a, const PmPairParams& p, int precision) {
  if (precision == 1) launch_pair_p<double>(s, box, a, p);
  else launch_pair_p<float>(s, box, a, p);
}

void pm_launch_diag(hipStream_t s, const PmBox& box, int n, const double* pos, const int* cell_of,
                    const int* cell_start, const int* cell_atoms, const int4* half_stencil, int nhalf,
                    const int* molecule, const int* ex_row, const int* ex_partner, const int* ex_cls,
                    double pair_cut2, int ncells, const int* cell_count, long long* cand, long long* inrc) {
  (void)n; (void)cell_of;
  hipLaunchKernelGGL(k_diag, dim3((ncells + 63) / 64), dim3(64), 0, s, box, pos, cell_start, cell_count,
                     cell_atoms, half_stencil, nhalf, molecule, ex_row, ex_partner, ex_cls, pair_cut2, ncells,
                     cand, inrc);
}

// batched pair kernels: clusters of 8, with or without energies
template <typename T>
static bool launch_pair_batched(hipStream_t s, int variant, const PairArgs* t, const int* pre, int nrep, dim3 g,
                                unsigned lds) {
  const int lj = (variant >> 1) & 3, ck = (variant >> 3) & 3, cs = (variant >> 5) & 3;
  const bool we = (variant >> 7) & 1;
  const bool jl = (variant >> 9) & 1;
  if ((variant >> 8) & 1) return false;   // 4-atom clusters
#define PM_CASE(L, C, S) \
  if (lj == L && ck == C && (C == PM_COUL_NONE || cs == S)) { \
    if (jl && we) hipLaunchKernelGGL((kb_pair<T, L, C, S, true, 8, true>), g, dim3(256), lds, s, t, pre, nrep); \
    else if (jl) hipLaunchKernelGGL((kb_pair<T, L, C, S, false, 8, true>), g, dim3(256), lds, s, t, pre, nrep); \
    else if (we) hipLaunchKernelGGL((kb_pair<T, L, C, S, true, 8, false>), g, dim3(256), lds, s, t, pre, nrep); \
    else hipLaunchKernelGGL((kb_pair<T, L, C, S, false, 8, false>), g, dim3(256), lds, s, t, pre, nrep); \
    return true; }
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_C1)
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_C2)
  PM_CASE(PM_SWF_NONE, PM_COUL_PLAIN, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_NONE, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C1, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_CUTOFF, PM_COUL_NONE, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C1, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_C1)
  PM_CASE(PM_SWF_C2, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_CUTOFF, PM_COUL_EWALD, PM_SWF_CUTOFF)
  PM_CASE(PM_SWF_C2, PM_COUL_PLAIN, PM_SWF_C2)
  PM_CASE(PM_SWF_C1, PM_COUL_PLAIN, PM_SWF_C1)
#undef PM_CASE
  return false;
}

bool pm_launch_batched_nb(hipStream_t s, int kind, int variant, const void* t, const int* pre, int nrep,
                          unsigned total, unsigned threads, unsigned lds) {
  const dim3 g(total), b(threads);
  switch (kind) {
    case PM_REC_BIN: hipLaunchKernelGGL(kb_bin, g, b, 0, s, (const BinArgs*)t, pre, nrep); return true;
    case PM_REC_SCAN_CELLS: hipLaunchKernelGGL(kb_scan_cells, g, b, 0, s, (const ScanCellsArgs*)t, pre, nrep); return true;
    case PM_REC_SCATTER: hipLaunchKernelGGL(kb_scatter, g, b, 0, s, (const ScatterArgs*)t, pre, nrep); return true;
    case PM_REC_FILL_SLOTS: hipLaunchKernelGGL(kb_fill_slots, g, b, 0, s, (const FillSlotsArgs*)t, pre, nrep); return true;
    case PM_REC_CHECK_DISP: hipLaunchKernelGGL(kb_check_disp, g, b, 0, s, (const CheckDispArgs*)t, pre, nrep); return true;
    case PM_REC_FOLD_SLOTS: hipLaunchKernelGGL(kb_fold_slots, g, b, 0, s, (const FoldSlotsArgs*)t, pre, nrep); return true;
    case PM_REC_LIST:
      if (variant == 4) hipLaunchKernelGGL(kb_list<4>, g, b, lds, s, (const ListArgs*)t, pre, nrep);
      else hipLaunchKernelGGL(kb_list<8>, g, b, lds, s, (const ListArgs*)t, pre, nrep);
      return true;
    case PM_REC_PAIR:
      if (variant & 1) return launch_pair_batched<double>(s, variant, (const PairArgs*)t, pre, nrep, g, lds);
      return launch_pair_batched<float>(s, variant, (const PairArgs*)t, pre, nrep, g, lds);
    case PM_REC_PRUNE:
      if (variant == 4) hipLaunchKernelGGL(kb_prune<4>, g, b, 0, s, (const PruneArgs*)t, pre, nrep);
      else hipLaunchKernelGGL(kb_prune<8>, g, b, 0, s, (const PruneArgs*)t, pre, nrep);
      return true;
    default: return false;
  }
}
