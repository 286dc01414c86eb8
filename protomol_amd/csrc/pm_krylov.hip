// pm_krylov.hip -- FP64 vector kernels of the device-resident Lanczos
// iteration over Hessian-vector products: scaled copies, the fixed-point to
// double combination of an H v result, blocked dot products against a stack
// of basis vectors (two passes, fixed order, so results are deterministic),
// the Gram-Schmidt update, normalisation and Ritz-vector assembly.
//
// Basis vectors live in one [nv][n3] array (row k = vector k).  The scalars a
// step needs (Gram-Schmidt coefficients, norms, the alpha/beta of the
// tridiagonal matrix) stay in device memory, so a whole run is queued
// without a host round trip.
#include "pm_internal.h"

#include <math.h>

namespace {

constexpr int KB = 256;          // threads per block
constexpr int KCHUNK = 4096;     // vector elements per block in the dot kernels

__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += red[k];
  return s;   // valid in thread 0
}

// out[k] = in[k] * w[k / 3] (w null: a plain copy)
__global__ void k_scale_copy(int n3, const double* __restrict__ in, const double* __restrict__ w,
                             double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n3) out[k] = w ? in[k] * w[k / 3] : in[k];
}

// out[k] = (fix[k] 2^-40 + dbl[k]) * w[k / 3]: the pair part of H v arrives as
// fixed point, the PME and correction parts as FP64
__global__ void k_fix_combine(int n3, const long long* __restrict__ fix, const double* __restrict__ dbl,
                              const double* __restrict__ w, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3) return;
  double v = (double)fix[k] * PM_FORCE_UNFIX + dbl[k];
  out[k] = w ? v * w[k / 3] : v;
}

// partial[j][b] = sum over block b's chunk of V[j][k] x[k]; grid (nblk, nv)
__global__ __launch_bounds__(KB) void k_dots_partial(int n3, const double* __restrict__ V, const double* __restrict__ x,
                                                     double* __restrict__ partial) {
  __shared__ double red[KB / 64];
  const int b = blockIdx.x, j = blockIdx.y, nblk = gridDim.x;
  const double* vj = V + (size_t)j * n3;
  const int lo = b * KCHUNK, hi = min(n3, lo + KCHUNK);
  double s = 0.0;
  for (int k = lo + threadIdx.x; k < hi; k += KB) s += vj[k] * x[k];
  s = block_sum(s, red);
  if (threadIdx.x == 0) partial[(size_t)j * nblk + b] = s;
}

// out[j] = sum_b partial[j][b]; the block of vector `acc_row` also adds (or
// stores, when `acc_set`) its value into *acc; with `take_sqrt` out holds the
// square root (a norm).  Grid: nv blocks.
__global__ __launch_bounds__(KB) void k_dots_final(int nblk, const double* __restrict__ partial, double* __restrict__ out,
                                                   int acc_row, double* __restrict__ acc, int acc_set, int take_sqrt) {
  __shared__ double red[KB / 64];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += KB) s += partial[(size_t)j * nblk + b];
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    if (take_sqrt) s = sqrt(s);
    out[j] = s;
    if (acc && j == acc_row) *acc = acc_set ? s : *acc + s;
  }
}

// w[k] -= sum_j c[j] V[j][k], j < nv
__global__ void k_gs_update(int n3, int nv, const double* __restrict__ V, const double* __restrict__ c,
                            double* __restrict__ w) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3) return;
  double s = w[k];
  for (int j = 0; j < nv; j++) s -= c[j] * V[(size_t)j * n3 + k];
  w[k] = s;
}

// out[k] = w[k] / *nrm (zero when the norm is below `tiny`: an invariant
// subspace was reached; the host then truncates the tridiagonal matrix)
__global__ void k_normalise(int n3, const double* __restrict__ w, const double* __restrict__ nrm, double tiny,
                            double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3) return;
  const double b = *nrm;
  out[k] = b > tiny ? w[k] / b : 0.0;
}

// Y[i][k] = sum_j S[i][j] V[j][k]: Ritz vectors from the tridiagonal
// eigenvectors (S row i = coefficients of Ritz vector i); grid (blocks, ny)
__global__ void k_ritz(int n3, int nv, const double* __restrict__ V, const double* __restrict__ S,
                       double* __restrict__ Y) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (k >= n3) return;
  double s = 0.0;
  for (int j = 0; j < nv; j++) s += S[(size_t)i * nv + j] * V[(size_t)j * n3 + k];
  Y[(size_t)i * n3 + k] = s;
}

inline dim3 blocks_for(long long n) { return dim3((unsigned)((n + KB - 1) / KB > 0 ? (n + KB - 1) / KB : 1)); }

}  // namespace

int pm_krylov_dot_blocks(int n3) { return (n3 + KCHUNK - 1) / KCHUNK; }

void pm_launch_scale_copy(hipStream_t s, int n3, const double* in, const double* w, double* out) {
  if (n3 > 0) hipLaunchKernelGGL(k_scale_copy, blocks_for(n3), dim3(KB), 0, s, n3, in, w, out);
}

void pm_launch_fix_combine(hipStream_t s, int n3, const long long* fix, const double* dbl, const double* w,
                           double* out) {
  if (n3 > 0) hipLaunchKernelGGL(k_fix_combine, blocks_for(n3), dim3(KB), 0, s, n3, fix, dbl, w, out);
}

void pm_launch_dots(hipStream_t s, int n3, int nv, const double* V, const double* x, double* partial, double* out,
                    int acc_row, double* acc, int acc_set, int take_sqrt) {
  if (n3 <= 0 || nv <= 0) return;
  const int nblk = pm_krylov_dot_blocks(n3);
  hipLaunchKernelGGL(k_dots_partial, dim3(nblk, nv), dim3(KB), 0, s, n3, V, x, partial);
  hipLaunchKernelGGL(k_dots_final, dim3(nv), dim3(KB), 0, s, nblk, partial, out, acc_row, acc, acc_set, take_sqrt);
}

void pm_launch_gs_update(hipStream_t s, int n3, int nv, const double* V, const double* c, double* w) {
  if (n3 > 0 && nv > 0) hipLaunchKernelGGL(k_gs_update, blocks_for(n3), dim3(KB), 0, s, n3, nv, V, c, w);
}

void pm_launch_normalise(hipStream_t s, int n3, const double* w, const double* nrm, double tiny, double* out) {
  if (n3 > 0) hipLaunchKernelGGL(k_normalise, blocks_for(n3), dim3(KB), 0, s, n3, w, nrm, tiny, out);
}

void pm_launch_ritz(hipStream_t s, int n3, int nv, const double* V, const double* S, int ny, double* Y) {
  if (n3 > 0 && nv > 0 && ny > 0) {
    dim3 g = blocks_for(n3);
    g.y = (unsigned)ny;
    hipLaunchKernelGGL(k_ritz, g, dim3(KB), 0, s, n3, nv, V, S, Y);
  }
}
