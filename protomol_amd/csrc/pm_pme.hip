// pm_pme.hip -- smooth particle-mesh Ewald reciprocal term on gfx950.
//
// Follows NonbondedPMEwaldSystemForce::reciprocalTerm
// (forces/NonbondedPMEwaldSystemForce.h:879-932) with Grid<BSpline>
// (base/Grid.h):
//   spread  = Grid::anterpolateCharge (:138-190), deterministic: every
//             stencil contribution is added as a 64-bit fixed-point integer;
//   forward = fftBack, exp(-i) (base/FFTComplex.cpp:79-82), as R2C along z
//             and C2C along y and x (Hermitian symmetry of a real grid);
//   scalarSum (:276-389), fused with the x transforms (forward, multiply by
//             theta(m), inverse) while the line sits in LDS;
//   inverse = fftForward, exp(+i), C2C along y, C2R along z;
//   gather  = Grid::interpolateForce (:193-228), one thread per atom.
// The 1D transforms are Stockham autosort mixed-radix (2, 3, 4, 5, 7) passes
// over lines staged in LDS.
#include "pm_internal.h"
#include "pm_terms.h"

#include <math.h>
#include <type_traits>

namespace {

template <typename T> struct Cx;
template <> struct Cx<float> { typedef float2 t; };
template <> struct Cx<double> { typedef double2 t; };

template <typename T2>
__device__ inline T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T2> __device__ inline T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2> __device__ inline T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by sign * i
template <typename T2> __device__ inline T2 mul_si(T2 a, int sign) {
  T2 r;
  if (sign < 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
  return r;
}

// R-point DFTs, out_k = sum_r v_r exp(sign 2 pi i r k / R)
template <typename T, typename T2, int R> struct Dft;
template <typename T, typename T2> struct Dft<T, T2, 2> {
  static __device__ inline void run(T2* v, int) {
    T2 a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  }
};
template <typename T, typename T2> struct Dft<T, T2, 4> {
  static __device__ inline void run(T2* v, int sign) {
    T2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    T2 c = cadd(v[1], v[3]), d = mul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(a, c); v[1] = cadd(b, d); v[2] = csub(a, c); v[3] = csub(b, d);
  }
};
template <typename T, typename T2> struct Dft<T, T2, 3> {
  static __device__ inline void run(T2* v, int sign) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    T2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    T2 t1; t1.x = v[0].x - (T)0.5 * s.x; t1.y = v[0].y - (T)0.5 * s.y;
    T2 t2 = mul_si(d, sign); t2.x *= h; t2.y *= h;
    v[0] = cadd(v[0], s); v[1] = cadd(t1, t2); v[2] = csub(t1, t2);
  }
};
template <typename T, typename T2> struct Dft<T, T2, 5> {
  static __device__ inline void run(T2* v, int sign) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    T2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    T2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    T2 x0 = v[0];
    T2 p1, p2;
    p1.x = x0.x + c1 * a1.x + c2 * a2.x; p1.y = x0.y + c1 * a1.y + c2 * a2.y;
    p2.x = x0.x + c2 * a1.x + c1 * a2.x; p2.y = x0.y + c2 * a1.y + c1 * a2.y;
    T2 m1; m1.x = s1 * b1.x + s2 * b2.x; m1.y = s1 * b1.y + s2 * b2.y;
    T2 m2; m2.x = s2 * b1.x - s1 * b2.x; m2.y = s2 * b1.y - s1 * b2.y;
    T2 q1 = mul_si(m1, sign), q2 = mul_si(m2, sign);
    v[0].x = x0.x + a1.x + a2.x; v[0].y = x0.y + a1.y + a2.y;
    v[1] = cadd(p1, q1); v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2); v[3] = csub(p2, q2);
  }
};

// the pairs (j, 7 - j) share cos(2 pi j k / 7) and flip sin(2 pi j k / 7):
// out_k = v0 + sum_j a_j cos + sign i sum_j b_j sin, out_(7-k) with -sin,
// a_j = v_j + v_(7-j), b_j = v_j - v_(7-j), j = 1..3 (jk mod 7 folded to 1..3)
template <typename T, typename T2> struct Dft<T, T2, 7> {
  static __device__ inline void run(T2* v, int sign) {
    const T c1 = (T)0.62348980185873353053, c2 = (T)-0.22252093395631440429, c3 = (T)-0.90096886790241912624;
    const T s1 = (T)0.78183148246802980871, s2 = (T)0.97492791218182360702, s3 = (T)0.43388373911755812048;
    const T2 x0 = v[0];
    const T2 a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    const T2 a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    const T2 a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    T2 p1, p2, p3, m1, m2, m3;
    p1.x = x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x; p1.y = x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y;
    p2.x = x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x; p2.y = x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y;
    p3.x = x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x; p3.y = x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y;
    m1.x = s1 * b1.x + s2 * b2.x + s3 * b3.x;       m1.y = s1 * b1.y + s2 * b2.y + s3 * b3.y;
    m2.x = s2 * b1.x - s3 * b2.x - s1 * b3.x;       m2.y = s2 * b1.y - s3 * b2.y - s1 * b3.y;
    m3.x = s3 * b1.x - s1 * b2.x + s2 * b3.x;       m3.y = s3 * b1.y - s1 * b2.y + s2 * b3.y;
    const T2 q1 = mul_si(m1, sign), q2 = mul_si(m2, sign), q3 = mul_si(m3, sign);
    v[0].x = x0.x + a1.x + a2.x + a3.x; v[0].y = x0.y + a1.y + a2.y + a3.y;
    v[1] = cadd(p1, q1); v[6] = csub(p1, q1);
    v[2] = cadd(p2, q2); v[5] = csub(p2, q2);
    v[3] = cadd(p3, q3); v[4] = csub(p3, q3);
  }
};

// Integer division by a block-uniform d >= 1 through a multiply-high:
// floor(v / d) = umulhi(v, ceil(2^32 / d)), exact while v * d <= 2^32
// (the host keeps FFT lengths and line counts within 4096, so every index
// here stays far below that bound).
struct UDiv {
  unsigned d, m;
};
__device__ inline UDiv udiv_make(int d) {
  UDiv u;
  u.d = (unsigned)d;
  u.m = d > 1 ? 0xFFFFFFFFu / (unsigned)d + 1u : 0u;
  return u;
}
__device__ inline int udiv(int v, const UDiv& u) { return u.d > 1 ? (int)__umulhi((unsigned)v, u.m) : v; }

// one Stockham pass of radix R over nl lines of length n held in LDS:
// element m of line l sits at buf[l * ls + m * es].  LF = true gives
// consecutive threads consecutive lines (the butterfly index outermost),
// false consecutive butterflies of one line.
template <typename T, typename T2, int R, bool LF>
__device__ inline void fft_stage(const T2* __restrict__ in, T2* __restrict__ out, int nl, int n, int ls, int es,
                                 int Ns, const T2* __restrict__ tw, int sign) {
  const int nR = n / R;
  const int tws = n / (Ns * R);
  const UDiv dt = udiv_make(LF ? nl : nR);
  const UDiv dn = udiv_make(Ns);
  const int total = nl * nR;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int q = udiv(t, dt);
    int l, j;
    if (LF) { j = q; l = t - q * nl; } else { l = q; j = t - q * nR; }
    const int jq = udiv(j, dn);
    const int k = j - jq * Ns;
    const T2* ip = in + l * ls + j * es;
    T2* op = out + l * ls + (jq * Ns * R + k) * es;
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = ip[r * nR * es];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        T2 w = tw[r * k * tws];          // r*k*tws < R*Ns*tws = n
        if (sign > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    Dft<T, T2, R>::run(v, sign);
#pragma unroll
    for (int r = 0; r < R; r++) op[r * Ns * es] = v[r];
  }
}

// Stockham passes over `nl` lines of length p.n held in LDS (element m of
// line l at buf[l * ls + m * es]); all threads of the block take part.  tw
// points at the n twiddles exp(-2 pi i m / n) already staged in LDS.
// Returns the buffer that holds the result.
template <typename T, typename T2, bool LF = false>
__device__ T2* fft_lines(T2* b0, T2* b1, int nl, const PmFftPlan& p, const T2* tw, int sign, int ls, int es = 1) {
  const int n = p.n;
  T2* in = b0;
  T2* out = b1;
  int Ns = 1;
  for (int st = 0; st < p.nst; st++) {
    const int R = p.rad[st];
    switch (R) {
      case 2: fft_stage<T, T2, 2, LF>(in, out, nl, n, ls, es, Ns, tw, sign); break;
      case 3: fft_stage<T, T2, 3, LF>(in, out, nl, n, ls, es, Ns, tw, sign); break;
      case 4: fft_stage<T, T2, 4, LF>(in, out, nl, n, ls, es, Ns, tw, sign); break;
      case 5: fft_stage<T, T2, 5, LF>(in, out, nl, n, ls, es, Ns, tw, sign); break;
      default: fft_stage<T, T2, 7, LF>(in, out, nl, n, ls, es, Ns, tw, sign); break;
    }
    __syncthreads();
    T2* tmp = in; in = out; out = tmp;
    Ns *= R;
  }
  return in;
}

// stage the twiddles of one transform length into LDS (followed by a barrier
// in the caller's first __syncthreads)
template <typename T2>
__device__ inline void load_twiddles(T2* dst, const PmFftPlan& p, const double* tw) {
  for (int m = threadIdx.x; m < p.n; m += blockDim.x) {
    T2 w;
    w.x = tw[2 * (p.tw + m)];
    w.y = tw[2 * (p.tw + m) + 1];
    dst[m] = w;
  }
}

// Issue priority of this wave over other kernels' waves on its CU (the pair
// kernel runs beside the PME chain on another stream and would otherwise take
// an equal share of every issue cycle): a wave-level setting, the streams
// themselves keep the default priority.
__device__ inline void pme_priority(int prio) {
  switch (prio) {
    case 1: __builtin_amdgcn_s_setprio(1); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 3: __builtin_amdgcn_s_setprio(3); break;
    default: break;
  }
}

// ---------------------------------------------------------- B-spline (P)
// BSpline::set (base/BSpline.cpp:64-117) for an even order P >= 2 (order 2:
// the linear case of :74-79).
template <typename T, int P>
__device__ inline void bspline(T w, T* th, T* dth) {
  if constexpr (P == 2) {
    th[1] = w;
    th[0] = T(1) - w;
    dth[1] = T(1);
    dth[0] = T(-1);
    return;
  }
  th[P - 1] = T(0);
  th[1] = w;
  th[0] = T(1) - w;
#pragma unroll
  for (int k = 2; k < P - 1; k++) {
    const T div = T(1) / T(k);
    th[k] = div * w * th[k - 1];
#pragma unroll
    for (int j = 1; j < k; j++)
      th[k - j] = div * ((w + T(j)) * th[k - j - 1] + (T(k + 1 - j) - w) * th[k - j]);
    th[0] = div * (T(1) - w) * th[0];
  }
  dth[0] = -th[0];
#pragma unroll
  for (int j = 1; j <= P - 1; j++) dth[j] = th[j - 1] - th[j];
  const T div = T(1) / T(P - 1);
  th[P - 1] = div * w * th[P - 2];
#pragma unroll
  for (int j = 1; j < P - 1; j++)
    th[P - 1 - j] = div * ((w + T(j)) * th[P - j - 2] + (T(P - j) - w) * th[P - 1 - j]);
  th[0] = div * (T(1) - w) * th[0];
}

// grid coordinate of a fixed-point position: integer cell and fraction
template <typename T>
__device__ inline void grid_coord(unsigned int X, int n, int& ix, T& w) {
  unsigned long long u = (unsigned long long)X * (unsigned long long)n;
  ix = (int)(u >> 32);
  w = (T)(unsigned int)(u & 0xffffffffull) * (T)2.3283064365386963e-10;  // 2^-32
}
template <typename T>
__device__ inline void grid_coord64(unsigned long long X, int n, int& ix, T& w) {
  unsigned long long hi = __umul64hi(X, (unsigned long long)n);
  unsigned long long lo = X * (unsigned long long)n;
  ix = (int)hi;
  w = (T)lo * (T)5.421010862427522e-20;  // 2^-64
}

// v mod n for v in [-n, 2n): stencil indices ix + off + i with ix in [0, n),
// |off + i| <= 4 (orders up to 8) and n >= 8 (the host refuses smaller grids)
__device__ inline int wrap_grid(int v, int n) {
  if (v < 0) v += n;
  if (v >= n) v -= n;
  return v;
}

template <typename T, int P>
__global__ __launch_bounds__(256) void k_spread(PmPmeLaunch a, const longlong4* __restrict__ pos64) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.counts[0]) return;
  const int at = a.atom_of_slot[s];
  if (at < 0) return;
  int ix, iy, iz;
  T wx, wy, wz;
  if constexpr (sizeof(T) == 4) {
    const int4 p = a.slot_pos[s];
    grid_coord<T>((unsigned)p.x, a.nx, ix, wx);
    grid_coord<T>((unsigned)p.y, a.ny, iy, wy);
    grid_coord<T>((unsigned)p.z, a.nz, iz, wz);
  } else {
    const longlong4 p = pos64[s];
    grid_coord64<T>((unsigned long long)p.x, a.nx, ix, wx);
    grid_coord64<T>((unsigned long long)p.y, a.ny, iy, wy);
    grid_coord64<T>((unsigned long long)p.z, a.nz, iz, wz);
  }
  T tx[P], ty[P], tz[P], dx[P], dy[P], dz[P];
  bspline<T, P>(wx, tx, dx);
  bspline<T, P>(wy, ty, dy);
  bspline<T, P>(wz, tz, dz);
  const int off = -(P - 1) / 2;
  const T q = (T)a.q[at];
  const T scale = (T)PM_GRID_FIX;
#pragma unroll
  for (int i = 0; i < P; i++) {
    const int gx = wrap_grid(ix + off + i, a.nx);
    const T qa = q * tx[i];
#pragma unroll
    for (int j = 0; j < P; j++) {
      const int gy = wrap_grid(iy + off + j, a.ny);
      const T qab = qa * ty[j];
      unsigned long long* row = a.grid_fix + ((long long)gx * a.ny + gy) * a.nz;
#pragma unroll
      for (int k = 0; k < P; k++) {
        const int gz = wrap_grid(iz + off + k, a.nz);
        long long v;
        if constexpr (sizeof(T) == 4) v = __float2ll_rn(qab * tz[k] * scale);
        else v = __double2ll_rn(qab * tz[k] * scale);
        atomicAdd(row + gz, (unsigned long long)v);
      }
    }
  }
}

// Spread by bricks of cells: one block accumulates the atoms of a brick into
// an LDS tile of 64-bit fixed-point charges (exact integer adds, so the
// result does not depend on the order), then adds the tile to the global grid
// with coalesced 64-bit atomics.  Stencil points outside the tile (atoms on a
// brick face whose rounding lands them one point further) go to the global
// grid directly.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_spread_tile(PmPmeLaunch a, const longlong4* __restrict__ pos64) {
  pme_priority(a.prio);
  extern __shared__ unsigned long long s_tile[];
  __shared__ int s_cs[64], s_cc[64], s_pre[65];
  const int bz = blockIdx.x % a.nbrick[2];
  const int by = (blockIdx.x / a.nbrick[2]) % a.nbrick[1];
  const int bx = blockIdx.x / (a.nbrick[1] * a.nbrick[2]);
  const int tnx = a.tile[0], tny = a.tile[1], tnz = a.tile[2];
  const int ntile = tnx * tny * tnz;
  for (int t = threadIdx.x; t < ntile; t += blockDim.x) s_tile[t] = 0ull;
  // cells of the brick
  const int cx0 = bx * a.brick[0], cy0 = by * a.brick[1], cz0 = bz * a.brick[2];
  const int ex = min(a.brick[0], a.ncell[0] - cx0), ey = min(a.brick[1], a.ncell[1] - cy0),
            ez = min(a.brick[2], a.ncell[2] - cz0);
  const int ncb = ex * ey * ez;
  if (threadIdx.x < ncb) {
    const int i = threadIdx.x / (ey * ez), j = (threadIdx.x / ez) % ey, k = threadIdx.x % ez;
    const int c = ((cx0 + i) * a.ncell[1] + (cy0 + j)) * a.ncell[2] + (cz0 + k);
    s_cs[threadIdx.x] = a.cell_slot[c];
    s_cc[threadIdx.x] = a.cell_count[c];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int u = 0; u < ncb; u++) { s_pre[u] = acc; acc += s_cc[u]; }
    s_pre[ncb] = acc;
  }
  // tile origin: grid point of the brick's lower corner, one point of margin
  // below the stencil start
  const int off = -(P - 1) / 2;
  int t0[3];
  const int nn[3] = {a.nx, a.ny, a.nz};
  const int c0[3] = {cx0, cy0, cz0};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (a.tile[d] >= nn[d]) { t0[d] = 0; continue; }
    double u = a.cell_lo_u[d] + c0[d] * a.cell_du[d];
    int g = (int)floor(u) + off - 1;
    g %= nn[d];
    if (g < 0) g += nn[d];
    t0[d] = g;
  }
  __syncthreads();
  const int natoms = s_pre[ncb];
  const T scale = (T)PM_GRID_FIX;
  for (int t = threadIdx.x; t < natoms; t += blockDim.x) {
    int cb = 0;
    while (cb + 1 < ncb && s_pre[cb + 1] <= t) cb++;
    const int sl = s_cs[cb] + (t - s_pre[cb]);
    const int at = a.atom_of_slot[sl];
    int ix, iy, iz;
    T wx, wy, wz;
    if constexpr (sizeof(T) == 4) {
      const int4 p = a.slot_pos[sl];
      grid_coord<T>((unsigned)p.x, a.nx, ix, wx);
      grid_coord<T>((unsigned)p.y, a.ny, iy, wy);
      grid_coord<T>((unsigned)p.z, a.nz, iz, wz);
    } else {
      const longlong4 p = pos64[sl];
      grid_coord64<T>((unsigned long long)p.x, a.nx, ix, wx);
      grid_coord64<T>((unsigned long long)p.y, a.ny, iy, wy);
      grid_coord64<T>((unsigned long long)p.z, a.nz, iz, wz);
    }
    T tx[P], ty[P], tz[P], dx[P], dy[P], dz[P];
    bspline<T, P>(wx, tx, dx);
    bspline<T, P>(wy, ty, dy);
    bspline<T, P>(wz, tz, dz);
    const T q = (T)a.q[at];
#pragma unroll
    for (int i = 0; i < P; i++) {
      const int gx = wrap_grid(ix + off + i, a.nx);
      int lx = gx - t0[0]; if (lx < 0) lx += a.nx;
      const T qa = q * tx[i];
#pragma unroll
      for (int j = 0; j < P; j++) {
        const int gy = wrap_grid(iy + off + j, a.ny);
        int ly = gy - t0[1]; if (ly < 0) ly += a.ny;
        const T qab = qa * ty[j];
#pragma unroll
        for (int k = 0; k < P; k++) {
          const int gz = wrap_grid(iz + off + k, a.nz);
          int lz = gz - t0[2]; if (lz < 0) lz += a.nz;
          long long v;
          if constexpr (sizeof(T) == 4) v = __float2ll_rn(qab * tz[k] * scale);
          else v = __double2ll_rn(qab * tz[k] * scale);
          if (lx < tnx && ly < tny && lz < tnz)
            atomicAdd(&s_tile[(lx * tny + ly) * tnz + lz], (unsigned long long)v);
          else
            atomicAdd(a.grid_fix + ((long long)gx * a.ny + gy) * a.nz + gz, (unsigned long long)v);
        }
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ntile; t += blockDim.x) {
    const unsigned long long v = s_tile[t];
    if (v == 0ull) continue;
    const int lz = t % tnz, ly = (t / tnz) % tny, lx = t / (tny * tnz);
    const int gx = (t0[0] + lx) % a.nx, gy = (t0[1] + ly) % a.ny, gz = (t0[2] + lz) % a.nz;
    atomicAdd(a.grid_fix + ((long long)gx * a.ny + gy) * a.nz + gz, v);
  }
}

// TAIL: the same thread also adds its atom's correction and bonded forces
template <typename T, int P, bool TAIL>
__global__ __launch_bounds__(256) void k_gather(PmPmeLaunch a, const longlong4* __restrict__ pos64, PmAtomTail tl) {
  pme_priority(a.prio);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.counts[0]) return;
  const int at = a.atom_of_slot[s];
  if (at < 0) return;
  int ix, iy, iz;
  T wx, wy, wz;
  if constexpr (sizeof(T) == 4) {
    const int4 p = a.slot_pos[s];
    grid_coord<T>((unsigned)p.x, a.nx, ix, wx);
    grid_coord<T>((unsigned)p.y, a.ny, iy, wy);
    grid_coord<T>((unsigned)p.z, a.nz, iz, wz);
  } else {
    const longlong4 p = pos64[s];
    grid_coord64<T>((unsigned long long)p.x, a.nx, ix, wx);
    grid_coord64<T>((unsigned long long)p.y, a.ny, iy, wy);
    grid_coord64<T>((unsigned long long)p.z, a.nz, iz, wz);
  }
  T tx[P], ty[P], tz[P], dx[P], dy[P], dz[P];
  bspline<T, P>(wx, tx, dx);
  bspline<T, P>(wy, ty, dy);
  bspline<T, P>(wz, tz, dz);
  const int off = -(P - 1) / 2;
  const T* g = reinterpret_cast<const T*>(a.grid_r);
  T fx = 0, fy = 0, fz = 0;
#pragma unroll
  for (int i = 0; i < P; i++) {
    const int gx = wrap_grid(ix + off + i, a.nx);
#pragma unroll
    for (int j = 0; j < P; j++) {
      const int gy = wrap_grid(iy + off + j, a.ny);
      const T xij = dx[i] * ty[j];
      const T yij = tx[i] * dy[j];
      const T zij = tx[i] * ty[j];
      const T* row = g + ((long long)gx * a.ny + gy) * a.nz;
#pragma unroll
      for (int k = 0; k < P; k++) {
        const int gz = wrap_grid(iz + off + k, a.nz);
        const T term = row[gz];
        fx -= term * xij * tz[k];
        fy -= term * yij * tz[k];
        fz -= term * zij * dz[k];
      }
    }
  }
  const double q = a.q[at];
  long long F[3] = {pm_force_fix((double)fx * a.hr[0] * q, a.err),
                    pm_force_fix((double)fy * a.hr[1] * q, a.err),
                    pm_force_fix((double)fz * a.hr[2] * q, a.err)};
  if (TAIL) {
    double f[3];
    if (tl.corr) {
      corr_atom_force(tl.box, at, tl.pos, tl.q, tl.row, tl.partner, tl.cls, tl.alpha, tl.csf, f);
      for (int k = 0; k < 3; k++) F[k] += pm_force_fix(f[k], a.err);
    }
    if (tl.bonded) {
      bonded_atom_force(tl.box, tl.b, at, f);
      for (int k = 0; k < 3; k++) F[k] += pm_force_fix(f[k], a.err);
    }
  }
  a.force_fix[3 * at] += F[0];
  a.force_fix[3 * at + 1] += F[1];
  a.force_fix[3 * at + 2] += F[2];
}

// ---------------------------------------------------------- PME adjoint
// B-spline weights with their first and second derivatives: the second
// derivative of the order-P spline is the order P-2 spline differenced twice,
// d2th[k] = m[k-2] - 2 m[k-1] + m[k] (m: order P-2, zero outside its range).
template <typename T, int P>
__device__ inline void bspline2(T w, T* th, T* dth, T* d2th) {
  if constexpr (P == 2) {
    // linear pieces: no second derivative
    bspline<T, 2>(w, th, dth);
    d2th[0] = d2th[1] = T(0);
    return;
  }
  T m[P];
#pragma unroll
  for (int k = 0; k < P; k++) m[k] = T(0);
  th[P - 1] = T(0);
  th[1] = w;
  th[0] = T(1) - w;
#pragma unroll
  for (int k = 2; k < P - 1; k++) {
    if (k == P - 2) {
#pragma unroll
      for (int j = 0; j < P - 2; j++) m[j] = th[j];
    }
    const T div = T(1) / T(k);
    th[k] = div * w * th[k - 1];
#pragma unroll
    for (int j = 1; j < k; j++)
      th[k - j] = div * ((w + T(j)) * th[k - j - 1] + (T(k + 1 - j) - w) * th[k - j]);
    th[0] = div * (T(1) - w) * th[0];
  }
#pragma unroll
  for (int k = 0; k < P; k++) {
    const T a0 = k >= 2 && k - 2 < P - 2 ? m[k - 2] : T(0);
    const T a1 = k >= 1 && k - 1 < P - 2 ? m[k - 1] : T(0);
    const T a2 = k < P - 2 ? m[k] : T(0);
    d2th[k] = a0 - T(2) * a1 + a2;
  }
  dth[0] = -th[0];
#pragma unroll
  for (int j = 1; j <= P - 1; j++) dth[j] = th[j - 1] - th[j];
  const T div = T(1) / T(P - 1);
  th[P - 1] = div * w * th[P - 2];
#pragma unroll
  for (int j = 1; j < P - 1; j++)
    th[P - 1 - j] = div * ((w + T(j)) * th[P - j - 2] + (T(P - j) - w) * th[P - 1 - j]);
  th[0] = div * (T(1) - w) * th[0];
}

template <typename T>
__device__ inline void slot_grid(const PmPmeLaunch& a, const longlong4* pos64, int s, int& ix, int& iy, int& iz,
                                 T& wx, T& wy, T& wz) {
  if constexpr (sizeof(T) == 4) {
    const int4 p = a.slot_pos[s];
    grid_coord<T>((unsigned)p.x, a.nx, ix, wx);
    grid_coord<T>((unsigned)p.y, a.ny, iy, wy);
    grid_coord<T>((unsigned)p.z, a.nz, iz, wz);
  } else {
    const longlong4 p = pos64[s];
    grid_coord64<T>((unsigned long long)p.x, a.nx, ix, wx);
    grid_coord64<T>((unsigned long long)p.y, a.ny, iy, wy);
    grid_coord64<T>((unsigned long long)p.z, a.nz, iz, wz);
  }
}

// Directional derivative of the charge grid along v:
//   dQ(g) = sum_j q_j grad_j W(g - x_j) . v_j   (64-bit fixed point, atomics)
template <typename T, int P>
__global__ __launch_bounds__(256) void k_spread_dir(PmPmeLaunch a, const longlong4* __restrict__ pos64,
                                                    const double4* __restrict__ slot_v) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.counts[0]) return;
  const int at = a.atom_of_slot[s];
  if (at < 0) return;
  int ix, iy, iz;
  T wx, wy, wz;
  slot_grid<T>(a, pos64, s, ix, iy, iz, wx, wy, wz);
  T tx[P], ty[P], tz[P], dx[P], dy[P], dz[P];
  bspline<T, P>(wx, tx, dx);
  bspline<T, P>(wy, ty, dy);
  bspline<T, P>(wz, tz, dz);
  const double4 v = slot_v[s];
  const T q = (T)a.q[at];
  const T vx = (T)(v.x * a.hr[0]) * q, vy = (T)(v.y * a.hr[1]) * q, vz = (T)(v.z * a.hr[2]) * q;
  const int off = -(P - 1) / 2;
  const T scale = (T)PM_GRID_FIX;
#pragma unroll
  for (int i = 0; i < P; i++) {
    const int gx = wrap_grid(ix + off + i, a.nx);
#pragma unroll
    for (int j = 0; j < P; j++) {
      const int gy = wrap_grid(iy + off + j, a.ny);
      const T cxy = vx * dx[i] * ty[j] + vy * tx[i] * dy[j];
      const T cz = vz * tx[i] * ty[j];
      unsigned long long* row = a.grid_fix + ((long long)gx * a.ny + gy) * a.nz;
#pragma unroll
      for (int k = 0; k < P; k++) {
        const int gz = wrap_grid(iz + off + k, a.nz);
        const T val = cxy * tz[k] + cz * dz[k];
        long long iv;
        if constexpr (sizeof(T) == 4) iv = __float2ll_rn(val * scale);
        else iv = __double2ll_rn(val * scale);
        atomicAdd(row + gz, (unsigned long long)iv);
      }
    }
  }
}

// MODE 0: out[at] += phi(x_at) (dE/dq of the reciprocal energy)
// MODE 1: out[3 at + c] += q (grad grad W . v)_c summed against phi
// MODE 2: out[3 at + c] += q (grad W)_c summed against dphi (the grid holds it)
template <typename T, int P, int MODE>
__global__ __launch_bounds__(256) void k_gather_adj(PmPmeLaunch a, const longlong4* __restrict__ pos64,
                                                    const double4* __restrict__ slot_v, double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.counts[0]) return;
  const int at = a.atom_of_slot[s];
  if (at < 0) return;
  int ix, iy, iz;
  T wx, wy, wz;
  slot_grid<T>(a, pos64, s, ix, iy, iz, wx, wy, wz);
  T tx[P], ty[P], tz[P], dx[P], dy[P], dz[P], ex[P], ey[P], ez[P];
  bspline2<T, P>(wx, tx, dx, ex);
  bspline2<T, P>(wy, ty, dy, ey);
  bspline2<T, P>(wz, tz, dz, ez);
  const T h0 = (T)a.hr[0], h1 = (T)a.hr[1], h2 = (T)a.hr[2];
  const double4 v = slot_v[s];
  const T vx = (T)v.x, vy = (T)v.y, vz = (T)v.z;
  const int off = -(P - 1) / 2;
  const T* g = reinterpret_cast<const T*>(a.grid_r);
  T r0 = 0, r1 = 0, r2 = 0;
#pragma unroll
  for (int i = 0; i < P; i++) {
    const int gx = wrap_grid(ix + off + i, a.nx);
#pragma unroll
    for (int j = 0; j < P; j++) {
      const int gy = wrap_grid(iy + off + j, a.ny);
      const T* row = g + ((long long)gx * a.ny + gy) * a.nz;
#pragma unroll
      for (int k = 0; k < P; k++) {
        const int gz = wrap_grid(iz + off + k, a.nz);
        const T phi = row[gz];
        if constexpr (MODE == 0) {
          r0 += phi * tx[i] * ty[j] * tz[k];
        } else if constexpr (MODE == 1) {
          // Hessian of W(g - x) with respect to x, applied to v
          const T hxx = h0 * h0 * ex[i] * ty[j] * tz[k], hyy = h1 * h1 * tx[i] * ey[j] * tz[k];
          const T hzz = h2 * h2 * tx[i] * ty[j] * ez[k];
          const T hxy = h0 * h1 * dx[i] * dy[j] * tz[k], hxz = h0 * h2 * dx[i] * ty[j] * dz[k];
          const T hyz = h1 * h2 * tx[i] * dy[j] * dz[k];
          r0 += phi * (hxx * vx + hxy * vy + hxz * vz);
          r1 += phi * (hxy * vx + hyy * vy + hyz * vz);
          r2 += phi * (hxz * vx + hyz * vy + hzz * vz);
        } else {
          r0 += phi * h0 * dx[i] * ty[j] * tz[k];
          r1 += phi * h1 * tx[i] * dy[j] * tz[k];
          r2 += phi * h2 * tx[i] * ty[j] * dz[k];
        }
      }
    }
  }
  if constexpr (MODE == 0) {
    out[at] += (double)r0;
  } else {
    const double q = a.q[at];
    out[3 * at] += q * (double)r0;
    out[3 * at + 1] += q * (double)r1;
    out[3 * at + 2] += q * (double)r2;
  }
}

// ------------------------------------------------------------- FFT passes
// The complex half grid is stored line-major, [x][y][kz], for the separate
// y and x passes, or plane-major, [kz][x][y], when one block transforms a
// whole kz plane (a.plane).  The z passes pad their LDS lines to n + 1
// points, so that the transposed plane-major reads and writes of
// consecutive lines fall in distinct LDS banks.

// z forward: fixed-point grid -> real -> complex half lines; clears the
// fixed-point grid for the next spread.
template <typename T>
__global__ __launch_bounds__(256) void k_fft_z_fwd(PmPmeLaunch a, PmFftPlan p, int TL) {
  pme_priority(a.prio);
  typedef typename Cx<T>::t T2;
  extern __shared__ unsigned char s_raw[];
  const int n = p.n, ls = n + 1, nl_tot = a.nx * a.ny, nzc = a.nzc;
  T2* b0 = reinterpret_cast<T2*>(s_raw);
  T2* b1 = b0 + TL * ls;
  T2* tws = b1 + TL * ls;
  load_twiddles(tws, p, a.tw);
  const int line0 = blockIdx.x * TL;
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
    const int l = e / n, k = e - l * n;
    const int line = line0 + l;
    T2 v; v.x = 0; v.y = 0;
    if (line < nl_tot) {
      unsigned long long* gp = a.grid_fix + (long long)line * n + k;
      v.x = (T)((double)(long long)(*gp) * PM_GRID_UNFIX);
      *gp = 0ull;
    }
    b0[l * ls + k] = v;
  }
  __syncthreads();
  T2* r = fft_lines<T, T2>(b0, b1, TL, p, tws, -1, ls);
  T2* gc = reinterpret_cast<T2*>(a.grid_c);
  if (a.plane) {
    for (int e = threadIdx.x; e < TL * nzc; e += blockDim.x) {
      const int k = e / TL, l = e - k * TL;
      const int line = line0 + l;
      if (line < nl_tot) gc[(long long)k * nl_tot + line] = r[l * ls + k];
    }
  } else {
    for (int e = threadIdx.x; e < TL * nzc; e += blockDim.x) {
      const int l = e / nzc, k = e - l * nzc;
      const int line = line0 + l;
      if (line < nl_tot) gc[(long long)line * nzc + k] = r[l * ls + k];
    }
  }
}

// Spread fused with the z forward transform.  A block owns PM_SZ_LX x
// a.spread_zfwd (4, 2 or 1: as many as fit LDS) whole z lines of the grid:
// it walks the slots of the cell columns whose atoms can reach those lines
// (the cells from the last layout build, widened by how far an atom may have
// drifted from its cell since: a.drift_u grid points), adds each atom's stencil points on its own lines into an LDS
// accumulator of 64-bit fixed-point charges (integer adds: the same sums as
// the spread kernels make, in any order), then converts the lines to real
// values and transforms them along z as k_fft_z_fwd does.  Every grid point
// is owned by one block, so there are no global atomics and no charge grid
// to clear.
#define PM_SZ_LX 4
#define PM_SZ_MAXCOL 256
template <typename T, int P>
__global__ __launch_bounds__(256) void k_spread_zfwd(PmPmeLaunch a, PmFftPlan p, const longlong4* __restrict__ pos64) {
  pme_priority(a.prio);
  typedef typename Cx<T>::t T2;
  constexpr int LX = PM_SZ_LX;
  const int LY = a.spread_zfwd, TL = LX * LY;
  extern __shared__ unsigned char s_raw[];
  __shared__ int s_start[PM_SZ_MAXCOL], s_pre[PM_SZ_MAXCOL + 1];
  __shared__ int s_ncol;
  const int n = p.n, ls = n + 1, nzc = a.nzc;
  T2* b0 = reinterpret_cast<T2*>(s_raw);
  T2* b1 = b0 + TL * ls;
  T2* tws = b1 + TL * ls;
  // the accumulator shares b1's space (TL * n 8-byte words <= TL * ls T2)
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(b1);
  load_twiddles(tws, p, a.tw);
  const int nby = (a.ny + LY - 1) / LY;
  const int x0 = (blockIdx.x / nby) * LX, y0 = (blockIdx.x % nby) * LY;
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) acc[e] = 0ull;
  constexpr int off = -(P - 1) / 2;
  if (threadIdx.x < 64) {
    // cell columns: per axis the cells that meet the grid coordinates
    // [x0 - off - P + 1, x0 + LX - off), widened by the drift margin (an atom
    // whose stencil reaches one of the block's lines lies there)
    const int lane = threadIdx.x;
    int c0[2], nc[2];
    const int g0[2] = {x0, y0}, lw[2] = {LX, LY};
#pragma unroll
    for (int d = 0; d < 2; d++) {
      const double ua = g0[d] - off - P + 1 - a.drift_u[d], ub = g0[d] + lw[d] - off + a.drift_u[d];
      const int lo = (int)floor((ua - a.cell_lo_u[d]) / a.cell_du[d]);
      const int hi = (int)floor((ub - a.cell_lo_u[d]) / a.cell_du[d]);
      c0[d] = lo;
      nc[d] = hi - lo + 1;
      if (nc[d] >= a.ncell[d]) { c0[d] = 0; nc[d] = a.ncell[d]; }
    }
    const int ncol = nc[0] * nc[1];   // <= PM_SZ_MAXCOL (checked by the host)
    const int nz = a.ncell[2];
    int carry = 0;
    for (int b = 0; b < ncol; b += 64) {
      const int q = b + lane;
      int st = 0, len = 0;
      if (q < ncol) {
        int cx = (c0[0] + q / nc[1]) % a.ncell[0], cy = (c0[1] + q % nc[1]) % a.ncell[1];
        if (cx < 0) cx += a.ncell[0];
        if (cy < 0) cy += a.ncell[1];
        const int col = (cx * a.ncell[1] + cy) * nz;
        st = a.cell_slot[col];
        len = a.cell_slot[col + nz - 1] + a.cell_count[col + nz - 1] - st;
        s_start[q] = st;
      }
      int inc = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
      }
      if (q < ncol) s_pre[q] = carry + inc - len;
      carry += __shfl(inc, 63, 64);
    }
    if (lane == 0) { s_pre[ncol] = carry; s_ncol = ncol; }
  }
  __syncthreads();
  const int ncol = s_ncol, total = s_pre[ncol];
  const T scale = (T)PM_GRID_FIX;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    // the column holding candidate t: the last prefix <= t
    int lo = 0, hi = ncol - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_pre[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int sl = s_start[lo] + (t - s_pre[lo]);
    int ix, iy, iz;
    T wx, wy, wz;
    if constexpr (sizeof(T) == 4) {
      const int4 q = a.slot_pos[sl];
      grid_coord<T>((unsigned)q.x, a.nx, ix, wx);
      grid_coord<T>((unsigned)q.y, a.ny, iy, wy);
      grid_coord<T>((unsigned)q.z, a.nz, iz, wz);
    } else {
      const longlong4 q = pos64[sl];
      grid_coord64<T>((unsigned long long)q.x, a.nx, ix, wx);
      grid_coord64<T>((unsigned long long)q.y, a.ny, iy, wy);
      grid_coord64<T>((unsigned long long)q.z, a.nz, iz, wz);
    }
    // the stencil's x and y points on this block's lines: offsets from the
    // block's first line, mod n
    int dx0 = ix + off - x0, dy0 = iy + off - y0;   // in (-n - P, n)
    if (dx0 < 0) dx0 += a.nx;
    if (dx0 < 0) dx0 += a.nx;
    if (dy0 < 0) dy0 += a.ny;
    if (dy0 < 0) dy0 += a.ny;
    int lxs[P], lys[P];
    unsigned int mx = 0, my = 0;
#pragma unroll
    for (int i = 0; i < P; i++) {
      lxs[i] = dx0 + i >= a.nx ? dx0 + i - a.nx : dx0 + i;
      if (lxs[i] < LX) mx |= 1u << i;
      lys[i] = dy0 + i >= a.ny ? dy0 + i - a.ny : dy0 + i;
      if (lys[i] < LY) my |= 1u << i;
    }
    if (mx == 0 || my == 0) continue;
    const int at = a.atom_of_slot[sl];
    T tx[P], ty[P], tz[P], dx[P], dy[P], dz[P];
    bspline<T, P>(wx, tx, dx);
    bspline<T, P>(wy, ty, dy);
    bspline<T, P>(wz, tz, dz);
    const T q = (T)a.q[at];
    int gzs[P];
#pragma unroll
    for (int k = 0; k < P; k++) gzs[k] = wrap_grid(iz + off + k, a.nz);
#pragma unroll
    for (int i = 0; i < P; i++) {
      if (!((mx >> i) & 1u)) continue;
      const T qa = q * tx[i];
#pragma unroll
      for (int j = 0; j < P; j++) {
        if (!((my >> j) & 1u)) continue;
        const T qab = qa * ty[j];
        unsigned long long* line = acc + (lxs[i] * LY + lys[j]) * n;
#pragma unroll
        for (int k = 0; k < P; k++) {
          long long v;
          if constexpr (sizeof(T) == 4) v = __float2ll_rn(qab * tz[k] * scale);
          else v = __double2ll_rn(qab * tz[k] * scale);
          atomicAdd(line + gzs[k], (unsigned long long)v);
        }
      }
    }
  }
  __syncthreads();
  // real lines into b0 (the accumulator in b1 is read first, then b1 is free)
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
    const int l = e / n, k = e - l * n;
    T2 v;
    v.x = (T)((double)(long long)acc[e] * PM_GRID_UNFIX);
    v.y = 0;
    b0[l * ls + k] = v;
  }
  __syncthreads();
  T2* r = fft_lines<T, T2>(b0, b1, TL, p, tws, -1, ls);
  T2* gc = reinterpret_cast<T2*>(a.grid_c);
  const int nl_tot = a.nx * a.ny;
  for (int e = threadIdx.x; e < TL * nzc; e += blockDim.x) {
    int l, k;
    if (a.plane) { k = e / TL; l = e - k * TL; }
    else { l = e / nzc; k = e - l * nzc; }
    const int gx = x0 + l / LY, gy = y0 + l % LY;
    if (gx >= a.nx || gy >= a.ny) continue;
    const int line = gx * a.ny + gy;
    if (a.plane) gc[(long long)k * nl_tot + line] = r[l * ls + k];
    else gc[(long long)line * nzc + k] = r[l * ls + k];
  }
}

// energy and virial partial sums of the influence-function step
struct PmERed {
  double en, vxx, vxy, vxz, vyy, vyz, vzz;
};

// Grid::scalarSum (base/Grid.h:276-389) at one wave vector (kx, ky, kz) of
// the half grid: adds v's energy and virial terms (weight 2 for the kz that
// also stand for their conjugates) and returns v * theta(m).
// theta(m) at (kx, ky, kz) != 0 with m = (mi, mj, mk) and m2 = |m|^2
__device__ inline double theta_at(const PmPmeLaunch& a, int kx, int ky, int kz, double& mi, double& mj, double& mk,
                                  double& m2) {
  const int i0 = kx <= a.nx / 2 ? kx : kx - a.nx;
  const int j0 = ky <= a.ny / 2 ? ky : ky - a.ny;
  mi = i0 * a.Lr[0];
  mj = j0 * a.Lr[1];
  mk = kz * a.Lr[2];
  m2 = mi * mi + mj * mj + mk * mk;
  const double* bm = a.bmod;
  const double* ex = a.expf_tab;
  return bm[kx] * bm[a.nx + ky] * bm[a.nx + a.ny + kz] * (ex[kx] * a.inv_pi_v) * ex[a.nx + ky] *
         ex[a.nx + a.ny + kz] / m2;
}

template <typename T, typename T2>
__device__ inline T2 influence(const PmPmeLaunch& a, int kx, int ky, int kz, T2 v, PmERed& s) {
  T theta = 0;
  if (kx != 0 || ky != 0 || kz != 0) {
    double mi, mj, mk, m2;
    const double th = theta_at(a, kx, ky, kz, mi, mj, mk, m2);
    theta = (T)th;
    const double w = (kz == 0 || ((a.nz & 1) == 0 && kz == a.nz / 2)) ? 1.0 : 2.0;
    const double q2 = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    const double eg = w * q2 * th;
    s.en += eg;
    if (a.want_virial) {
      const double vv = 2.0 * (1.0 / m2 + a.fac);
      s.vxx += eg * (1.0 - vv * mi * mi);
      s.vxy -= eg * (vv * mi * mj);
      s.vxz -= eg * (vv * mi * mk);
      s.vyy += eg * (1.0 - vv * mj * mj);
      s.vyz -= eg * (vv * mj * mk);
      s.vzz += eg * (1.0 - vv * mk * mk);
    }
  }
  v.x *= theta;
  v.y *= theta;
  return v;
}

// theta(m) over the line-major half grid, rounded as influence() rounds it
template <typename T>
__global__ __launch_bounds__(256) void k_theta(PmPmeLaunch a) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ng = (long long)a.nx * a.ny * a.nzc;
  if (e >= ng) return;
  const int kz = (int)(e % a.nzc), ky = (int)((e / a.nzc) % a.ny), kx = (int)(e / ((long long)a.nzc * a.ny));
  T th = 0;
  if (kx != 0 || ky != 0 || kz != 0) {
    double mi, mj, mk, m2;
    th = (T)theta_at(a, kx, ky, kz, mi, mj, mk, m2);
  }
  const_cast<T*>(reinterpret_cast<const T*>(a.theta))[e] = th;
}

// block sum (fixed order: lanes, then waves) of the partials into this
// block's energy record; ends with the block's threads synchronised
__device__ inline void store_block_energy(const PmPmeLaunch& a, const PmERed& s, double (*s_red)[8]) {
  const double part[7] = {s.en, s.vxx, s.vxy, s.vxz, s.vyy, s.vyz, s.vzz};
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int u = 0; u < 7; u++) {
    double t = part[u];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (lane == 0) s_red[wv][u] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[7];
    for (int u = 0; u < 7; u++) {
      double t = 0;
      for (int w2 = 0; w2 < (int)(blockDim.x >> 6); w2++) t += s_red[w2][u];
      tot[u] = t;
    }
    double* ep = a.epart + (long long)blockIdx.x * PM_NE;
    ep[PM_E_COUL] = tot[0] * 0.5;
    for (int u = 1; u < PM_E_VIR; u++) ep[u] = 0.0;
    ep[PM_E_VIR + 0] = tot[1] * 0.5; ep[PM_E_VIR + 1] = tot[2] * 0.5; ep[PM_E_VIR + 2] = tot[3] * 0.5;
    ep[PM_E_VIR + 3] = tot[2] * 0.5; ep[PM_E_VIR + 4] = tot[4] * 0.5; ep[PM_E_VIR + 5] = tot[5] * 0.5;
    ep[PM_E_VIR + 6] = tot[3] * 0.5; ep[PM_E_VIR + 7] = tot[5] * 0.5; ep[PM_E_VIR + 8] = tot[6] * 0.5;
  }
}

// y pass (either direction) on the line-major grid: the lines (x, kz) in
// x-major order, TL consecutive lines per block (a block may span two x)
template <typename T>
__global__ __launch_bounds__(256) void k_fft_y(PmPmeLaunch a, PmFftPlan p, int TL, int sign) {
  pme_priority(a.prio);
  typedef typename Cx<T>::t T2;
  extern __shared__ unsigned char s_raw[];
  T2* b0 = reinterpret_cast<T2*>(s_raw);
  T2* b1 = b0 + TL * p.n;
  T2* tws = b1 + TL * p.n;
  load_twiddles(tws, p, a.tw);
  const int n = p.n;  // ny
  const int nl = a.nx * a.nzc, line0 = blockIdx.x * TL;
  __shared__ long long s_off[64];   // grid offset of each line's y = 0 point (-1: past the grid)
  if (threadIdx.x < TL) {
    const int line = line0 + threadIdx.x;
    const int x = line / a.nzc, kz = line - x * a.nzc;
    s_off[threadIdx.x] = line < nl ? (long long)x * a.ny * a.nzc + kz : -1;
  }
  __syncthreads();
  T2* gc = reinterpret_cast<T2*>(a.grid_c);
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
    const int y = e / TL, l = e - y * TL;
    const long long o = s_off[l];
    T2 v; v.x = 0; v.y = 0;
    if (o >= 0) v = gc[o + (long long)y * a.nzc];
    b0[l * n + y] = v;
  }
  __syncthreads();
  T2* r = fft_lines<T, T2>(b0, b1, TL, p, tws, sign, n);
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
    const int y = e / TL, l = e - y * TL;
    const long long o = s_off[l];
    if (o >= 0) gc[o + (long long)y * a.nzc] = r[l * n + y];
  }
}

// x pass with the influence function on the line-major grid: forward
// transform, multiply by theta(m), inverse transform; the lines (y, kz) in
// y-major order, TL consecutive lines per block
template <typename T>
__global__ __launch_bounds__(256) void k_fft_x_conv(PmPmeLaunch a, PmFftPlan p, int TL) {
  pme_priority(a.prio);
  typedef typename Cx<T>::t T2;
  extern __shared__ unsigned char s_raw[];
  __shared__ double s_red[16][8];
  T2* b0 = reinterpret_cast<T2*>(s_raw);
  T2* b1 = b0 + TL * p.n;
  T2* tws = b1 + TL * p.n;
  load_twiddles(tws, p, a.tw);
  const int n = p.n;  // nx
  const int nl = a.ny * a.nzc, line0 = blockIdx.x * TL;
  __shared__ int s_y[64], s_kz[64];   // each line's y (-1: past the grid) and kz
  if (threadIdx.x < TL) {
    const int line = line0 + threadIdx.x;
    const int y = line / a.nzc;
    s_y[threadIdx.x] = line < nl ? y : -1;
    s_kz[threadIdx.x] = line - y * a.nzc;
  }
  __syncthreads();
  const long long xs = (long long)a.ny * a.nzc;   // grid stride of x
  T2* gc = reinterpret_cast<T2*>(a.grid_c);
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
    const int x = e / TL, l = e - x * TL;
    const int y = s_y[l];
    T2 v; v.x = 0; v.y = 0;
    if (y >= 0) v = gc[x * xs + (long long)y * a.nzc + s_kz[l]];
    b0[l * n + x] = v;
  }
  __syncthreads();
  T2* r = fft_lines<T, T2>(b0, b1, TL, p, tws, -1, n);
  if (a.want_energy || a.want_virial || !a.theta) {
    PmERed s = {0, 0, 0, 0, 0, 0, 0};
    for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
      const int l = e / n, kx = e - l * n;
      const int y = s_y[l];
      if (y < 0) continue;
      r[e] = influence<T, T2>(a, kx, y, s_kz[l], r[e], s);
    }
    store_block_energy(a, s, s_red);   // its barrier also orders the writes to r[]
  } else {
    // no energy records: theta(m) from the table
    const T* th = reinterpret_cast<const T*>(a.theta);
    for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
      const int l = e / n, kx = e - l * n;
      const int y = s_y[l];
      if (y < 0) continue;
      const T t = th[kx * xs + (long long)y * a.nzc + s_kz[l]];
      r[e].x *= t;
      r[e].y *= t;
    }
    __syncthreads();
  }
  T2* r2 = fft_lines<T, T2>(r, r == b0 ? b1 : b0, TL, p, tws, +1, n);
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
    const int x = e / TL, l = e - x * TL;
    const int y = s_y[l];
    if (y >= 0) gc[x * xs + (long long)y * a.nzc + s_kz[l]] = r2[l * n + x];
  }
}

// One kz plane of the plane-major grid per block: y forward, x forward, the
// influence function, x inverse and y inverse with the whole plane in LDS
// (row stride ny + 1, so that lines along y and lines along x are both read
// by consecutive threads on consecutive lines without bank conflicts).
template <typename T>
__global__ __launch_bounds__(1024) void k_fft_plane(PmPmeLaunch a, PmFftPlan px, PmFftPlan py) {
  pme_priority(a.prio);
  typedef typename Cx<T>::t T2;
  extern __shared__ unsigned char s_raw[];
  __shared__ double s_red[16][8];
  const int nx = a.nx, ny = a.ny, rs = ny + 1, kz = blockIdx.x;
  T2* b0 = reinterpret_cast<T2*>(s_raw);
  T2* b1 = b0 + nx * rs;
  T2* twx = b1 + nx * rs;
  T2* twy = twx + nx;
  load_twiddles(twx, px, a.tw);
  load_twiddles(twy, py, a.tw);
  T2* gp = reinterpret_cast<T2*>(a.grid_c) + (long long)kz * nx * ny;
  for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
    const int x = e / ny, y = e - x * ny;
    b0[x * rs + y] = gp[e];
  }
  __syncthreads();
  T2* r = fft_lines<T, T2, true>(b0, b1, nx, py, twy, -1, rs, 1);      // y lines
  r = fft_lines<T, T2, true>(r, r == b0 ? b1 : b0, ny, px, twx, -1, 1, rs);   // x lines
  PmERed s = {0, 0, 0, 0, 0, 0, 0};
  for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
    const int x = e / ny, y = e - x * ny;
    r[x * rs + y] = influence<T, T2>(a, x, y, kz, r[x * rs + y], s);
  }
  store_block_energy(a, s, s_red);
  r = fft_lines<T, T2, true>(r, r == b0 ? b1 : b0, ny, px, twx, +1, 1, rs);
  r = fft_lines<T, T2, true>(r, r == b0 ? b1 : b0, nx, py, twy, +1, rs, 1);
  for (int e = threadIdx.x; e < nx * ny; e += blockDim.x) {
    const int x = e / ny, y = e - x * ny;
    gp[e] = r[x * rs + y];
  }
}

// z inverse: complex half lines -> full Hermitian lines -> real grid
template <typename T>
__global__ __launch_bounds__(256) void k_fft_z_inv(PmPmeLaunch a, PmFftPlan p, int TL) {
  pme_priority(a.prio);
  typedef typename Cx<T>::t T2;
  extern __shared__ unsigned char s_raw[];
  const int n = p.n, ls = n + 1, nl_tot = a.nx * a.ny, nzc = a.nzc;
  T2* b0 = reinterpret_cast<T2*>(s_raw);
  T2* b1 = b0 + TL * ls;
  T2* tws = b1 + TL * ls;
  load_twiddles(tws, p, a.tw);
  const int line0 = blockIdx.x * TL;
  const T2* gc = reinterpret_cast<const T2*>(a.grid_c);
  if (a.plane) {
    for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
      const int k = e / TL, l = e - k * TL;
      const int line = line0 + l;
      T2 v; v.x = 0; v.y = 0;
      if (line < nl_tot) {
        if (k < nzc) v = gc[(long long)k * nl_tot + line];
        else { v = gc[(long long)(n - k) * nl_tot + line]; v.y = -v.y; }
      }
      b0[l * ls + k] = v;
    }
  } else {
    for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
      const int l = e / n, k = e - l * n;
      const int line = line0 + l;
      T2 v; v.x = 0; v.y = 0;
      if (line < nl_tot) {
        if (k < nzc) v = gc[(long long)line * nzc + k];
        else { v = gc[(long long)line * nzc + (n - k)]; v.y = -v.y; }
      }
      b0[l * ls + k] = v;
    }
  }
  __syncthreads();
  T2* r = fft_lines<T, T2>(b0, b1, TL, p, tws, +1, ls);
  T* g = reinterpret_cast<T*>(a.grid_r);
  for (int e = threadIdx.x; e < TL * n; e += blockDim.x) {
    const int l = e / n, k = e - l * n;
    const int line = line0 + l;
    if (line < nl_tot) g[(long long)line * n + k] = r[l * ls + k].x;
  }
}

}  // namespace

// ----------------------------------------------------------------- launcher
static int lines_per_block(int n, size_t elem) {
  int tl = (int)(32768 / ((size_t)(n + 1) * elem * 2));
  if (tl > 16) tl = 16;
  if (tl < 1) tl = 1;
  return tl;
}

// LDS of the z passes (padded lines) and of the separate y and x passes
static size_t zpass_lds(int n, int tl, size_t elem) { return ((size_t)2 * tl * (n + 1) + n) * elem; }
static size_t line_lds(int n, int tl, size_t elem) { return ((size_t)2 * tl + 1) * n * elem; }
// dynamic LDS of the plane kernel: two padded planes and both twiddle tables
static size_t plane_lds(int nx, int ny, size_t elem) { return ((size_t)2 * nx * (ny + 1) + nx + ny) * elem; }

static size_t fft_elem(int precision) { return precision == 1 ? sizeof(double2) : sizeof(float2); }

int pm_pme_plane_fits(int nx, int ny, int precision, int lds_max) {
  const size_t stat = sizeof(double) * 16 * 8;   // s_red of k_fft_plane
  return plane_lds(nx, ny, fft_elem(precision)) + stat <= (size_t)lds_max ? 1 : 0;
}

// threads of the plane kernel: the block size in {1024, 768, 512} that
// leaves the fewest idle threads over the butterfly loops of the four
// transforms
static int plane_threads(const PmFftPlan& px, const PmFftPlan& py) {
  static const int cand[3] = {1024, 768, 512};
  int best = 1024;
  long long best_cost = -1;
  for (int c : cand) {
    long long cost = 0;
    for (int t = 0; t < 2; t++) {
      const PmFftPlan& p = t == 0 ? px : py;
      const int nl = t == 0 ? py.n : px.n;
      for (int st = 0; st < p.nst; st++) {
        const long long work = (long long)nl * (p.n / p.rad[st]);
        cost += (work + c - 1) / c * c;
      }
    }
    if (best_cost < 0 || cost < best_cost) { best_cost = cost; best = c; }
  }
  return best;
}

// z forward, then either y forward, x forward + influence function + x
// inverse, y inverse as separate passes over the line-major grid or all of
// them in one kernel per kz plane, then z inverse: the fixed-point charge
// grid to the real potential grid
template <typename T>
static void launch_fft_chain(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans, bool z_done = false) {
  const size_t el = sizeof(typename Cx<T>::t);
  const int nthreads = 256;
  if (!z_done) {
    const int tl = lines_per_block(a.nz, el);
    const int nb = (a.nx * a.ny + tl - 1) / tl;
    hipLaunchKernelGGL(k_fft_z_fwd<T>, dim3(nb), dim3(nthreads), zpass_lds(a.nz, tl, el), s, a, plans[2], tl);
  }
  if (a.plane) {
    hipLaunchKernelGGL(k_fft_plane<T>, dim3(a.nzc), dim3(plane_threads(plans[0], plans[1])),
                       plane_lds(a.nx, a.ny, el), s, a, plans[0], plans[1]);
  } else {
    {
      const int tl = lines_per_block(a.ny, el);
      const int nb = (a.nx * a.nzc + tl - 1) / tl;
      hipLaunchKernelGGL(k_fft_y<T>, dim3(nb), dim3(nthreads), line_lds(a.ny, tl, el), s, a, plans[1], tl, -1);
    }
    {
      const int tl = lines_per_block(a.nx, el);
      const int nb = (a.ny * a.nzc + tl - 1) / tl;
      hipLaunchKernelGGL(k_fft_x_conv<T>, dim3(nb), dim3(nthreads), line_lds(a.nx, tl, el), s, a, plans[0], tl);
    }
    {
      const int tl = lines_per_block(a.ny, el);
      const int nb = (a.nx * a.nzc + tl - 1) / tl;
      hipLaunchKernelGGL(k_fft_y<T>, dim3(nb), dim3(nthreads), line_lds(a.ny, tl, el), s, a, plans[1], tl, +1);
    }
  }
  {
    const int tl = lines_per_block(a.nz, el);
    const int nb = (a.nx * a.ny + tl - 1) / tl;
    hipLaunchKernelGGL(k_fft_z_inv<T>, dim3(nb), dim3(nthreads), zpass_lds(a.nz, tl, el), s, a, plans[2], tl);
  }
}

template <typename T, int P>
static void launch_spread_tile(hipStream_t s, const PmPmeLaunch& a, const longlong4* pos64) {
  const int nbk = a.nbrick[0] * a.nbrick[1] * a.nbrick[2];
  const size_t lds = (size_t)a.tile[0] * a.tile[1] * a.tile[2] * sizeof(unsigned long long);
  hipLaunchKernelGGL((k_spread_tile<T, P>), dim3(nbk), dim3(256), lds, s, a, pos64);
}

template <typename T, int P>
static void launch_pme_t(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans,
                         const longlong4* pos64, const PmAtomTail* tail) {
  if (a.spread_zfwd) {
    const size_t el = sizeof(typename Cx<T>::t);
    const int ly = a.spread_zfwd;
    const int nb = ((a.nx + PM_SZ_LX - 1) / PM_SZ_LX) * ((a.ny + ly - 1) / ly);
    hipLaunchKernelGGL((k_spread_zfwd<T, P>), dim3(nb), dim3(256), zpass_lds(a.nz, PM_SZ_LX * ly, el), s, a,
                       plans[2], pos64);
    launch_fft_chain<T>(s, a, plans, true);
  } else {
    launch_spread_tile<T, P>(s, a, pos64);
    launch_fft_chain<T>(s, a, plans);
  }
  const dim3 g((a.nslot_max + 255) / 256), b(256);
  if (tail) hipLaunchKernelGGL((k_gather<T, P, true>), g, b, 0, s, a, pos64, *tail);
  else hipLaunchKernelGGL((k_gather<T, P, false>), g, b, 0, s, a, pos64, PmAtomTail{});
}

// adjoint: mode 0 the reciprocal potential at every atom (dE/dq), mode 1 the
// reciprocal H v (the gather of grad grad W against phi, then the spread of
// the directional charge derivative, its potential and the gather of grad W)
template <typename T, int P>
static void launch_pme_adj_t(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans, const longlong4* pos64,
                             int mode, const double4* slot_v, double* out) {
  const dim3 g((a.nslot_max + 255) / 256), b(256);
  launch_spread_tile<T, P>(s, a, pos64);
  launch_fft_chain<T>(s, a, plans);
  if (mode == 0) {
    hipLaunchKernelGGL((k_gather_adj<T, P, 0>), g, b, 0, s, a, pos64, slot_v, out);
    return;
  }
  hipLaunchKernelGGL((k_gather_adj<T, P, 1>), g, b, 0, s, a, pos64, slot_v, out);
  hipLaunchKernelGGL((k_spread_dir<T, P>), g, b, 0, s, a, pos64, slot_v);
  launch_fft_chain<T>(s, a, plans);
  hipLaunchKernelGGL((k_gather_adj<T, P, 2>), g, b, 0, s, a, pos64, slot_v, out);
}

int pm_pme_spread_zfwd_fits(const PmPmeLaunch& a, int precision) {
  for (int ly = 4; ly >= 1; ly /= 2) {
    const size_t lds = zpass_lds(a.nz, PM_SZ_LX * ly, fft_elem(precision));
    if (lds + sizeof(int) * (2 * PM_SZ_MAXCOL + 2) > 64 * 1024) continue;
    const int g[2] = {PM_SZ_LX, ly};
    long long cols = 1;
    for (int d = 0; d < 2; d++) {
      const double w = g[d] + a.order - 1 + 2.0 * a.drift_u[d];
      // the kernel's range spans at most floor(w / du) + 2 cells; one more for rounding
      long long c = (long long)floor(w / a.cell_du[d]) + 3;
      if (c > a.ncell[d]) c = a.ncell[d];
      cols *= c;
    }
    return cols <= PM_SZ_MAXCOL ? ly : 0;
  }
  return 0;
}

int pm_pme_xpass_blocks(int nx, int ny, int nz, int precision, int plane) {
  const int nzc = nz / 2 + 1;
  if (plane) return nzc;
  const int tl = lines_per_block(nx, fft_elem(precision));
  return (ny * nzc + tl - 1) / tl;
}

template <typename T>
static bool launch_pme_p(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans, const longlong4* pos64,
                         const PmAtomTail* tail) {
  switch (a.order) {
    case 2: launch_pme_t<T, 2>(s, a, plans, pos64, tail); return true;
    case 4: launch_pme_t<T, 4>(s, a, plans, pos64, tail); return true;
    case 6: launch_pme_t<T, 6>(s, a, plans, pos64, tail); return true;
    case 8: launch_pme_t<T, 8>(s, a, plans, pos64, tail); return true;
    case 10: launch_pme_t<T, 10>(s, a, plans, pos64, tail); return true;
    case 12: launch_pme_t<T, 12>(s, a, plans, pos64, tail); return true;
    default: return false;
  }
}

template <typename T>
static bool launch_pme_adj_p(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans, const longlong4* pos64,
                             int mode, const double4* sv, double* out) {
  switch (a.order) {
    case 2: launch_pme_adj_t<T, 2>(s, a, plans, pos64, mode, sv, out); return true;
    case 4: launch_pme_adj_t<T, 4>(s, a, plans, pos64, mode, sv, out); return true;
    case 6: launch_pme_adj_t<T, 6>(s, a, plans, pos64, mode, sv, out); return true;
    case 8: launch_pme_adj_t<T, 8>(s, a, plans, pos64, mode, sv, out); return true;
    case 10: launch_pme_adj_t<T, 10>(s, a, plans, pos64, mode, sv, out); return true;
    case 12: launch_pme_adj_t<T, 12>(s, a, plans, pos64, mode, sv, out); return true;
    default: return false;
  }
}

bool pm_launch_pme_adjoint(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans, const longlong4* pos64,
                           int precision, int mode, const double4* slot_v, double* out) {
  if (precision == 1) return launch_pme_adj_p<double>(s, a, plans, pos64, mode, slot_v, out);
  return launch_pme_adj_p<float>(s, a, plans, pos64, mode, slot_v, out);
}

void pm_launch_pme_theta(hipStream_t s, const PmPmeLaunch& a, int precision) {
  const long long ng = (long long)a.nx * a.ny * a.nzc;
  const dim3 g((unsigned)((ng + 255) / 256)), b(256);
  if (precision == 1) hipLaunchKernelGGL(k_theta<double>, g, b, 0, s, a);
  else hipLaunchKernelGGL(k_theta<float>, g, b, 0, s, a);
}

bool pm_launch_pme_plans(hipStream_t s, const PmPmeLaunch& a, const PmFftPlan* plans, const longlong4* pos64,
                         int precision, const PmAtomTail* tail) {
  if (precision == 1) return launch_pme_p<double>(s, a, plans, pos64, tail);
  return launch_pme_p<float>(s, a, plans, pos64, tail);
}
