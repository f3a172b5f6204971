// pmg_band.hip -- banded direct solve on the coarsest level (mgsolv = 1).
//
// Reference: Vbuildband (pmgc/buildBd.cc:251-304) builds the LINPACK band
// form of the coarsest operator and factors it with Vdpbfa; every V-cycle
// then solves with Vdpbsl (pmgc/mlinpckd.cc:167-235, used at
// pmgc/mgcsd.cc:391-407), both on the host.  Here all three run on the GPU
// in FP64:
//
//  * k_band_build: one thread per band column writes the column as
//    Vbuildband1_7 / Vbuildband1_27 do, in the same order, so offsets that
//    coincide on narrow grids keep the reference's last write.
//  * k_band_factor: one workgroup, right-looking band Cholesky.  The
//    (m+1) x (m+1) trailing window of the Schur complement stays in LDS
//    (circular in both indices); each step takes the pivot row of R, applies
//    the rank-1 update and brings in the next column of A.  R is written in
//    the reference's band layout (abd(m+1+i-j, j) = R(i,j)) and twice more
//    for the substitution, by rows and by columns, each row k indexed by
//    position modulo W: rt[k*W + p%W] = R(k, p), ct[k*W + p%W] = R(p, k),
//    1/R(k,k) in slot k%W, zero outside the band.
//  * k_band_solve: R^T y = b, then R x = y, in one wavefront.  Positions of
//    the sliding window live in registers (position p in lane p % 64, slot
//    (p / 64) % S); each step broadcasts the new unknown with readlane and
//    updates the window with one FMA per slot.  The coefficients of a block
//    of D steps, and the right-hand-side values entering the window, are
//    loaded one block ahead, so the per-step chain is FMA -> readlane ->
//    FMA.  The rest of the workgroup packs the right-hand side before and
//    writes the level (interior values, zero boundary) after.
//
// Layouts: abd is column-major (lda = m + 1) over the interior unknowns in
// lexicographic order (i fastest), as the reference's acB.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "pmg_internal.h"

namespace pmg {

namespace {

void band_check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    fprintf(stderr, "pmg: launch of %s failed: %s\n", what,
            hipGetErrorString(e));
}

constexpr int BAND_WMAX = 140;  // largest LDS window (m + 1) of the factor
constexpr int FAC_NT = 1024;
constexpr int SOLVE_NT = 256;
constexpr int SOLVE_D = 32;  // steps per prefetch block

__device__ inline double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Vbuildband1_7 (buildBd.cc:62-124) / Vbuildband1_27 (:126-249).  Column jj
// (0-based) holds the couplings of interior point jj with the points ii <=
// jj at row m - (jj - ii).
__global__ void k_band_build(int nx, int ny, int nz, int numdia,
                             const double *__restrict__ ac,
                             double *__restrict__ abd, int m, int lda) {
  const int sx = nx - 2, sy = ny - 2, sxy = sx * sy;
  const int nb = sxy * (nz - 2);
  const int jj = blockIdx.x * blockDim.x + threadIdx.x;
  if (jj >= nb) return;
  const int i = jj % sx + 2, j = (jj / sx) % sy + 2, k = jj / sxy + 2;
  const size_t N = (size_t)nx * ny * nz;
  auto at = [&](int pl, int a, int b, int c) {
    return ac[(size_t)pl * N + (size_t)(c - 1) * nx * ny +
              (size_t)(b - 1) * nx + (a - 1)];
  };
  double *col = abd + (size_t)jj * lda;
  for (int r = 0; r < lda; r++) col[r] = 0.0;
  col[m] = at(0, i, j, k);                  // diagonal: oC only
  col[m - 1] = -at(1, i - 1, j, k);         // east
  col[m - sx] = -at(2, i, j - 1, k);        // north
  if (numdia == 27) {
    col[m - (sx - 1)] = -at(4, i, j - 1, k);  // north-east: oNE(i, j-1, k)
    col[m - (sx + 1)] = -at(5, i, j - 1, k);  // north-west: oNW(i, j-1, k)
  }
  col[m - sxy] = -at(3, i, j, k - 1);       // up
  if (numdia == 27) {
    col[m - (sxy - 1)] = -at(6, i, j, k - 1);        // uE
    col[m - (sxy + 1)] = -at(7, i, j, k - 1);        // uW
    col[m - (sxy - sx)] = -at(8, i, j, k - 1);       // uN
    col[m - (sxy + sx)] = -at(9, i, j, k - 1);       // uS
    col[m - (sxy - sx - 1)] = -at(10, i, j, k - 1);  // uNE
    col[m - (sxy - sx + 1)] = -at(11, i, j, k - 1);  // uNW
    col[m - (sxy + sx - 1)] = -at(12, i, j, k - 1);  // uSE
    col[m - (sxy + sx + 1)] = -at(13, i, j, k - 1);  // uSW
  }
}

__global__ void k_band_zero(double *__restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// Right-looking band Cholesky (same factor as Vdpbfa, mlinpckd.cc:167-206).
// info_out[0] = first column (1-based) whose pivot is <= 0, else 0.
__global__ __launch_bounds__(FAC_NT) void k_band_factor(
    double *__restrict__ abd, double *__restrict__ rt, double *__restrict__ ct,
    int n, int m, int lda, int W, double *__restrict__ info_out) {
  __shared__ double win[BAND_WMAX * BAND_WMAX];
  __shared__ double rrow[BAND_WMAX];
  const int w = m + 1;
  const int tid = threadIdx.x;
  // window over rows/columns 0..w-1 (slot = index mod w)
  for (int e = tid; e < w * w; e += FAC_NT) {
    const int a = e / w, b = e - (e / w) * w;
    const int lo = min(a, b), hi = max(a, b);
    win[e] = hi < n ? abd[(size_t)hi * lda + m + lo - hi] : 0.0;
  }
  // (a, b) pairs of the m x m update handled by this thread
  const int mm = m * m;
  const int step_a = FAC_NT / max(m, 1), step_b = FAC_NT % max(m, 1);
  __syncthreads();
  int info = 0;
  for (int j = 0; j < n; j++) {
    const int js = j % w;
    const double d = win[js * w + js];
    if (d <= 0.0) {  // the test of Vdpbfa (uniform: one LDS value)
      info = j + 1;
      break;
    }
    const double r = sqrt(d);
    if (tid < m) {
      const int b = j + 1 + tid;
      double v = 0.0;
      if (b < n) {
        int sb = js + 1 + tid;
        if (sb >= w) sb -= w;
        v = win[js * w + sb] / r;
        abd[(size_t)b * lda + m - (b - j)] = v;
        rt[(size_t)j * W + b % W] = v;
        ct[(size_t)b * W + j % W] = v;
      }
      rrow[tid] = v;
    }
    if (tid == FAC_NT - 1) {
      abd[(size_t)j * lda + m] = r;
      rt[(size_t)j * W + j % W] = 1.0 / r;
      ct[(size_t)j * W + j % W] = 1.0 / r;
    }
    __syncthreads();
    // rank-1 update of rows / columns j+1 .. j+m (every slot but js)
    const int j1 = js + 1 == w ? 0 : js + 1;
    if (m > 0) {
      int a = tid / m, b = tid - (tid / m) * m;
      for (int e = tid; e < mm; e += FAC_NT) {
        int sa = j1 + a, sb = j1 + b;
        if (sa >= w) sa -= w;
        if (sb >= w) sb -= w;
        win[sa * w + sb] -= rrow[a] * rrow[b];
        a += step_a;
        b += step_b;
        if (b >= m) {
          b -= m;
          a++;
        }
      }
    }
    // column c = j + w of A enters slot js (rows c-m .. c; zero past n)
    const int c = j + w;
    for (int a = tid; a <= m; a += FAC_NT) {
      int sa = j1 + a;
      if (sa >= w) sa -= w;
      const double v = c < n ? abd[(size_t)c * lda + a] : 0.0;
      win[sa * w + js] = v;
      win[js * w + sa] = v;
    }
    __syncthreads();
  }
  if (tid == 0) info_out[0] = (double)info;
}

// ---- substitution ----
// Position p of the window: lane p % 64, slot (p / 64) % S; W = 64 S > m.

template <int S>
__device__ inline double pick(const double (&v)[S], int s) {
  double r = v[0];
#pragma unroll
  for (int q = 1; q < S; q++)
    if (s == q) r = v[q];
  return r;
}

// Forward: y = R^{-T} b (Vdpbsl's first loop, mlinpckd.cc:216-222).
// Loads for the block starting at k1: rt rows and the values that enter the
// window when this lane's position retires inside the block.
template <int S>
__device__ inline void fwd_load(int k1, int n, int lane,
                                const double *__restrict__ rt,
                                const double *y, double (&c)[SOLVE_D][S],
                                double &nv) {
  constexpr int W = 64 * S;
#pragma unroll
  for (int u = 0; u < SOLVE_D; u++) {
    const double *row = rt + (size_t)min(k1 + u, n - 1) * W + lane;
#pragma unroll
    for (int s = 0; s < S; s++) c[u][s] = row[s * 64];
  }
  const int kr = k1 + ((lane - k1) & 63);  // this lane's retirement step
  const int src = kr + W;
  const double v = y[min(src, n - 1)];
  nv = (kr < k1 + SOLVE_D && src < n) ? v : 0.0;
}

template <int S>
__device__ inline void fwd_block(int k0, int n, int lane, double (&b)[S],
                                 double *y, const double (&c)[SOLVE_D][S],
                                 double nv) {
#pragma unroll
  for (int u = 0; u < SOLVE_D; u++) {
    const int k = k0 + u;
    if (k >= n) break;
    const int ps = (k >> 6) % S, pl = k & 63;
    const double yk = readlane_f64(pick<S>(b, ps) * pick<S>(c[u], ps), pl);
#pragma unroll
    for (int s = 0; s < S; s++) b[s] = fma(-c[u][s], yk, b[s]);
    if (lane == pl) y[k] = yk;
#pragma unroll
    for (int s = 0; s < S; s++)
      if (s == ps && lane == pl) b[s] = nv;
  }
}

// Backward: x = R^{-1} y (Vdpbsl's second loop, mlinpckd.cc:225-234).
template <int S>
__device__ inline void bwd_load(int k1, int lane, const double *__restrict__ ct,
                                const double *y, double (&c)[SOLVE_D][S],
                                double &nv) {
  constexpr int W = 64 * S;
#pragma unroll
  for (int u = 0; u < SOLVE_D; u++) {
    const double *col = ct + (size_t)max(k1 - u, 0) * W + lane;
#pragma unroll
    for (int s = 0; s < S; s++) c[u][s] = col[s * 64];
  }
  const int kr = k1 - ((k1 - lane) & 63);
  const int src = kr - W;
  const double v = y[max(src, 0)];
  nv = (kr > k1 - SOLVE_D && src >= 0) ? v : 0.0;
}

template <int S>
__device__ inline void bwd_block(int k0, int lane, double (&b)[S], double *y,
                                 const double (&c)[SOLVE_D][S], double nv) {
#pragma unroll
  for (int u = 0; u < SOLVE_D; u++) {
    const int k = k0 - u;
    if (k < 0) break;
    const int ps = (k >> 6) % S, pl = k & 63;
    const double xk = readlane_f64(pick<S>(b, ps) * pick<S>(c[u], ps), pl);
#pragma unroll
    for (int s = 0; s < S; s++) b[s] = fma(-c[u][s], xk, b[s]);
    if (lane == pl) y[k] = xk;
#pragma unroll
    for (int s = 0; s < S; s++)
      if (s == ps && lane == pl) b[s] = nv;
  }
}

template <class T, int S>
__global__ __launch_bounds__(SOLVE_NT) void k_band_solve(
    int nx, int ny, int nz, int n, const double *__restrict__ rt,
    const double *__restrict__ ct, double *__restrict__ y,
    const T *__restrict__ f, T *__restrict__ x) {
  constexpr int W = 64 * S;
  const int tid = threadIdx.x;
  const int sx = nx - 2, sy = ny - 2, sxy = sx * sy;
  // Vxcopy_small (mikpckd.cc:78-92): interior of f, packed
  for (int p = tid; p < n; p += SOLVE_NT) {
    const int i = p % sx + 1, j = (p / sx) % sy + 1, k = p / sxy + 1;
    y[p] = (double)f[(size_t)k * nx * ny + (size_t)j * nx + i];
  }
  __syncthreads();
  if (tid < 64) {
    const int lane = tid;
    double b[S];
    double cA[SOLVE_D][S], cB[SOLVE_D][S], nvA, nvB;
    // forward, window = positions 0 .. W-1
#pragma unroll
    for (int s = 0; s < S; s++) {
      const int p = s * 64 + lane;
      const double v = y[min(p, n - 1)];
      b[s] = p < n ? v : 0.0;
    }
    fwd_load<S>(0, n, lane, rt, y, cA, nvA);
    for (int k0 = 0; k0 < n; k0 += 2 * SOLVE_D) {
      fwd_load<S>(k0 + SOLVE_D, n, lane, rt, y, cB, nvB);
      fwd_block<S>(k0, n, lane, b, y, cA, nvA);
      if (k0 + SOLVE_D >= n) break;
      fwd_load<S>(k0 + 2 * SOLVE_D, n, lane, rt, y, cA, nvA);
      fwd_block<S>(k0 + SOLVE_D, n, lane, b, y, cB, nvB);
    }
    // the stores of y above precede the loads below
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // backward, window = positions n-W .. n-1
#pragma unroll
    for (int s = 0; s < S; s++) {
      int o = (n - 1) % W - s * 64 - lane;
      if (o < 0) o += W;
      const int p = n - 1 - o;
      const double v = y[max(p, 0)];
      b[s] = p >= 0 ? v : 0.0;
    }
    bwd_load<S>(n - 1, lane, ct, y, cA, nvA);
    for (int k0 = n - 1; k0 >= 0; k0 -= 2 * SOLVE_D) {
      bwd_load<S>(k0 - SOLVE_D, lane, ct, y, cB, nvB);
      bwd_block<S>(k0, lane, b, y, cA, nvA);
      if (k0 - SOLVE_D < 0) break;
      bwd_load<S>(k0 - 2 * SOLVE_D, lane, ct, y, cA, nvA);
      bwd_block<S>(k0 - SOLVE_D, lane, b, y, cB, nvB);
    }
  }
  __syncthreads();
  // Vxcopy_large + VfboundPMG00: interior from y, boundary zero
  const int nall = nx * ny * nz;
  for (int q = tid; q < nall; q += SOLVE_NT) {
    const int i = q % nx, j = (q / nx) % ny, k = q / (nx * ny);
    const bool in = i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 &&
                    k < nz - 1;
    const int p = in ? (k - 1) * sxy + (j - 1) * sx + (i - 1) : 0;
    const double v = y[p];
    x[q] = in ? (T)v : (T)0;
  }
}

}  // namespace

BandDims band_dims(const Dim &d, int numdia) {
  BandDims b;
  const int sx = d.nx - 2, sy = d.ny - 2;
  b.n = sx * sy * (d.nz - 2);
  b.m = numdia == 7 ? sx * sy : sx * sy + sx + 1;
  b.lda = b.m + 1;
  b.W = 64 * ((b.m + 1 + 63) / 64);
  // one interior point in x or y puts offsets of buildBd.cc:192-245 on the
  // diagonal or outside the column; the factor window must fit in LDS
  b.ok = d.nx >= 3 && d.ny >= 3 && d.nz >= 3 && b.n >= 1 &&
         (numdia == 7 || (sx >= 2 && sy >= 2)) && b.m + 1 <= BAND_WMAX;
  return b;
}

void band_build_factor(const Dim &d, int numdia, const double *ac, double *abd,
                       double *rt, double *ct, double *info, hipStream_t s) {
  const BandDims b = band_dims(d, numdia);
  if (!b.ok) {
    fprintf(stderr, "pmg: coarsest level unsuitable for the band solve\n");
    return;
  }
  const int nt = 256;
  const size_t nw = (size_t)b.n * b.W;
  k_band_zero<<<(unsigned)((nw + nt - 1) / nt), nt, 0, s>>>(rt, nw);
  k_band_zero<<<(unsigned)((nw + nt - 1) / nt), nt, 0, s>>>(ct, nw);
  k_band_build<<<(b.n + nt - 1) / nt, nt, 0, s>>>(d.nx, d.ny, d.nz, numdia,
                                                  ac, abd, b.m, b.lda);
  band_check_launch("band_build");
  k_band_factor<<<1, FAC_NT, 0, s>>>(abd, rt, ct, b.n, b.m, b.lda, b.W, info);
  band_check_launch("band_factor");
}

template <class T>
void band_solve(const Dim &d, int numdia, const double *rt, const double *ct,
                double *y, const T *f, T *x, hipStream_t s) {
  const BandDims b = band_dims(d, numdia);
  if (!b.ok) {
    fprintf(stderr, "pmg: coarsest level unsuitable for the band solve\n");
    return;
  }
  switch (b.W / 64) {
    case 1:
      k_band_solve<T, 1><<<1, SOLVE_NT, 0, s>>>(d.nx, d.ny, d.nz, b.n, rt, ct,
                                                y, f, x);
      break;
    case 2:
      k_band_solve<T, 2><<<1, SOLVE_NT, 0, s>>>(d.nx, d.ny, d.nz, b.n, rt, ct,
                                                y, f, x);
      break;
    default:
      k_band_solve<T, 3><<<1, SOLVE_NT, 0, s>>>(d.nx, d.ny, d.nz, b.n, rt, ct,
                                                y, f, x);
      break;
  }
  band_check_launch("band_solve");
}

template void band_solve<double>(const Dim &, int, const double *,
                                 const double *, double *, const double *,
                                 double *, hipStream_t);
template void band_solve<float>(const Dim &, int, const double *,
                                const double *, double *, const float *,
                                float *, hipStream_t);

}  // namespace pmg
