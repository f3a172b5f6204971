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
//    updates the window with one FMA per slot, so the per-step chain is
//    FMA -> readlane -> FMA.  The packed vector sits in LDS; coefficient
//    rows are loaded a fixed number of steps ahead.  The rest of the
//    workgroup packs the right-hand side before and writes the level
//    (interior values, zero boundary) after.
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

constexpr int BAND_WMAX = 128;  // largest window (m + 1) of the factor
constexpr int FAC_NT = 1024;
constexpr int SOLVE_NT = 256;
constexpr int BAND_NMAX = 8192;  // unknowns of the LDS-resident vector

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
//
// Window slot s holds absolute index j + ((s - j) mod w) at step j.  Thread
// t < G*w (G = FAC_NT / w row groups) owns column slot t % w and row slots
// g, g + G, ... (g = t / w) in registers.  Each step reads the pivot row
// published in LDS by its owners at the end of the step before, so one
// barrier per step suffices.  The columns of A that enter the window are
// staged in LDS a chunk of FAC_K steps ahead, and the rows of R are kept in
// LDS and written out a chunk at a time, so no step waits on global memory.
constexpr int FAC_K = 16;
constexpr int FAC_QMAX = (BAND_WMAX + FAC_NT / BAND_WMAX - 1) / (FAC_NT / BAND_WMAX);

__global__ __launch_bounds__(FAC_NT) void k_band_factor(
    double *__restrict__ abd, double *__restrict__ rt, double *__restrict__ ct,
    int n, int m, int lda, int W, double *__restrict__ info_out) {
  __shared__ double prow[2][BAND_WMAX];
  __shared__ double colbuf[2][FAC_K][BAND_WMAX];
  __shared__ double rbuf[2][FAC_K][BAND_WMAX];  // rows of R: r, then R(j, j+o)
  const int w = m + 1;
  const int tid = threadIdx.x;
  const int G = FAC_NT / w;
  const bool own = tid < G * w;
  const int cb = own ? tid % w : 0, g = own ? tid / w : 0;
  // column c of A (band entries 0..m, zero past n) into a chunk buffer
  auto col_entry = [&](int e, int c0) {
    const int kk = e / w, r = e - (e / w) * w;
    const int c = c0 + kk;
    return c < n ? abd[(size_t)c * lda + r] : 0.0;
  };
  // initial window: absolute index = slot
  double wv[FAC_QMAX];
#pragma unroll
  for (int q = 0; q < FAC_QMAX; q++) {
    const int ra = g + G * q;
    double v = 0.0;
    if (own && ra < w) {
      const int lo = min(ra, cb), hi = max(ra, cb);
      if (hi < n) v = abd[(size_t)hi * lda + m + lo - hi];
    }
    wv[q] = v;
  }
  for (int e = tid; e < FAC_K * w; e += FAC_NT)
    colbuf[0][e / w][e - (e / w) * w] = col_entry(e, w);
  if (own && g == 0) prow[0][cb] = wv[0];
  __syncthreads();
  int info = 0;
  const int wm = W - 1;  // W is a power of two
  // rows j0 .. j0+FAC_K-1 of R (those below jend) from rbuf to the band and
  // the row / column copies
  auto flush = [&](int j0, int jend) {
    const double *rb = &rbuf[(j0 / FAC_K) & 1][0][0];
    for (int e = tid; e < FAC_K * w; e += FAC_NT) {
      const int jj = j0 + e / w, o = e - (e / w) * w;
      if (jj >= jend) continue;
      const double v = rb[(e / w) * BAND_WMAX + o];
      const int bcol = jj + o;
      if (o == 0) {
        const double vi = 1.0 / v;
        abd[(size_t)jj * lda + m] = v;
        rt[(size_t)jj * W + (jj & wm)] = vi;
        ct[(size_t)jj * W + (jj & wm)] = vi;
      } else if (bcol < n) {
        abd[(size_t)bcol * lda + m - o] = v;
        rt[(size_t)jj * W + (bcol & wm)] = v;
        ct[(size_t)bcol * W + (jj & wm)] = v;
      }
    }
  };
  for (int j = 0; j < n; j++) {
    const int js = j % w, jk = j % FAC_K, par = j & 1;
    if (jk == 0) {  // the next chunk's columns, used from step j + FAC_K on
      double *nb = &colbuf[((j / FAC_K) + 1) & 1][0][0];
      for (int e = tid; e < FAC_K * w; e += FAC_NT)
        nb[(e / w) * BAND_WMAX + e - (e / w) * w] = col_entry(e, j + FAC_K + w);
      if (j >= FAC_K) flush(j - FAC_K, n);
    }
    const double d = prow[par][js];
    if (d <= 0.0) {  // the test of Vdpbfa (uniform: one LDS value)
      info = j + 1;
      break;
    }
    const double r = sqrt(d), rinv = 1.0 / r;
    if (own) {
      int relc = cb - js;
      if (relc < 0) relc += w;
      const double vc = prow[par][cb] * rinv;  // R(j, j + relc)
      if (g == 0) rbuf[(j / FAC_K) & 1][jk][relc] = relc == 0 ? r : vc;
      // rank-1 update of every owned element (the row and the column of
      // the pivot slot are overwritten below; rows past w are unused)
      const double cf = vc * rinv;
#pragma unroll
      for (int q = 0; q < FAC_QMAX; q++) {
        const int ra = min(g + G * q, w - 1);
        wv[q] = fma(-prow[par][ra], cf, wv[q]);
      }
      // the pivot slot takes column c = j + w of A; in the new window
      // slot s holds j + 1 + ((s - js - 1) mod w)
      const double *cnew = colbuf[(j / FAC_K) & 1][jk];
      if (cb == js) {
#pragma unroll
        for (int q = 0; q < FAC_QMAX; q++) {
          int rn = g + G * q - js - 1;
          if (rn < 0) rn += w;
          wv[q] = cnew[min(rn, w - 1)];
        }
      }
      if (g == js % G) {  // this thread holds row js at q = js / G
        int reln = cb - js - 1;
        if (reln < 0) reln += w;
        const double v = cnew[reln];
        const int qj = js / G;
#pragma unroll
        for (int q = 0; q < FAC_QMAX; q++)
          if (q == qj) wv[q] = v;
      }
      const int js1 = js + 1 == w ? 0 : js + 1;
      if (g == js1 % G) {  // publish the next pivot row
        const int q1 = js1 / G;
        double v = wv[0];
#pragma unroll
        for (int q = 1; q < FAC_QMAX; q++)
          if (q == q1) v = wv[q];
        prow[par ^ 1][cb] = v;
      }
    }
    __syncthreads();
  }
  if (info == 0 && n > 0) flush(((n - 1) / FAC_K) * FAC_K, n);
  if (tid == 0) info_out[0] = (double)info;
}

// ---- substitution ----
// Position p of the window: lane p % 64, slot (p / 64) % S; W = 64 S > m.
// The unknowns are padded to npad (a multiple of W) with zero rows and a
// zero right-hand side, so the positions fall in whole groups of 64: in the
// group of positions 64g .. 64g+63 every lane retires its slot g % S once,
// position 64g + u in lane u.  A group is unrolled with the slot and the
// lane of each step known at compile time.  A retiring slot is zeroed and
// collects the updates of its next position (W further on); that
// position's right-hand side is added, and the group's results stored, at
// the end of the group.  The vector lives in LDS; coefficient rows are
// loaded SOLVE_RING steps ahead into a register ring.
constexpr int SOLVE_RING = 16;

template <int S, bool FWD, int PS>
__device__ __attribute__((always_inline)) void band_group(
    int g, int npad, int lane, double (&b)[S], double (&c)[SOLVE_RING][S],
    const double *__restrict__ co, double *ys) {
  constexpr int W = 64 * S, D = SOLVE_RING;
  double zs = 0.0;  // this lane's result in the group
  // row of step t + D (the padding rows around co are zero)
  const double *nrow =
      co + (ptrdiff_t)(FWD ? 64 * g + D : 64 * g + 63 - D) * W + lane;
#pragma unroll
  for (int u = 0; u < 64; u++) {
    const int pl = FWD ? u : 63 - u;  // lane of position k
    const double zk =
        readlane_f64(b[PS] * c[u % D][PS], pl);  // b_k / R(k,k)
#pragma unroll
    for (int s = 0; s < S; s++) b[s] = fma(-c[u % D][s], zk, b[s]);
    const bool me = lane == pl;
    b[PS] = me ? 0.0 : b[PS];
    zs = me ? zk : zs;
    const double *row = nrow + (ptrdiff_t)(FWD ? u : -u) * W;
#pragma unroll
    for (int s = 0; s < S; s++) c[u % D][s] = row[s * 64];
    // keep each step's loads in that step (the scheduler would otherwise
    // sink them next to their use, D steps later, and expose the latency)
    __builtin_amdgcn_sched_barrier(0);
  }
  ys[64 * g + lane] = zs;
  const int pe = FWD ? 64 * g + lane + W : 64 * g + lane - W;
  const double ve = ys[min(max(pe, 0), npad - 1)];
  b[PS] += (pe >= 0 && pe < npad) ? ve : 0.0;
}

// One sweep of the substitution (FWD: positions 0 .. npad-1 with the rows
// of R from rt, R^T y = b; else npad-1 .. 0 with its columns from ct,
// R x = y), in place in ys.  npad is a multiple of W, so the slots of the
// groups come in a fixed order.  co points at row 0 of a copy with
// SOLVE_RING zero rows on either side.
template <int S, bool FWD>
__device__ __attribute__((always_inline)) void band_sweep(
    int npad, int lane, const double *__restrict__ co, double *ys) {
  constexpr int W = 64 * S, D = SOLVE_RING;
  const int ng = npad / 64;
  double b[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    // FWD: positions 0 .. W-1; else npad-W .. npad-1 (slot s = (p/64) % S)
    const int p = FWD ? s * 64 + lane : npad - W + s * 64 + lane;
    b[s] = ys[p];
  }
  double c[D][S];
#pragma unroll
  for (int u = 0; u < D; u++) {
    const double *row = co + (ptrdiff_t)(FWD ? u : npad - 1 - u) * W + lane;
#pragma unroll
    for (int s = 0; s < S; s++) c[u][s] = row[s * 64];
  }
  if constexpr (S == 1) {
    for (int q = 0; q < ng; q++)
      band_group<1, FWD, 0>(FWD ? q : ng - 1 - q, npad, lane, b, c, co, ys);
  } else {
    for (int q = 0; q < ng; q += 2) {
      if (FWD) {
        band_group<2, FWD, 0>(q, npad, lane, b, c, co, ys);
        band_group<2, FWD, 1>(q + 1, npad, lane, b, c, co, ys);
      } else {
        band_group<2, FWD, 1>(ng - 1 - q, npad, lane, b, c, co, ys);
        band_group<2, FWD, 0>(ng - 2 - q, npad, lane, b, c, co, ys);
      }
    }
  }
}

template <class T, int S>
__global__ __launch_bounds__(SOLVE_NT) void k_band_solve(
    int nx, int ny, int nz, int n, int npad, const double *__restrict__ rt,
    const double *__restrict__ ct, const T *__restrict__ f,
    T *__restrict__ x) {
  extern __shared__ double ys[];
  const int tid = threadIdx.x;
  const int sx = nx - 2, sy = ny - 2, sxy = sx * sy;
  // Vxcopy_small (mikpckd.cc:78-92): interior of f, packed; zero padding
  for (int p = tid; p < npad; p += SOLVE_NT) {
    double v = 0.0;
    if (p < n) {
      const int i = p % sx + 1, j = (p / sx) % sy + 1, k = p / sxy + 1;
      v = (double)f[(size_t)k * nx * ny + (size_t)j * nx + i];
    }
    ys[p] = v;
  }
  __syncthreads();
  if (tid < 64) {
    band_sweep<S, true>(npad, tid, rt, ys);   // R^T y = b
    band_sweep<S, false>(npad, tid, ct, ys);  // R x = y
  }
  __syncthreads();
  // Vxcopy_large + VfboundPMG00: interior from the packed vector, boundary 0
  const int nall = nx * ny * nz;
  for (int q = tid; q < nall; q += SOLVE_NT) {
    const int i = q % nx, j = (q / nx) % ny, k = q / (nx * ny);
    const bool in = i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 &&
                    k < nz - 1;
    const int p = in ? (k - 1) * sxy + (j - 1) * sx + (i - 1) : 0;
    const double v = ys[p];
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
  b.npad = b.W * ((b.n + b.W - 1) / b.W);
  b.nrow = b.npad + 2 * SOLVE_RING;
  // one interior point in x or y puts offsets of buildBd.cc:192-245 on the
  // diagonal or outside the column; the factor window and the substitution
  // vector must fit in LDS
  b.ok = d.nx >= 3 && d.ny >= 3 && d.nz >= 3 && b.n >= 1 &&
         (numdia == 7 || (sx >= 2 && sy >= 2)) && b.m + 1 <= BAND_WMAX &&
         b.npad <= BAND_NMAX;
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
  const size_t nw = (size_t)b.nrow * b.W;
  k_band_zero<<<(unsigned)((nw + nt - 1) / nt), nt, 0, s>>>(rt, nw);
  k_band_zero<<<(unsigned)((nw + nt - 1) / nt), nt, 0, s>>>(ct, nw);
  k_band_build<<<(b.n + nt - 1) / nt, nt, 0, s>>>(d.nx, d.ny, d.nz, numdia,
                                                  ac, abd, b.m, b.lda);
  band_check_launch("band_build");
  const size_t off = (size_t)SOLVE_RING * b.W;  // leading zero rows
  k_band_factor<<<1, FAC_NT, 0, s>>>(abd, rt + off, ct + off, b.n, b.m, b.lda,
                                     b.W, info);
  band_check_launch("band_factor");
}

template <class T>
void band_solve(const Dim &d, int numdia, const double *rt, const double *ct,
                const T *f, T *x, hipStream_t s) {
  const BandDims b = band_dims(d, numdia);
  if (!b.ok) {
    fprintf(stderr, "pmg: coarsest level unsuitable for the band solve\n");
    return;
  }
  const size_t lds = (size_t)b.npad * sizeof(double);
  const size_t off = (size_t)SOLVE_RING * b.W;  // leading zero rows
  if (b.W == 64)
    k_band_solve<T, 1><<<1, SOLVE_NT, lds, s>>>(d.nx, d.ny, d.nz, b.n, b.npad,
                                                rt + off, ct + off, f, x);
  else
    k_band_solve<T, 2><<<1, SOLVE_NT, lds, s>>>(d.nx, d.ny, d.nz, b.n, b.npad,
                                                rt + off, ct + off, f, x);
  band_check_launch("band_solve");
}

template void band_solve<double>(const Dim &, int, const double *,
                                 const double *, const double *, double *,
                                 hipStream_t);
template void band_solve<float>(const Dim &, int, const double *,
                                const double *, const float *, float *,
                                hipStream_t);

}  // namespace pmg
