// pmg_kernels.hip -- HIP kernels of the PMG hot path for gfx950 (CDNA4).
//
// Every kernel works on the reference's 1-based Fortran-order arrays with i
// fastest (pmgc/vmatrix.h:79-80).  Work is mapped with i on the 64-lane
// wavefront dimension so every stencil load is a coalesced row segment;
// the plain stencil kernels use blocks of 64 x 4 threads over (i, j) that
// march KCH planes in k (one plane on small levels), the smoothers their own
// tilings described at each kernel.  Reductions are deterministic:
// wavefront sums, then LDS in a fixed order, one partial per block, and a
// single-block final pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <utility>
#include <cmath>
#include <cstdio>

#include "pmg_internal.h"

namespace pmg {

constexpr int BX = 64, BY = 4, KCH = 4;

__host__ __device__ inline size_t ix(int nx, int ny, int i, int j, int k) {
  return ((size_t)(k - 1) * (size_t)ny + (size_t)(j - 1)) * (size_t)nx +
         (size_t)(i - 1);
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// XCD-aware tiling.  Workgroups are dispatched round-robin over the eight
// XCDs (workgroup b runs on XCD b % 8), each with its own L2.  Stencil
// kernels are launched as 1-D grids of xcd_blocks(g) workgroups and map
// workgroup b to tile (b % 8) * per + b / 8 of the logical g.x * g.y * g.z
// tile grid (x fastest): every XCD works on one contiguous slab of tiles, so
// the halo rows and planes that neighbouring tiles share are fetched into
// one L2 once.  Workgroups past the last tile do no work (reductions still
// store a zero partial).
struct XG {
  int gx, gy, gz;
};
inline XG xg_of(const dim3 &g) { return XG{(int)g.x, (int)g.y, (int)g.z}; }
inline unsigned xcd_blocks(const dim3 &g) {
  return 8u * ((g.x * g.y * g.z + 7u) / 8u);
}
// bid: the workgroup's index in the launch of one problem (blockIdx.x, or
// its index within its problem's share of a grouped launch).
__device__ inline bool xcd_tile_at(const XG &g, unsigned bid, int &bx,
                                   int &by, int &bz) {
  const int nt = g.gx * g.gy * g.gz, per = (nt + 7) / 8;
  const int t = (int)(bid % 8u) * per + (int)(bid / 8u);
  bx = t % g.gx;
  by = (t / g.gx) % g.gy;
  bz = t / (g.gx * g.gy);
  return t < nt;
}
__device__ inline bool xcd_tile(const XG &g, int &bx, int &by, int &bz) {
  return xcd_tile_at(g, blockIdx.x, bx, by, bz);
}

static void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    fprintf(stderr, "pmg: launch of %s failed: %s\n", what,
            hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Wavefront and block reductions (deterministic for a fixed block shape)
// ---------------------------------------------------------------------------
// One DPP lane move of a double (both 32-bit halves with the same control);
// lanes without a source, or in rows outside ROWMASK, read 0.
template <int CTRL, int ROWMASK>
__device__ inline double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL,
                                             ROWMASK, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL,
                                             ROWMASK, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront, returned in every lane.  Row
// prefix sums by row_shr 1, 2, 4, 8 leave each 16-lane row's total in its
// lane 15; row_bcast15 and row_bcast31 fold the four rows into lane 63,
// which is then read out.  All moves stay in the VALU (no LDS round trips
// as with lane shuffles), and the order of additions is fixed.
__device__ inline double wave_sum(double v) {
  v += dpp_f64<0x111, 0xF>(v);  // row_shr:1
  v += dpp_f64<0x112, 0xF>(v);  // row_shr:2
  v += dpp_f64<0x114, 0xF>(v);  // row_shr:4
  v += dpp_f64<0x118, 0xF>(v);  // row_shr:8
  v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

template <int NT>
__device__ inline double block_sum(double v, double *red) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  v = wave_sum(v);
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (wid == 0) {
    t = lane < NT / 64 ? red[lane] : 0.0;
    t = wave_sum(t);
  }
  __syncthreads();
  return t;  // valid in thread 0
}

// Reduction epilogue: each block stores its sums (valid in thread 0) as a
// partial; red_final() then sums the partials in block order in a second,
// single-workgroup kernel.  (A "last block finishes" scheme needs one
// device-scope atomic per block; across the eight XCDs, whose L2 caches are
// not coherent, those serialise and cost more than the extra launch.)
template <int NV>
__device__ inline void red_store_at(const Red &r, size_t b, double v0,
                                   double v1) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (tid == 0) {
    r.part[b * NV] = v0;
    if (NV == 2) r.part[b * NV + 1] = v1;
  }
}
template <int NV>
__device__ inline void red_store(const Red &r, double v0, double v1) {
  red_store_at<NV>(r, blockIdx.x + gridDim.x * (blockIdx.y +
                                               (size_t)gridDim.y * blockIdx.z),
                   v0, v1);
}

// out[0] = sum0 (MODE 0), sum0 / sum1 (MODE 1; 0 when sum1 == 0, a zero
// coarse correction) or sum1 (MODE 2).  Fixed order: thread t sums blocks
// t, t+1024, ...
template <int NV, int MODE>
__device__ __forceinline__ void red_final_body(const double *part,
                                               unsigned nblk, double *out) {
  __shared__ double lds[16];
  double a = 0.0, c = 0.0;
  for (unsigned q = threadIdx.x; q < nblk; q += 1024) {
    a += part[(size_t)q * NV];
    if (NV == 2) c += part[(size_t)q * NV + 1];
  }
  const double sa = block_sum<1024>(a, lds);
  const double sc = NV == 2 ? block_sum<1024>(c, lds) : 0.0;
  if (threadIdx.x == 0) {
    if (MODE == 1)
      out[0] = sc != 0.0 ? sa / sc : 0.0;
    else if (MODE == 2)
      out[0] = sc;
    else
      out[0] = sa;
  }
}
template <int NV, int MODE>
__global__ __launch_bounds__(1024) void k_red_final(const double *part,
                                                    unsigned nblk,
                                                    double *out) {
  red_final_body<NV, MODE>(part, nblk, out);
}

template <int NV, int MODE>
static void red_final(const Red &r, const dim3 &g, hipStream_t s) {
  k_red_final<NV, MODE><<<1, 1024, 0, s>>>(r.part, g.x * g.y * g.z, r.out);
}

// ---------------------------------------------------------------------------
// 27-point neighbour sums in the reference's grouping (matvecd.cc:137-170)
// ---------------------------------------------------------------------------
template <class T, class XT>
__device__ inline void sum27(const T *__restrict__ ac, size_t n, size_t sy,
                             size_t sz, size_t p, const XT *x, double &sO,
                             double &sU, double &sD) {
  const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n, *oNE = ac + 4 * n,
          *oNW = ac + 5 * n, *uE = ac + 6 * n, *uW = ac + 7 * n,
          *uN = ac + 8 * n, *uS = ac + 9 * n, *uNE = ac + 10 * n,
          *uNW = ac + 11 * n, *uSE = ac + 12 * n, *uSW = ac + 13 * n;
  using A = T;
  A o = oN[p] * (A)x[p + sy] + oN[p - sy] * (A)x[p - sy] +
        oE[p] * (A)x[p + 1] + oE[p - 1] * (A)x[p - 1] +
        oNE[p] * (A)x[p + 1 + sy] + oNW[p] * (A)x[p - 1 + sy] +
        oNW[p + 1 - sy] * (A)x[p + 1 - sy] +
        oNE[p - 1 - sy] * (A)x[p - 1 - sy];
  A u = uC[p] * (A)x[p + sz] + uN[p] * (A)x[p + sy + sz] +
        uS[p] * (A)x[p - sy + sz] + uE[p] * (A)x[p + 1 + sz] +
        uW[p] * (A)x[p - 1 + sz] + uNE[p] * (A)x[p + 1 + sy + sz] +
        uNW[p] * (A)x[p - 1 + sy + sz] + uSE[p] * (A)x[p + 1 - sy + sz] +
        uSW[p] * (A)x[p - 1 - sy + sz];
  A dd = uC[p - sz] * (A)x[p - sz] + uS[p + sy - sz] * (A)x[p + sy - sz] +
         uN[p - sy - sz] * (A)x[p - sy - sz] +
         uW[p + 1 - sz] * (A)x[p + 1 - sz] +
         uE[p - 1 - sz] * (A)x[p - 1 - sz] +
         uSW[p + 1 + sy - sz] * (A)x[p + 1 + sy - sz] +
         uSE[p - 1 + sy - sz] * (A)x[p - 1 + sy - sz] +
         uNW[p + 1 - sy - sz] * (A)x[p + 1 - sy - sz] +
         uNE[p - 1 - sy - sz] * (A)x[p - 1 - sy - sz];
  sO = o;
  sU = u;
  sD = dd;
}

template <class T, class XT>
__device__ inline T sum7(const T *__restrict__ ac, size_t n, size_t sy,
                         size_t sz, size_t p, const XT *x) {
  const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  return oN[p] * (T)x[p + sy] + oN[p - sy] * (T)x[p - sy] +
         oE[p] * (T)x[p + 1] + oE[p - 1] * (T)x[p - 1] +
         uC[p - sz] * (T)x[p - sz] + uC[p] * (T)x[p + sz];
}

size_t max_partials(const Dim &d) {
  // bound for any k chunk (kchunk() may pick one plane per block)
  return (size_t)cdiv(d.nx, BX) * cdiv(d.ny, BY) * d.nz + 64;
}

// Planes per block for the residual and matvec kernels: KCH on large
// levels; one on levels of at most `small` interior planes, whose few
// blocks would otherwise wait for KCH dependent memory round trips each.
// The matvec's xdamp partials (one pair per block) are summed again by
// every workgroup of the following prolongation, so it keeps KCH above 40
// planes; the residual inside the cycle has no partials.
static int kchunk(const Dim &d, int small = 40) {
  return d.nz - 2 <= small ? 1 : KCH;
}

// ---------------------------------------------------------------------------
// Smoothers: Vgsrb7x (gsd.cc:60-111) and Vgsrb27x (gsd.cc:113-212)
// ---------------------------------------------------------------------------

// One red or black half-sweep.  Each thread owns a pair of i-adjacent
// points of which exactly one has the active colour.
template <class T>
__global__ __launch_bounds__(256) void k_gs7(int nx, int ny, int nz,
                                             const T *__restrict__ ac,
                                             const T *__restrict__ cc,
                                             const T *__restrict__ fc, T *x,
                                             int colour) {
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  const int ip = blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k0 = 2 + blockIdx.z * KCH;
  if (j > ny - 1) return;
  for (int kk = 0; kk < KCH; kk++) {
    const int k = k0 + kk;
    if (k > nz - 1) break;
    const int i = 2 + 2 * ip + ((colour + j + k) & 1);
    if (i > nx - 1) continue;
    const size_t p = ix(nx, ny, i, j, k);
    const T s = sum7<T, T>(ac, n, sy, sz, p, x);
    x[p] = (fc[p] + s) / (ac[p] + cc[p]);
  }
}

// One colour of the 8-colour sweep: (i%2) + 2(j%2) + 4(k%2) == colour.
template <class T>
__global__ __launch_bounds__(256) void k_gs27(int nx, int ny, int nz,
                                              const T *__restrict__ ac,
                                              const T *__restrict__ cc,
                                              const T *__restrict__ fc, T *x,
                                              int colour) {
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  const int i = 2 + (colour & 1) + 2 * (int)(blockIdx.x * BX + threadIdx.x);
  const int j = 2 + ((colour >> 1) & 1) + 2 * (int)(blockIdx.y * BY + threadIdx.y);
  const int k = 2 + ((colour >> 2) & 1) + 2 * (int)blockIdx.z;
  if (i > nx - 1 || j > ny - 1 || k > nz - 1) return;
  const size_t p = ix(nx, ny, i, j, k);
  double sO, sU, sD;
  sum27<T, T>(ac, n, sy, sz, p, x, sO, sU, sD);
  x[p] = (fc[p] + (T)((T)((T)sO + (T)sU) + (T)sD)) / (ac[p] + cc[p]);
}

// One full red-black sweep in a single pass: xout = black(red(xin)), the
// two half-sweeps of Vgsrb7x (gsd.cc:81-110) in the reference colour order.
//
// A workgroup owns a (TX-4) x (TY-4) tile of (i, j) columns and marches a
// chunk of KC planes in k; each thread holds one column of a TX x TY tile
// that includes two rings of halo.  Per plane each thread loads only its own
// point of the six operator arrays and of xin (prefetched into registers,
// see the rings below) and shares the in-plane operands through LDS:
//   - first-colour values ("mid") are computed on the tile minus the outer
//     ring, reading xin and the west/south couplings oE(i-1), oN(j-1) of
//     their neighbours from LDS;
//   - second-colour values are computed on the owned tile (minus two rings)
//     from the neighbours' mid values in LDS and the column's mid values of
//     planes k-1 and k+1 in registers.
// So every array element is fetched once per tile (plus the halo), all
// loads are coalesced rows, and one barrier per plane orders the LDS
// traffic: xin planes rotate over 2 slots, couplings and mid values over 3
// (a plane's slot is rewritten two planes later, after a barrier every
// reader has passed).
constexpr int SW_TX = 64, SW_TY = 16;
template <class T>
struct SwCo {  // own-point operator values of one plane
  T oC, cc, fc, oE, oN, uC;
};

// ZX: xin is zero everywhere (the sweep after Vazeros) and is not read.
// DG: the diagonal oC + cc comes precomputed in dg (the same sum, so the
// same bits; one array less to read); ac and cc then serve only for the
// couplings.
// ADD: the sweep's result is not stored; it is added to the FP64 iterate
// instead, uadd[p] += (double)x[p] at each interior point, with the
// arithmetic of k_addcorr (the last sweep of a mixed-precision correction
// cycle; every interior point is read and written by its owning thread
// only, so no other thread sees the update).
template <class T, bool ZX, bool DG, bool ADD = false>
__device__ __forceinline__ void gs7_sweep_body(
    int nx, int ny, int nz, const T *__restrict__ ac, const T *__restrict__ cc,
    const T *__restrict__ fc, const T *__restrict__ xin, T *__restrict__ xout,
    int c1, int kc, int gx, int gy, int gz, unsigned bid,
    const T *__restrict__ dg, double *__restrict__ uadd = nullptr) {
  __shared__ T SX[2][SW_TY][SW_TX], SE[3][SW_TY][SW_TX], SN[3][SW_TY][SW_TX],
      SM[3][SW_TY][SW_TX];
  // XCD-aware tile order: workgroups are dispatched round-robin over the 8
  // XCDs, so workgroup b runs on XCD b % 8.  Each XCD gets a contiguous run
  // of tiles (x fastest, then y, then k chunk): neighbouring tiles, which
  // read each other's halo rows and columns, share that XCD's L2.
  const int ntile = gx * gy * gz, per = (ntile + 7) / 8;
  const int t = (int)(bid % 8) * per + (int)(bid / 8);
  if (t >= ntile) return;  // whole workgroup, before any barrier
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = (t % gx) * (SW_TX - 4) + tx;
  const int j = ((t / gx) % gy) * (SW_TY - 4) + ty;
  const int kb = 2 + (t / (gx * gy)) * kc;
  const int ke = min(kb + kc - 1, nz - 1);
  const size_t n = (size_t)nx * ny * nz, sz = (size_t)nx * ny;
  const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  const bool inner = i >= 2 && i <= nx - 1 && j >= 2 && j <= ny - 1;
  const bool midr = inner && tx >= 1 && tx <= SW_TX - 2 && ty >= 1 &&
                    ty <= SW_TY - 2;
  const bool owned = inner && tx >= 2 && tx <= SW_TX - 3 && ty >= 2 &&
                     ty <= SW_TY - 3;
  // column base, clamped into the grid: loads stay unconditional
  const size_t p0 = ix(nx, ny, min(max(i, 1), nx), min(max(j, 1), ny), 1);
  auto at = [&](int q) { return p0 + (size_t)(min(max(q, 1), nz) - 1) * sz; };
  auto load = [&](int q, SwCo<T> &P) {
    const size_t p = at(q);
    if constexpr (DG) {
      P.oC = dg[p], P.cc = (T)0;  // oC + cc == dg (dg >= 0)
    } else {
      P.oC = ac[p], P.cc = cc[p];
    }
    P.fc = fc[p], P.oE = oE[p], P.oN = oN[p], P.uC = uC[p];
  };
  auto xat = [&](int q) -> T { return ZX ? (T)0 : xin[at(q)]; };
  auto first = [&](int q) {
    return midr && q >= 2 && q <= nz - 1 && ((i + j + q) & 1) == c1;
  };
  // first-colour value at plane q (reference term order, gsd.cc:98-105)
  auto fval = [&](int q, const SwCo<T> &P, T xm, T xp, T uCm) -> T {
    const int s = q % 3, sx = q & 1;
    const T sm = P.oN * SX[sx][ty + 1][tx] + SN[s][ty - 1][tx] * SX[sx][ty - 1][tx] +
                 P.oE * SX[sx][ty][tx + 1] + SE[s][ty][tx - 1] * SX[sx][ty][tx - 1] +
                 uCm * xm + P.uC * xp;
    return (P.fc + sm) / (P.oC + P.cc);
  };
  // Register rings of five planes, the operator values three planes ahead
  // and xin four planes ahead: step k reads the operator of planes k-1
  // (uC), k and k+1 and xin of planes k, k+1 and k+2, all loaded at least
  // two steps earlier.  Plane q sits in slot (q - kb) mod 5 of each ring and
  // the loop is unrolled by 5, so every slot is one fixed set of registers
  // (rotating values between registers instead would copy registers whose
  // loads are still in flight and make each step wait for the step before).
  SwCo<T> C[5];
  T X[5];
  T mprev, mcur;
  // ADD: the FP64 iterate of the owned column, two planes ahead (plane q in
  // slot (q - kb) mod 5, as in the rings above)
  double U[5];
  if constexpr (ADD) {
    if (owned) {
      U[0] = uadd[at(kb)];
      U[1] = uadd[at(min(kb + 1, ke))];
    }
  }
  {
    // Load order as in the loop's steady state: the last loads issued here
    // (C[2], X[3]) are the ones the first step does not read, so that step
    // waits for the same loads whether it is entered from here or from the
    // loop's back edge.
    // plane kb-2: its uC and x enter plane kb-1's first colour
    const T uCq = uC[at(kb - 2)];
    const T xq = xat(kb - 2);
    load(kb - 1, C[4]);
    X[4] = xat(kb - 1);
    load(kb, C[0]);
    X[0] = xat(kb);
    X[1] = xat(kb + 1);
    load(kb + 1, C[1]);
    X[2] = xat(kb + 2);
    load(kb + 2, C[2]);
    X[3] = xat(kb + 3);
    SX[(kb - 1) & 1][ty][tx] = X[4];
    SE[(kb - 1) % 3][ty][tx] = C[4].oE;
    SN[(kb - 1) % 3][ty][tx] = C[4].oN;
    SX[kb & 1][ty][tx] = X[0];
    SE[kb % 3][ty][tx] = C[0].oE;
    SN[kb % 3][ty][tx] = C[0].oN;
    __syncthreads();
    mprev = first(kb - 1) ? fval(kb - 1, C[4], xq, X[0], uCq) : X[4];
  }
  mcur = first(kb) ? fval(kb, C[0], X[4], X[1], C[4].uC) : X[0];
  SM[kb % 3][ty][tx] = mcur;
  // the first step rewrites xin slot (kb+1)&1 == (kb-1)&1, which the
  // plane kb-1 values above read
  __syncthreads();
  auto step = [&](int k, auto slot) __attribute__((always_inline)) {
    constexpr int r = decltype(slot)::value;
    const SwCo<T> &P0 = C[r], &P1 = C[(r + 1) % 5];
    const T uCm = C[(r + 4) % 5].uC;  // plane k-1
    // the operator is read up to plane ke+1 and xin up to ke+2; the last
    // steps re-read those planes (L2 hits) so that the loads stay
    // unconditional
    load(min(k + 3, ke + 1), C[(r + 3) % 5]);
    X[(r + 4) % 5] = xat(min(k + 4, ke + 2));
    if constexpr (ADD) {
      if (owned) U[(r + 2) % 5] = uadd[at(min(k + 2, ke))];
    }
    const T x0 = X[r], x1 = X[(r + 1) % 5], x2 = X[(r + 2) % 5];
    const int s1 = (k + 1) % 3, s0 = k % 3;
    SX[(k + 1) & 1][ty][tx] = x1;
    SE[s1][ty][tx] = P1.oE;
    SN[s1][ty][tx] = P1.oN;
    __syncthreads();
    const T mnext = first(k + 1) ? fval(k + 1, P1, x0, x2, P0.uC) : x1;
    SM[s1][ty][tx] = mnext;
    if (owned) {
      T v = mcur;
      if (((i + j + k) & 1) != c1) {
        const T sm = P0.oN * SM[s0][ty + 1][tx] +
                     SN[s0][ty - 1][tx] * SM[s0][ty - 1][tx] +
                     P0.oE * SM[s0][ty][tx + 1] +
                     SE[s0][ty][tx - 1] * SM[s0][ty][tx - 1] +
                     uCm * mprev + P0.uC * mnext;
        v = (P0.fc + sm) / (P0.oC + P0.cc);
      }
      if constexpr (ADD)
        uadd[p0 + (size_t)(k - 1) * sz] = U[r] + (double)v;
      else
        xout[p0 + (size_t)(k - 1) * sz] = v;
    }
    mprev = mcur;
    mcur = mnext;
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  int k = kb;
  for (; k + 4 <= ke; k += 5) {
    step(k, I0{});
    step(k + 1, I1{});
    step(k + 2, I2{});
    step(k + 3, I3{});
    step(k + 4, I4{});
  }
  if (k <= ke) step(k, I0{});
  if (k + 1 <= ke) step(k + 1, I1{});
  if (k + 2 <= ke) step(k + 2, I2{});
  if (k + 3 <= ke) step(k + 3, I3{});
}

template <class T, bool ZX, bool DG, bool ADD = false>
__global__ __launch_bounds__(SW_TX *SW_TY) void k_gs7_sweep(
    int nx, int ny, int nz, const T *__restrict__ ac, const T *__restrict__ cc,
    const T *__restrict__ fc, const T *__restrict__ xin, T *__restrict__ xout,
    int c1, int kc, int gx, int gy, int gz, const T *__restrict__ dg,
    double *__restrict__ uadd) {
  gs7_sweep_body<T, ZX, DG, ADD>(nx, ny, nz, ac, cc, fc, xin, xout, c1, kc,
                                 gx, gy, gz, blockIdx.x, dg, uadd);
}

// Compute units of the current device (cached per device).
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess || v <= 0)
      v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

// nsweeps full sweeps, ping-ponging between x and tmp (same boundary
// values); the result is left in x.  The k chunk is sized so that one round
// of workgroups (one per compute unit: the LDS and registers of a
// 1024-thread workgroup fill a CU) covers the grid, and is at least 4 planes
// (small grids, whose sweeps are latency-bound, take more workgroups with
// shorter marches: 65^3 mixed solve 1.70 -> 1.62 ms against an 8-plane
// floor).  zero_init: x starts
// at zero (Vazeros), boundary included, and its old contents are not read.
// uadd (FP32 with dg only): the last sweep adds its result to this FP64
// iterate instead of storing it, and x is left as it was (returns true).
template <class T>
bool gs7_fused(const Dim &d, const T *ac, const T *cc, const T *fc, T *x,
               T *tmp, int nsweeps, int iadjoint, hipStream_t s,
               bool zero_init, const T *dg, double *uadd) {
  if (d.nx < 3 || d.ny < 3 || d.nz < 3 || nsweeps <= 0) return false;
  const bool add = std::is_same<T, float>::value && uadd && dg && !zero_init;
  const int gx = cdiv(d.nx - 2, SW_TX - 4), gy = cdiv(d.ny - 2, SW_TY - 4);
  const int zb = std::max(1, device_cus() / (gx * gy));
  const int kc = std::max(4, cdiv(d.nz - 2, zb));
  const int gz = cdiv(d.nz - 2, kc);
  const int nblk = 8 * cdiv(gx * gy * gz, 8);
  dim3 b(SW_TX, SW_TY);
  const int c1 = iadjoint ? 1 : 0;
  T *src = x, *dst = tmp;
  for (int it = 0; it < nsweeps; it++) {
#define SWEEP(ZX, DG, ADD)                                                    \
  k_gs7_sweep<T, ZX, DG, ADD><<<nblk, b, 0, s>>>(                             \
      d.nx, d.ny, d.nz, ac, cc, fc, src, dst, c1, kc, gx, gy, gz, dg, uadd)
    if (add && it == nsweeps - 1) {
      if constexpr (std::is_same<T, float>::value) SWEEP(false, true, true);
      check_launch("gs7_sweep");
      return true;
    }
    if (zero_init && it == 0) {
      if (dg) SWEEP(true, true, false);
      else SWEEP(true, false, false);
    } else {
      if (dg) SWEEP(false, true, false);
      else SWEEP(false, false, false);
    }
#undef SWEEP
    T *t = src;
    src = dst;
    dst = t;
  }
  if (src != x)
    hipMemcpyAsync(x, src, d.n() * sizeof(T), hipMemcpyDeviceToDevice, s);
  check_launch("gs7_sweep");
  return false;
}

template <class T>
void gs7(const Dim &d, const T *ac, const T *cc, const T *fc, T *x,
         int nsweeps, int iadjoint, hipStream_t s) {
  if (d.nx < 3 || d.ny < 3 || d.nz < 3) return;
  const int npairs = cdiv(d.nx - 2, 2);
  dim3 g(cdiv(npairs, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  dim3 b(BX, BY);
  for (int it = 0; it < nsweeps; it++)
    for (int c = 0; c <= 1; c++) {
      const int col = (1 - iadjoint) * c + iadjoint * (1 - c);
      k_gs7<T><<<g, b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc, fc, x, col);
    }
  check_launch("gs7");
}

// A whole 8-colour sweep (Vgsrb27x, gsd.cc:160-211) as four row-parallel
// launches.  Colour (i%2) + 2(j%2) + 4(k%2) in order 0..7 visits the row
// classes (k%2, j%2) = (0,0), (0,1), (1,0), (1,1), each holding the colours
// i%2 = 0 then 1.  A 27-point stencil couples a row only to rows of other
// classes (fixed while this class is updated) and to itself, so the rows of
// one class are independent; inside a row the second colour reads its i+-1
// neighbours of the first colour, carried between adjacent lanes.  One
// wavefront per row, lane l owns the points i0 = 2+2l and i0+1.  Every point
// sees the neighbour values of the one-launch-per-colour order: identical
// results.  iadjoint reverses the class order and the in-row order.
// zmask: row classes (bit 2*(k%2) + (j%2)) whose values are still zero in
// this sweep (zero start, Vazeros before the smoothing): read as 0.
// Rows up to 128 * PPL + 1 points: lane l owns the point pairs
// m = l + 64 q (q < PPL), i0 = 2 + 2m and i0 + 1.  Pair m's west neighbour
// is pair m-1's second point (the previous lane, or lane 63 of the previous
// segment for lane 0); its east neighbour is pair m+1's first point.
constexpr int ROWS_MAXN = 128 * 4 + 1;
// A point's operands for its update (gsd.cc:160-211), as loaded: the
// couplings and neighbour values of its own plane (all but the two in-row
// neighbours), of the plane above (u) and of the plane below (d).
template <class T>
struct G27Raw {
  T cN[2], xN[2];  // oN(p) x(p+sy), oN(p-sy) x(p-sy)
  T cE, cW;        // oE(p), oE(p-1): the in-row couplings
  T c[4], xv[4];   // the other four in-plane terms
  T cu[9], xu[9], cd[9], xd[9];
  T f, oC, cc;
};

// The terms that do not involve the two in-row neighbours, summed in the
// reference's order as far as they go.
template <class T>
struct G27Part {
  T a;             // oN(p) x(p+sy) + oN(p-sy) x(p-sy): the first two terms
  T cE, cW;
  T c[4], xv[4];   // the four in-plane terms after the in-row ones
  T u, dd, f, dg;  // upper and lower plane sums, fc, oC + cc
};

// One row (j, k) of class (pz, py) of a 27-point sweep, by one wavefront.
// The neighbour rows come from four arrays by their offset (dy, dz) from
// this row: xO (dy, dz even: the row itself), xN (dy odd, dz = 0), xU
// (dy = 0, dz odd), xNU (dy and dz odd); the new values go to xW.  A
// sweep kept in one array passes it for all five.
template <class T, int ADJ, int PPL>
__device__ __forceinline__ void gs27_row(
    int nx, int ny, int nz, const T *__restrict__ ac, const T *__restrict__ cc,
    const T *__restrict__ fc, int pz, int py, int j, int k, int zmask,
    const T *xO, const T *xN, const T *xU, const T *xNU, T *xW) {
  const int lane = threadIdx.x & 63;
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n, *oNE = ac + 4 * n,
          *oNW = ac + 5 * n, *uE = ac + 6 * n, *uW = ac + 7 * n,
          *uN = ac + 8 * n, *uS = ac + 9 * n, *uNE = ac + 10 * n,
          *uNW = ac + 11 * n, *uSE = ac + 12 * n, *uSW = ac + 13 * n;
  // neighbour row classes: (dy, dz) -> zero?
  auto zr = [&](int dy, int dz) -> bool {
    const int cls = 2 * (pz ^ (dz != 0)) + (py ^ (dy != 0));
    return (zmask >> cls) & 1;
  };
  const bool zN = zr(1, 0), zU = zr(0, 1), zNU = zr(1, 1);  // classes
  const bool zown = (zmask >> (2 * pz + py)) & 1;
  const size_t rb = ix(nx, ny, 1, j, k);  // row base (i = 1)
  // The loads are unconditional: rows that are still zero in a zero-start
  // sweep are loaded and replaced by 0 afterwards, and points past the row
  // end read the last interior point's operands and are not stored.
  auto load = [&](int i) -> G27Raw<T> {
    const size_t p = rb + min(i, nx - 1) - 1;
    G27Raw<T> r;
    r.cN[0] = oN[p], r.xN[0] = xN[p + sy];
    r.cN[1] = oN[p - sy], r.xN[1] = xN[p - sy];
    r.cE = oE[p];
    r.cW = oE[p - 1];
    r.c[0] = oNE[p], r.xv[0] = xN[p + 1 + sy];
    r.c[1] = oNW[p], r.xv[1] = xN[p - 1 + sy];
    r.c[2] = oNW[p + 1 - sy], r.xv[2] = xN[p + 1 - sy];
    r.c[3] = oNE[p - 1 - sy], r.xv[3] = xN[p - 1 - sy];
    const T *cu[9] = {uC, uN, uS, uE, uW, uNE, uNW, uSE, uSW};
    const long long du[9] = {0,
                             (long long)sy,
                             -(long long)sy,
                             1,
                             -1,
                             1 + (long long)sy,
                             -1 + (long long)sy,
                             1 - (long long)sy,
                             -1 - (long long)sy};
    // the plane below couples through the mirrored plane of the neighbour
    const T *cd[9] = {uC, uS, uN, uW, uE, uSW, uSE, uNW, uNE};
    // terms 0, 3, 4 stay in the row's column of planes (dy = 0)
    const bool dy0[9] = {true, false, false, true, true,
                         false, false, false, false};
#pragma unroll
    for (int m = 0; m < 9; m++) {
      const T *xz = dy0[m] ? xU : xNU;
      r.cu[m] = cu[m][p];
      r.xu[m] = xz[p + du[m] + sz];
      r.cd[m] = cd[m][p + du[m] - sz];
      r.xd[m] = xz[p + du[m] - sz];
    }
    r.f = fc[p];
    r.oC = ac[p];
    r.cc = cc[p];
    return r;
  };
  auto part = [&](const G27Raw<T> &r) -> G27Part<T> {
    auto Z = [](T v, bool z) -> T { return z ? (T)0 : v; };
    G27Part<T> g;
    g.a = r.cN[0] * Z(r.xN[0], zN) + r.cN[1] * Z(r.xN[1], zN);
    g.cE = r.cE;
    g.cW = r.cW;
#pragma unroll
    for (int m = 0; m < 4; m++) g.c[m] = r.c[m], g.xv[m] = Z(r.xv[m], zN);
    // u: uC with the plane above's own column, the others diagonal (zNU)
    // except uE, uW (same row, zU)
    const bool zu[9] = {zU, zNU, zNU, zU, zU, zNU, zNU, zNU, zNU};
    T u = r.cu[0] * Z(r.xu[0], zu[0]);
#pragma unroll
    for (int m = 1; m < 9; m++) u += r.cu[m] * Z(r.xu[m], zu[m]);
    T dd = r.cd[0] * Z(r.xd[0], zu[0]);
#pragma unroll
    for (int m = 1; m < 9; m++) dd += r.cd[m] * Z(r.xd[m], zu[m]);
    g.u = u;
    g.dd = dd;
    g.f = r.f;
    g.dg = r.oC + r.cc;
    return g;
  };
  // the update from the part and the in-row neighbours xw (i-1), xe (i+1)
  auto fin = [&](const G27Part<T> &g, T xw, T xe) -> T {
    const T o = g.a + g.cE * xe + g.cW * xw + g.c[0] * g.xv[0] +
                g.c[1] * g.xv[1] + g.c[2] * g.xv[2] + g.c[3] * g.xv[3];
    return (g.f + ((o + g.u) + g.dd)) / g.dg;
  };
  int i0[PPL], i1[PPL];
  bool v0[PPL], v1[PPL];
  T a0[PPL], a1[PPL], n0[PPL], n1[PPL];
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    i0[q] = 2 + 2 * (lane + 64 * q);
    i1[q] = i0[q] + 1;
    v0[q] = i0[q] <= nx - 1;
    v1[q] = i1[q] <= nx - 1;
    // current values of the own row, boundary included (i <= nx)
    const T r0 = xO[rb + min(i0[q], nx) - 1], r1 = xO[rb + min(i1[q], nx) - 1];
    a0[q] = (i0[q] <= nx) ? (zown && v0[q] ? (T)0 : r0) : (T)0;
    a1[q] = (i1[q] <= nx) ? (zown && v1[q] ? (T)0 : r1) : (T)0;
    n0[q] = a0[q];
    n1[q] = a1[q];
  }
  const T xb1 = xO[rb];  // i = 1 boundary
  // One lane pair per lane (rows up to 129 points), and two in FP32 (rows
  // up to 257 points, 4 x 53 values): every load of the lane's points is
  // issued before any arithmetic, so the row waits for memory once; only
  // the in-row terms wait for the first colour.  FP64 rows of up to 257
  // points load each colour's two points together (two round trips instead
  // of four); longer rows load each point's operands just before its
  // update.  Same arithmetic in every case.
  G27Part<T> g0[PPL], g1[PPL];
  constexpr bool PRE = PPL == 1 || (PPL == 2 && sizeof(T) == 4);
  if constexpr (PPL == 1) {
    const G27Raw<T> r0 = load(i0[0]), r1 = load(i1[0]);
    __builtin_amdgcn_sched_barrier(0);
    g0[0] = part(r0);
    g1[0] = part(r1);
  } else if constexpr (PRE) {
    const G27Raw<T> r00 = load(i0[0]), r10 = load(i1[0]), r01 = load(i0[1]),
                    r11 = load(i1[1]);
    __builtin_amdgcn_sched_barrier(0);
    g0[0] = part(r00);
    g1[0] = part(r10);
    g0[1] = part(r01);
    g1[1] = part(r11);
  }
  constexpr bool PRE1 = !PRE && PPL == 2;
  auto part0 = [&](int q) -> G27Part<T> {
    return PRE ? g0[q] : part(load(i0[q]));
  };
  auto part1 = [&](int q) -> G27Part<T> {
    return PRE ? g1[q] : part(load(i1[q]));
  };
  auto batch = [&](const int *ii, G27Part<T> *g) {
    const G27Raw<T> ra = load(ii[0]), rb = load(ii[1]);
    __builtin_amdgcn_sched_barrier(0);
    g[0] = part(ra);
    g[1] = part(rb);
  };
  // value of the previous pair's second point / next pair's first point
  // (every lane takes part in the shuffles)
  auto west = [&](const T *s, int q) -> T {
    T w = __shfl_up(s[q], 1, 64);
    const T last = q > 0 ? __shfl(s[q > 0 ? q - 1 : 0], 63, 64) : xb1;
    return lane == 0 ? last : w;
  };
  auto east = [&](const T *s, int q) -> T {
    T e = __shfl_down(s[q], 1, 64);
    const T first = q + 1 < PPL ? __shfl(s[q + 1 < PPL ? q + 1 : q], 0, 64)
                                : e;  // beyond the last interior point
    return lane == 63 ? first : e;
  };
  if (ADJ == 0) {
    // colour i%2 = 0 (points i0) from the old odd values, then i%2 = 1
    if constexpr (PRE1) batch(i0, g0);
#pragma unroll
    for (int q = 0; q < PPL; q++) {
      const T w = west(a1, q);
      const G27Part<T> g = PRE1 ? g0[q] : part0(q);
      if (v0[q]) n0[q] = fin(g, w, a1[q]);
    }
    if constexpr (PRE1) batch(i1, g1);
#pragma unroll
    for (int q = 0; q < PPL; q++) {
      const T e = east(n0, q);
      const G27Part<T> g = PRE1 ? g1[q] : part1(q);
      if (v1[q]) n1[q] = fin(g, n0[q], e);
    }
  } else {
    if constexpr (PRE1) batch(i1, g1);
#pragma unroll
    for (int q = 0; q < PPL; q++) {
      const T e = east(a0, q);
      const G27Part<T> g = PRE1 ? g1[q] : part1(q);
      if (v1[q]) n1[q] = fin(g, a0[q], e);
    }
    if constexpr (PRE1) batch(i0, g0);
#pragma unroll
    for (int q = 0; q < PPL; q++) {
      const T w = west(n1, q);
      const G27Part<T> g = PRE1 ? g0[q] : part0(q);
      if (v0[q]) n0[q] = fin(g, w, n1[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    if (v0[q]) xW[rb + i0[q] - 1] = n0[q];
    if (v1[q]) xW[rb + i1[q] - 1] = n1[q];
  }
}

template <class T, int ADJ, int PPL>
__global__ __launch_bounds__(256) void k_gs27_rows(
    int nx, int ny, int nz, const T *__restrict__ ac, const T *__restrict__ cc,
    const T *__restrict__ fc, T *__restrict__ x, int pz, int py, int nj,
    int nrows, int zmask) {
  // XCD-contiguous runs of rows (see xcd_tile): the row blocks of one XCD
  // are neighbours in (j, k), in every class launch of the sweep
  const int nrb = (nrows + 3) / 4, per = (nrb + 7) / 8;
  const int rblk = (int)(blockIdx.x % 8u) * per + (int)(blockIdx.x / 8u);
  const int row = rblk * 4 + (threadIdx.x >> 6);
  if (rblk >= nrb || row >= nrows) return;  // whole wavefronts; no barrier
  const int j = 2 + py + 2 * (row % nj), k = 2 + pz + 2 * (row / nj);
  gs27_row<T, ADJ, PPL>(nx, ny, nz, ac, cc, fc, pz, py, j, k, zmask, x, x, x,
                        x, x);
}

// Two row classes of one k parity in one launch: the classes (pz, 1-bp)
// and then (pz, bp) of the sweep order, bp being the y parity of the class
// that comes second.  A workgroup of RS2 + 1 wavefronts owns a strip of
// RS2 rows of class
// (pz, bp) in one plane k and first updates the rows of class (pz, 1-bp)
// next to them (one row more than its own; a strip's edge row is also
// computed by the neighbouring strip, from the same values), then, after a
// barrier, its own rows.  The sweep reads cur and writes every new row to
// nxt: a row's old values, and the rows of classes the sweep has not
// reached, come from cur, which no launch of the sweep writes; the classes
// already updated come from nxt (the first class of this launch, read by
// the second after the barrier) and from po, the other k parity's array (cur
// in the sweep's first launch, nxt in its second).  So no workgroup reads a
// value another workgroup writes in the same launch, except the duplicated
// edge rows, which both write with the same bits.
template <class T, int ADJ, int PPL, int RS2>
__device__ __forceinline__ void gs27_pair_body(
    int nx, int ny, int nz, const T *__restrict__ ac, const T *__restrict__ cc,
    const T *__restrict__ fc, const T *cur, T *nxt, const T *po, int pz,
    int bp, int nplanes, int nstrips, int zmask, unsigned bid) {
  // XCD-contiguous runs of strips, planes outer
  const int nt = nplanes * nstrips, per = (nt + 7) / 8;
  const int t = (int)(bid % 8u) * per + (int)(bid / 8u);
  if (t >= nt) return;  // whole workgroup, before the barrier
  const int k = 2 + pz + 2 * (t / nstrips), st = t % nstrips;
  const int w = threadIdx.x >> 6;
  const int jf = 2 + bp + 2 * RS2 * st;            // first own row
  const int jl = min(jf + 2 * (RS2 - 1), ny - 1);  // own rows are <= jl
  // first class: rows jf-1, jf+1, ..., up to jl+1, inside 2..ny-1
  const int j1 = jf - 1 + 2 * w;
  if (j1 >= 2 && j1 <= ny - 1 && j1 <= jl + 1)
    gs27_row<T, ADJ, PPL>(nx, ny, nz, ac, cc, fc, pz, 1 - bp, j1, k, zmask,
                          cur, cur, po, po, nxt);
  __syncthreads();
  const int j2 = jf + 2 * w;
  if (w < RS2 && j2 <= jl)
    gs27_row<T, ADJ, PPL>(nx, ny, nz, ac, cc, fc, pz, bp, j2, k,
                          zmask & ~(1 << (2 * pz + 1 - bp)), cur, nxt, po, po,
                          nxt);
}

template <class T, int ADJ, int PPL, int RS2>
__global__ __launch_bounds__(64 * (RS2 + 1)) void k_gs27_pair(
    int nx, int ny, int nz, const T *__restrict__ ac, const T *__restrict__ cc,
    const T *__restrict__ fc, const T *cur, T *nxt, const T *po, int pz,
    int bp, int nplanes, int nstrips, int zmask) {
  gs27_pair_body<T, ADJ, PPL, RS2>(nx, ny, nz, ac, cc, fc, cur, nxt, po, pz,
                                   bp, nplanes, nstrips, zmask, blockIdx.x);
}

static int count_parity(int n, int par) {  // #{v in [2, n-1] : v%2 == par}
  const int first = 2 + par;
  return first > n - 1 ? 0 : (n - 1 - first) / 2 + 1;
}

template <class T>
void gs27(const Dim &d, const T *ac, const T *cc, const T *fc, T *x,
          int nsweeps, int iadjoint, hipStream_t s, bool zero_init, T *tmp) {
  if (d.nx < 3 || d.ny < 3 || d.nz < 3) return;
  const int ppl = d.nx <= 129 ? 1 : (d.nx <= 257 ? 2 : 4);
  // Two launches per sweep (k_gs27_pair) for pairs of sweeps, ping-ponging
  // between x and tmp, which must hold the same boundary values.  An odd
  // last sweep uses the four class launches in x.
  int done = 0;
  if (tmp && tmp != x && d.nx <= ROWS_MAXN && d.ny >= 4) {
    const int bp = iadjoint ? 0 : 1;
    constexpr int RSA = 3, RSB = 7;
    // strips of 7 rows (8 wavefronts) on levels of 65^3 to 129^3, of 3 rows
    // (4 wavefronts, more workgroups) below, where the launch is latency-
    // rather than bandwidth-bound, and above, where the rows of 7-row strips
    // overflow the L2 (measured at 145^3, 81^3, 41^3, 21^3)
    const bool big = d.nx >= 65 && d.nx <= 129;
    const int nstrips = cdiv(count_parity(d.ny, bp), big ? RSB : RSA);
    T *cur = x, *nxt = tmp;
    for (; done < (nsweeps & ~1); done++) {
      int zmask = (zero_init && done == 0) ? 0xF : 0;
      for (int h = 0; h < 2; h++) {
        const int pz = iadjoint ? 1 - h : h;
        const int npl = count_parity(d.nz, pz);
        if (npl > 0) {
          const int nb = 8 * cdiv(npl * nstrips, 8);
          const T *po = h == 0 ? cur : nxt;
#define PAIR_LAUNCH(A, Q)                                                     \
  do {                                                                        \
    if (big)                                                                  \
      k_gs27_pair<T, A, Q, RSB><<<nb, 64 * (RSB + 1), 0, s>>>(                \
          d.nx, d.ny, d.nz, ac, cc, fc, cur, nxt, po, pz, bp, npl, nstrips,   \
          zmask);                                                             \
    else                                                                      \
      k_gs27_pair<T, A, Q, RSA><<<nb, 64 * (RSA + 1), 0, s>>>(                \
          d.nx, d.ny, d.nz, ac, cc, fc, cur, nxt, po, pz, bp, npl, nstrips,   \
          zmask);                                                             \
  } while (0)
          if (iadjoint) {
            if (ppl == 1) PAIR_LAUNCH(1, 1);
            else if (ppl == 2) PAIR_LAUNCH(1, 2);
            else PAIR_LAUNCH(1, 4);
          } else {
            if (ppl == 1) PAIR_LAUNCH(0, 1);
            else if (ppl == 2) PAIR_LAUNCH(0, 2);
            else PAIR_LAUNCH(0, 4);
          }
#undef PAIR_LAUNCH
        }
        zmask &= ~(0x3 << (2 * pz));
      }
      T *t = cur;  // the sweep's result is in nxt
      cur = nxt;
      nxt = t;
    }
    // an even number of sweeps leaves the result in x
    check_launch("gs27_pair");
  }
  if (done == nsweeps) return;
  if (d.nx <= ROWS_MAXN) {
    for (int it = done; it < nsweeps; it++) {
      int zmask = (zero_init && it == 0) ? 0xF : 0;
      for (int ph = 0; ph < 4; ph++) {
        const int cls = iadjoint ? 3 - ph : ph;
        const int pz = cls >> 1, py = cls & 1;
        const int nj = count_parity(d.ny, py), nk = count_parity(d.nz, pz);
        const int nrows = nj * nk;
        if (nrows > 0) {
          const int nb = 8 * (((nrows + 3) / 4 + 7) / 8);
#define ROWS_LAUNCH(A, Q)                                                     \
  k_gs27_rows<T, A, Q><<<nb, 256, 0, s>>>(d.nx, d.ny, d.nz, ac, cc, fc, x, pz, \
                                          py, nj, nrows, zmask)
          if (iadjoint) {
            if (ppl == 1) ROWS_LAUNCH(1, 1);
            else if (ppl == 2) ROWS_LAUNCH(1, 2);
            else ROWS_LAUNCH(1, 4);
          } else {
            if (ppl == 1) ROWS_LAUNCH(0, 1);
            else if (ppl == 2) ROWS_LAUNCH(0, 2);
            else ROWS_LAUNCH(0, 4);
          }
#undef ROWS_LAUNCH
        }
        zmask &= ~(1 << cls);
      }
    }
    check_launch("gs27_rows");
    return;
  }
  if (zero_init) fill<T>(x, d.n(), (T)0, s);
  dim3 b(BX, BY);
  for (int it = 0; it < nsweeps; it++)
    for (int c = 0; c < 8; c++) {
      const int col = c * (1 - iadjoint) + iadjoint * (7 - c);
      const int px = col & 1, py = (col >> 1) & 1, pz = (col >> 2) & 1;
      // number of points of this parity in 2..n-1
      const int ni = (d.nx - 2 - px + 1) / 2, nj = (d.ny - 2 - py + 1) / 2,
                nk = (d.nz - 2 - pz + 1) / 2;
      if (ni <= 0 || nj <= 0 || nk <= 0) continue;
      dim3 g(cdiv(ni, BX), cdiv(nj, BY), nk);
      k_gs27<T><<<g, b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc, fc, x, col);
    }
  check_launch("gs27");
}

// ---------------------------------------------------------------------------
// Residual (Vmresid7_1s / Vmresid27_1s) with optional L1 partial sums: one
// thread per (i, j) column of a chunk of kc planes (7-point levels with the
// precomputed diagonal take the march k_resid7m below instead)
// ---------------------------------------------------------------------------
template <class T, class TO, int ND>
__device__ __forceinline__ void resid_body(
    int nx, int ny, int nz, const T *__restrict__ ac,
    const T *__restrict__ cc, const T *__restrict__ fc,
    const T *__restrict__ x, TO *__restrict__ r, Red rd, int do_red, int kc,
    XG xg, unsigned bid) {
  __shared__ double red[4];
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  int bx, by, bz;
  const bool tok = xcd_tile_at(xg, bid, bx, by, bz);
  const int i = 2 + bx * BX + threadIdx.x;
  const int j = 2 + by * BY + threadIdx.y;
  const int k0 = 2 + bz * kc;
  double acc = 0.0;
  if (ND == 7 && kc == KCH && tok && i <= nx - 1 && j <= ny - 1) {
    // the chunk's planes with a fixed trip count and clamped (always valid)
    // loads, so every plane's loads can be issued before the first result
    T vv[KCH];
#pragma unroll
    for (int kk = 0; kk < KCH; kk++) {
      const int k = min(k0 + kk, nz - 1);
      const size_t p = ix(nx, ny, i, j, k);
      const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
      const T dgp = ac[p] + cc[p];
      vv[kk] = fc[p] + oN[p] * x[p + sy] + oN[p - sy] * x[p - sy] +
               oE[p] * x[p + 1] + oE[p - 1] * x[p - 1] +
               uC[p - sz] * x[p - sz] + uC[p] * x[p + sz] - dgp * x[p];
    }
#pragma unroll
    for (int kk = 0; kk < KCH; kk++) {
      if (k0 + kk > nz - 1) break;
      if (r) r[ix(nx, ny, i, j, k0 + kk)] = (TO)vv[kk];
      acc += fabs((double)vv[kk]);
    }
  } else if (tok && i <= nx - 1 && j <= ny - 1) {
    for (int kk = 0; kk < kc; kk++) {
      const int k = k0 + kk;
      if (k > nz - 1) break;
      const size_t p = ix(nx, ny, i, j, k);
      T v;
      if (ND == 7) {
        const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
        const T dgp = ac[p] + cc[p];
        v = fc[p] + oN[p] * x[p + sy] + oN[p - sy] * x[p - sy] +
            oE[p] * x[p + 1] + oE[p - 1] * x[p - 1] +
            uC[p - sz] * x[p - sz] + uC[p] * x[p + sz] - dgp * x[p];
      } else {
        double sO, sU, sD;
        sum27<T, T>(ac, n, sy, sz, p, x, sO, sU, sD);
        v = (T)((T)((T)(fc[p] + (T)sO) + (T)sU) + (T)sD) -
            (ac[p] + cc[p]) * x[p];
      }
      if (r) r[p] = (TO)v;
      acc += fabs((double)v);
    }
  }
  if (do_red) {
    const double t = block_sum<256>(acc, red);
    red_store_at<1>(rd, bid, t, 0.0);
  }
}

template <class T, class TO, int ND>
__global__ __launch_bounds__(256) void k_resid(
    int nx, int ny, int nz, const T *__restrict__ ac,
    const T *__restrict__ cc, const T *__restrict__ fc,
    const T *__restrict__ x, TO *__restrict__ r, Red rd, int do_red, int kc,
    XG xg) {
  resid_body<T, TO, ND>(nx, ny, nz, ac, cc, fc, x, r, rd, do_red, kc, xg,
                        blockIdx.x);
}

// The 7-point residual with the precomputed diagonal as a march in k (the
// fine levels' residual, the FP64 outer residual of the mixed solve among
// them).  A 1024-thread workgroup owns a (RM_TX-2) x (RM_TY-2) tile of
// (i, j) columns plus one halo ring and marches a chunk of planes.  Each
// thread loads only its own point of dg, fc, oE, oN, uC and x, into
// five-slot register rings (the operator three planes ahead, x four, as in
// the fused sweep), and takes its in-plane neighbours' x and the west and
// south couplings oE(i-1), oN(j-1) from LDS: two slots, one barrier per
// plane (a slot is rewritten two planes later, after a barrier every reader
// has passed).  Against one thread per point with seven x loads through
// the caches, every element is fetched once per tile.  The sum is
// resid_body's, term for term, so r has the same bits; the L1 norm is
// summed per thread over its planes, then over the workgroup.
constexpr int RM_TX = 64, RM_TY = 16;
template <class T, class TO>
__device__ __forceinline__ void resid7m_body(
    int nx, int ny, int nz, const T *__restrict__ ac,
    const T *__restrict__ fc, const T *__restrict__ x, TO *__restrict__ r,
    const T *__restrict__ dg, const Red &rd, int do_red, int kc, int gx,
    int gy, int gz, unsigned bid) {
  __shared__ T SX[2][RM_TY][RM_TX], SE[2][RM_TY][RM_TX], SN[2][RM_TY][RM_TX];
  __shared__ double red[RM_TX * RM_TY / 64];
  // XCD-aware tile order, as in the fused sweep
  const int ntile = gx * gy * gz, per = (ntile + 7) / 8;
  const int t = (int)(bid % 8) * per + (int)(bid / 8);
  if (t >= ntile) {  // whole workgroup, before any barrier: a zero partial
    if (do_red) red_store_at<1>(rd, bid, 0.0, 0.0);
    return;
  }
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = (t % gx) * (RM_TX - 2) + tx + 1;
  const int j = ((t / gx) % gy) * (RM_TY - 2) + ty + 1;
  const int kb = 2 + (t / (gx * gy)) * kc;
  const int ke = min(kb + kc - 1, nz - 1);
  const size_t n = (size_t)nx * ny * nz, sz = (size_t)nx * ny;
  const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  const bool owned = tx >= 1 && tx <= RM_TX - 2 && ty >= 1 &&
                     ty <= RM_TY - 2 && i <= nx - 1 && j <= ny - 1;
  // column base (i, j >= 1 by construction), clamped into the grid: loads
  // stay unconditional
  const size_t p0 = ix(nx, ny, min(i, nx), min(j, ny), 1);
  auto at = [&](int q) { return p0 + (size_t)(min(max(q, 1), nz) - 1) * sz; };
  struct Co {
    T dg, fc, oE, oN, uC;
  };
  auto load = [&](int q, Co &P) {
    const size_t p = at(q);
    P.dg = dg[p], P.fc = fc[p], P.oE = oE[p], P.oN = oN[p], P.uC = uC[p];
  };
  // plane q in slot (q - kb) mod 5 of both rings; step k reads the operator
  // of planes k-1 (uC) and k and x of planes k-1, k, k+1; the operator is
  // read up to plane ke and x up to ke+1 (later steps re-read those)
  Co C[5];
  T X[5];
  load(kb - 1, C[4]);
  X[4] = x[at(kb - 1)];
  load(min(kb, ke), C[0]);
  X[0] = x[at(kb)];
  load(min(kb + 1, ke), C[1]);
  X[1] = x[at(min(kb + 1, ke + 1))];
  load(min(kb + 2, ke), C[2]);
  X[2] = x[at(min(kb + 2, ke + 1))];
  X[3] = x[at(min(kb + 3, ke + 1))];
  double acc = 0.0;
  auto step = [&](int k, auto slot) __attribute__((always_inline)) {
    constexpr int s5 = decltype(slot)::value;
    const T uCm = C[(s5 + 4) % 5].uC, xm = X[(s5 + 4) % 5];
    load(min(k + 3, ke), C[(s5 + 3) % 5]);
    X[(s5 + 4) % 5] = x[at(min(k + 4, ke + 1))];
    const Co &P = C[s5];
    const T x0 = X[s5], xp = X[(s5 + 1) % 5];
    const int s = k & 1;
    SX[s][ty][tx] = x0;
    SE[s][ty][tx] = P.oE;
    SN[s][ty][tx] = P.oN;
    __syncthreads();
    if (owned) {
      const T v = P.fc + P.oN * SX[s][ty + 1][tx] +
                  SN[s][ty - 1][tx] * SX[s][ty - 1][tx] +
                  P.oE * SX[s][ty][tx + 1] +
                  SE[s][ty][tx - 1] * SX[s][ty][tx - 1] + uCm * xm +
                  P.uC * xp - P.dg * x0;
      if (r) r[p0 + (size_t)(k - 1) * sz] = (TO)v;
      acc += fabs((double)v);
    }
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  int k = kb;
  for (; k + 4 <= ke; k += 5) {
    step(k, I0{});
    step(k + 1, I1{});
    step(k + 2, I2{});
    step(k + 3, I3{});
    step(k + 4, I4{});
  }
  if (k <= ke) step(k, I0{});
  if (k + 1 <= ke) step(k + 1, I1{});
  if (k + 2 <= ke) step(k + 2, I2{});
  if (k + 3 <= ke) step(k + 3, I3{});
  if (do_red) {
    const double tsum = block_sum<RM_TX * RM_TY>(acc, red);
    red_store_at<1>(rd, bid, tsum, 0.0);
  }
}
template <class T, class TO>
__global__ __launch_bounds__(RM_TX *RM_TY) void k_resid7m(
    int nx, int ny, int nz, const T *__restrict__ ac,
    const T *__restrict__ fc, const T *__restrict__ x, TO *__restrict__ r,
    const T *__restrict__ dg, Red rd, int do_red, int kc, int gx, int gy,
    int gz) {
  resid7m_body<T, TO>(nx, ny, nz, ac, fc, x, r, dg, rd, do_red, kc, gx, gy,
                      gz, blockIdx.x);
}

// Launch shape of k_resid7m: one round of workgroups (one per compute
// unit) covers the grid, with chunks of at least 4 planes.  Returns the
// number of workgroups (and norm partials).
static unsigned resid7m_shape(const Dim &d, int &kc, int &gx, int &gy,
                              int &gz) {
  gx = cdiv(d.nx - 2, RM_TX - 2), gy = cdiv(d.ny - 2, RM_TY - 2);
  const int zb = std::max(1, device_cus() / (gx * gy));
  kc = std::max(4, cdiv(d.nz - 2, zb));
  gz = cdiv(d.nz - 2, kc);
  return 8u * (unsigned)cdiv(gx * gy * gz, 8);
}

// The march serves a 7-point level with the precomputed diagonal when its
// tiles of RM_TX - 2 owned columns need no more tiles across i than
// resid_body's blocks of BX columns, and its k chunks are long enough to
// hide the four planes its rings load before the first result.  With 63
// or 127 interior columns (65^3, 129^3) a third of its workgroups' columns
// would be idle (129^3 Newton 7.61 -> 7.79 ms); at 97^3 its chunks are 6
// planes.  resid_body, computing the same sum in-kernel, serves there.
constexpr int RM_MIN_PLANES = 16;
static bool resid7m_fits(const Dim &d) {
  int kc, gx, gy, gz;
  resid7m_shape(d, kc, gx, gy, gz);
  return cdiv(d.nx - 2, RM_TX - 2) == cdiv(d.nx - 2, BX) &&
         kc >= RM_MIN_PLANES;
}

size_t resid_partials(const Dim &d) {
  const int kc = kchunk(d, 40);
  return xcd_blocks(
      dim3(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, kc)));
}

template <class T, class TO>
void resid(const Dim &d, int numdia, const T *ac, const T *cc, const T *fc,
           const T *x, TO *r, const Red *red, hipStream_t s, const T *dg) {
  if (numdia == 7 && dg && resid7m_fits(d)) {
    int kcm, gx, gy, gz;
    const unsigned nb = resid7m_shape(d, kcm, gx, gy, gz);
    const Red rdm = red ? *red : Red{};
    k_resid7m<T, TO><<<nb, dim3(RM_TX, RM_TY), 0, s>>>(
        d.nx, d.ny, d.nz, ac, fc, x, r, dg, rdm, red ? 1 : 0, kcm, gx, gy,
        gz);
    if (red) red_final<1, 0>(rdm, dim3(nb), s);
    check_launch("resid");
    return;
  }
  const int kc = kchunk(d, red ? 40 : 80);
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, kc));
  dim3 b(BX, BY);
  const Red rd = red ? *red : Red{};
  if (numdia == 7)
    k_resid<T, TO, 7><<<xcd_blocks(g), b, 0, s>>>(
        d.nx, d.ny, d.nz, ac, cc, fc, x, r, rd, red ? 1 : 0, kc, xg_of(g));
  else
    k_resid<T, TO, 27><<<xcd_blocks(g), b, 0, s>>>(
        d.nx, d.ny, d.nz, ac, cc, fc, x, r, rd, red ? 1 : 0, kc, xg_of(g));
  if (red) red_final<1, 0>(rd, dim3(xcd_blocks(g)), s);
  check_launch("resid");
}

// ---------------------------------------------------------------------------
// Ion terms (Vc_vecpmg / Vdc_vecpmg, mypdec.cc:136-212, 348-414)
// ---------------------------------------------------------------------------
constexpr int MAXIONS = 50;
__constant__ double c_charge[MAXIONS];
__constant__ double c_sconc[MAXIONS];
__constant__ int c_nion;
static int h_nion = 0;

void set_ions(int nion, const double *charge, const double *sconc) {
  if (nion > MAXIONS) nion = MAXIONS;
  double q[MAXIONS] = {0}, c[MAXIONS] = {0};
  for (int i = 0; i < nion; i++) q[i] = charge[i], c[i] = sconc[i];
  hipMemcpyToSymbol(HIP_SYMBOL(c_charge), q, sizeof(q));
  hipMemcpyToSymbol(HIP_SYMBOL(c_sconc), c, sizeof(c));
  hipMemcpyToSymbol(HIP_SYMBOL(c_nion), &nion, sizeof(int));
  h_nion = nion;
}

__device__ inline double ion_term(double coef, double u, int deriv) {
  const double ZSMALL = 1.0e-20, ZLARGE = 1.0e20, SMIN = -85.0, SMAX = 85.0;
  double out = 0.0;
  for (int s = 0; s < c_nion; s++) {
    const double q = c_charge[s];
    const double zcf2 = deriv ? c_sconc[s] * q * q : -1.0 * c_sconc[s] * q;
    const double zu2 = -1.0 * q;
    const double az = fmin(ZSMALL, fabs(zcf2 * coef)) * ZLARGE;
    const double an = fmax(fmin(zu2 * u, 0.0), SMIN);
    const double ap = fmin(fmax(zu2 * u, 0.0), SMAX);
    out = out + zcf2 * coef * exp(az * (an + ap));
  }
  return out;
}

// out32: the FP32 copy for the mixed-precision inner solves, rounded as a
// separate conversion would.
// dg, dg32: the Jacobian's diagonal oC + out, summed as the sweeps and
// residuals would (FP32: the sum of the FP32 copies).
__global__ void k_dc_vec(const double *coef, const double *u, double *out,
                         size_t n, float *out32, const double *oc,
                         double *dg, const float *oc32, float *dg32) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    const double v = ion_term(coef[t], u[t], 1);
    out[t] = v;
    if (out32) out32[t] = (float)v;
    if (dg) dg[t] = oc[t] + v;
    if (dg32) dg32[t] = oc32[t] + (float)v;
  }
}

void dc_vec(const double *coef, const double *u, double *out, size_t n,
            hipStream_t s, float *out32, const double *oc, double *dg,
            const float *oc32, float *dg32) {
  int nb = (int)((n + 255) / 256);
  if (nb > 65535) nb = 65535;
  k_dc_vec<<<nb, 256, 0, s>>>(coef, u, out, n, out32, oc, dg, oc32, dg32);
  check_launch("dc_vec");
}

template <int ND>
__global__ __launch_bounds__(256) void k_nmresid(
    int nx, int ny, int nz, const double *__restrict__ ac,
    const double *__restrict__ cc, const double *__restrict__ fc,
    const double *__restrict__ x, double *__restrict__ r, Red rd,
    int do_red, XG xg) {
  __shared__ double red[4];
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  int bx, by, bz;
  const bool tok = xcd_tile(xg, bx, by, bz);
  const int i = 2 + bx * BX + threadIdx.x;
  const int j = 2 + by * BY + threadIdx.y;
  const int k0 = 2 + bz * KCH;
  double acc = 0.0;
  if (tok && i <= nx - 1 && j <= ny - 1) {
    const double *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
    for (int kk = 0; kk < KCH; kk++) {
      const int k = k0 + kk;
      if (k > nz - 1) break;
      const size_t p = ix(nx, ny, i, j, k);
      const double w1 = ion_term(cc[p], x[p], 0);
      double v;
      if (ND == 7) {
        v = fc[p] + oN[p] * x[p + sy] + oN[p - sy] * x[p - sy] +
            oE[p] * x[p + 1] + oE[p - 1] * x[p - 1] + uC[p - sz] * x[p - sz] +
            uC[p] * x[p + sz] - ac[p] * x[p] - w1;
      } else {  // Vnmresid27_1s (matvecd.cc:578-693)
        double sO, sU, sD;
        sum27<double, double>(ac, n, sy, sz, p, x, sO, sU, sD);
        v = sO + sU + sD + fc[p] - ac[p] * x[p] - w1;
      }
      if (r) r[p] = v;
      acc += fabs(v);
    }
  }
  if (do_red) {
    const double t = block_sum<256>(acc, red);
    red_store<1>(rd, t, 0.0);
  }
}

void nmresid(const Dim &d, int numdia, const double *ac, const double *cc,
             const double *fc, const double *x, double *r, const Red *red,
             hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  const Red rd = red ? *red : Red{};
  if (numdia == 7)
    k_nmresid<7><<<xcd_blocks(g), dim3(BX, BY), 0, s>>>(
        d.nx, d.ny, d.nz, ac, cc, fc, x, r, rd, red ? 1 : 0, xg_of(g));
  else
    k_nmresid<27><<<xcd_blocks(g), dim3(BX, BY), 0, s>>>(
        d.nx, d.ny, d.nz, ac, cc, fc, x, r, rd, red ? 1 : 0, xg_of(g));
  if (red) red_final<1, 0>(rd, dim3(xcd_blocks(g)), s);
  check_launch("nmresid");
}

// ---------------------------------------------------------------------------
// y = A x with partial sums of x.w0 and x.y (Vmatvec + 2 Vxdot, Vxdot
// being x . y; the reference's squares x)
// ---------------------------------------------------------------------------
template <class T, int ND>
__device__ __forceinline__ void matvec_dots_body(
    int nx, int ny, int nz, const T *__restrict__ ac,
    const T *__restrict__ cc, const T *__restrict__ x,
    const T *__restrict__ w0, T *__restrict__ y, Red rd, int do_red, int kc,
    XG xg, unsigned bid) {
  __shared__ double red[4];
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  int bx, by, bz;
  const bool tok = xcd_tile_at(xg, bid, bx, by, bz);
  const int i = 2 + bx * BX + threadIdx.x;
  const int j = 2 + by * BY + threadIdx.y;
  const int k0 = 2 + bz * kc;
  double a0 = 0.0, a1 = 0.0;
  if (tok && i <= nx - 1 && j <= ny - 1) {
    for (int kk = 0; kk < kc; kk++) {
      const int k = k0 + kk;
      if (k > nz - 1) break;
      const size_t p = ix(nx, ny, i, j, k);
      T v;
      if (ND == 7) {
        const T *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
        v = -oN[p] * x[p + sy] - oN[p - sy] * x[p - sy] - oE[p] * x[p + 1] -
            oE[p - 1] * x[p - 1] - uC[p - sz] * x[p - sz] -
            uC[p] * x[p + sz] + (ac[p] + cc[p]) * x[p];
      } else {
        double sO, sU, sD;
        sum27<T, T>(ac, n, sy, sz, p, x, sO, sU, sD);
        v = (T)((T)(-(T)sO + -(T)sU) + -(T)sD) + (ac[p] + cc[p]) * x[p];
      }
      if (y) y[p] = v;
      if (w0) a0 += (double)x[p] * (double)w0[p];
      a1 += (double)x[p] * (double)v;
    }
  }
  if (do_red) {
    const double t0 = block_sum<256>(a0, red);
    const double t1 = block_sum<256>(a1, red);
    red_store_at<2>(rd, bid, t0, t1);
  }
}

template <class T, int ND>
__global__ __launch_bounds__(256) void k_matvec_dots(
    int nx, int ny, int nz, const T *__restrict__ ac,
    const T *__restrict__ cc, const T *__restrict__ x,
    const T *__restrict__ w0, T *__restrict__ y, Red rd, int do_red, int kc,
    XG xg) {
  matvec_dots_body<T, ND>(nx, ny, nz, ac, cc, x, w0, y, rd, do_red, kc, xg,
                          blockIdx.x);
}

template <class T>
void matvec_dots(const Dim &d, int numdia, const T *ac, const T *cc,
                 const T *x, const T *w0, T *y, const Red *red,
                 hipStream_t s) {
  const int kc = kchunk(d);
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, kc));
  dim3 b(BX, BY);
  const Red rd = red ? *red : Red{};
  if (numdia == 7)
    k_matvec_dots<T, 7><<<xcd_blocks(g), b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc, x, w0, y,
                                        rd, red ? 1 : 0, kc, xg_of(g));
  else
    k_matvec_dots<T, 27><<<xcd_blocks(g), b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc, x, w0, y,
                                         rd, red ? 1 : 0, kc, xg_of(g));
  if (red) red_final<2, 1>(rd, dim3(xcd_blocks(g)), s);
  check_launch("matvec_dots");
}

template <class T>
unsigned matvec_dots_part(const Dim &d, int numdia, const T *ac, const T *cc,
                          const T *x, const T *w0, T *y, double *part,
                          hipStream_t s) {
  const int kc = kchunk(d);
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, kc));
  dim3 b(BX, BY);
  Red rd;
  rd.part = part;
  if (numdia == 7)
    k_matvec_dots<T, 7><<<xcd_blocks(g), b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc, x, w0, y,
                                        rd, 1, kc, xg_of(g));
  else
    k_matvec_dots<T, 27><<<xcd_blocks(g), b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc, x, w0, y,
                                         rd, 1, kc, xg_of(g));
  check_launch("matvec_dots");
  return xcd_blocks(g);
}

// ---------------------------------------------------------------------------
// Grid transfer: Vrestrc2 (matvecd.cc:755-795), VinterpPMG2 (:878-960)
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void restrict_body(int nxf, int nyf, int nxc,
                                              int nyc, int nzc,
                                              const T *__restrict__ xin,
                                              T *__restrict__ xout,
                                              const T *__restrict__ pc, XG xg,
                                              float *__restrict__ out32,
                                              unsigned bid) {
  // no multiply-add contraction: the setup restricts cc and fc with this
  // kernel, and a source of both signs cancels in those sums, so the
  // coarse cc and fc are the oracle's to the bit only with the same
  // roundings (same term order, separate multiply and add)
#pragma clang fp contract(off)
  int bx, by, bz;
  if (!xcd_tile_at(xg, bid, bx, by, bz)) return;
  const int ii = 2 + bx * BX + threadIdx.x;
  const int jj = 2 + by * BY + threadIdx.y;
  const int kk = 2 + bz;
  if (ii > nxc - 1 || jj > nyc - 1) return;
  const size_t nc = (size_t)nxc * nyc * nzc, pC = ix(nxc, nyc, ii, jj, kk);
  const size_t sy = nxf, sz = (size_t)nxf * nyf;
  const size_t f = ix(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1);
  // in-plane offsets in P plane order C,N,S,E,W,NE,NW,SE,SW
  const long long off[9] = {0,
                            (long long)sy,
                            -(long long)sy,
                            1,
                            -1,
                            1 + (long long)sy,
                            -1 + (long long)sy,
                            1 - (long long)sy,
                            -1 - (long long)sy};
  const long long zo[3] = {0, (long long)sz, -(long long)sz};
  T t[3];
#pragma unroll
  for (int b = 0; b < 3; b++) {
    T acc = pc[(size_t)(9 * b) * nc + pC] * xin[f + zo[b]];
#pragma unroll
    for (int q = 1; q < 9; q++)
      acc += pc[(size_t)(9 * b + q) * nc + pC] * xin[f + zo[b] + off[q]];
    t[b] = acc;
  }
  const T v = t[0] + t[1] + t[2];
  xout[pC] = v;
  if (out32) out32[pC] = (float)v;
}

template <class T>
__global__ __launch_bounds__(256) void k_restrict(int nxf, int nyf, int nxc,
                                                  int nyc, int nzc,
                                                  const T *__restrict__ xin,
                                                  T *__restrict__ xout,
                                                  const T *__restrict__ pc, XG xg,
                                                  float *__restrict__ out32) {
  restrict_body<T>(nxf, nyf, nxc, nyc, nzc, xin, xout, pc, xg, out32,
                   blockIdx.x);
}

template <class T>
void restrict_(const Dim &f, const Dim &c, const T *xin, T *xout, const T *pc,
               hipStream_t s, float *out32) {
  if (c.nx < 3 || c.ny < 3 || c.nz < 3) return;
  if (f.nx != 2 * c.nx - 1 || f.ny != 2 * c.ny - 1 || f.nz != 2 * c.nz - 1) {
    fprintf(stderr, "pmg: restrict shape mismatch\n");
    return;
  }
  dim3 g(cdiv(c.nx - 2, BX), cdiv(c.ny - 2, BY), c.nz - 2);
  k_restrict<T><<<xcd_blocks(g), dim3(BX, BY), 0, s>>>(f.nx, f.ny, c.nx, c.ny, c.nz, xin,
                                           xout, pc, xg_of(g), out32);
  check_launch("restrict");
}

// Two sources restricted in one pass (the setup's cc and fc): every P plane
// is loaded once for both sums, each of which keeps restrict_body's term
// order, so the results are those of two k_restrict launches to the bit.
__global__ __launch_bounds__(256) void k_restrict2(
    int nxf, int nyf, int nxc, int nyc, int nzc, const double *__restrict__ a,
    const double *__restrict__ b, double *__restrict__ ao,
    double *__restrict__ bo, const double *__restrict__ pc, XG xg,
    float *__restrict__ ao32, float *__restrict__ bo32) {
#pragma clang fp contract(off)
  int bx, by, bz;
  if (!xcd_tile(xg, bx, by, bz)) return;
  const int ii = 2 + bx * BX + threadIdx.x;
  const int jj = 2 + by * BY + threadIdx.y;
  const int kk = 2 + bz;
  if (ii > nxc - 1 || jj > nyc - 1) return;
  const size_t nc = (size_t)nxc * nyc * nzc, pC = ix(nxc, nyc, ii, jj, kk);
  const size_t sy = nxf, sz = (size_t)nxf * nyf;
  const size_t f = ix(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1);
  const long long off[9] = {0,
                            (long long)sy,
                            -(long long)sy,
                            1,
                            -1,
                            1 + (long long)sy,
                            -1 + (long long)sy,
                            1 - (long long)sy,
                            -1 - (long long)sy};
  const long long zo[3] = {0, (long long)sz, -(long long)sz};
  double ta[3], tb[3];
#pragma unroll
  for (int z = 0; z < 3; z++) {
    const double w0 = pc[(size_t)(9 * z) * nc + pC];
    double sa = w0 * a[f + zo[z]], sb = w0 * b[f + zo[z]];
#pragma unroll
    for (int q = 1; q < 9; q++) {
      const double w = pc[(size_t)(9 * z + q) * nc + pC];
      sa += w * a[f + zo[z] + off[q]];
      sb += w * b[f + zo[z] + off[q]];
    }
    ta[z] = sa;
    tb[z] = sb;
  }
  const double va = ta[0] + ta[1] + ta[2];
  ao[pC] = va;
  if (ao32) ao32[pC] = (float)va;
  const double vb = tb[0] + tb[1] + tb[2];
  bo[pC] = vb;
  if (bo32) bo32[pC] = (float)vb;
}

void restrict2(const Dim &f, const Dim &c, const double *a, const double *b,
               double *ao, double *bo, const double *pc, hipStream_t s,
               float *ao32, float *bo32) {
  if (c.nx < 3 || c.ny < 3 || c.nz < 3) return;
  if (f.nx != 2 * c.nx - 1 || f.ny != 2 * c.ny - 1 || f.nz != 2 * c.nz - 1) {
    fprintf(stderr, "pmg: restrict shape mismatch\n");
    return;
  }
  dim3 g(cdiv(c.nx - 2, BX), cdiv(c.ny - 2, BY), c.nz - 2);
  k_restrict2<<<xcd_blocks(g), dim3(BX, BY), 0, s>>>(
      f.nx, f.ny, c.nx, c.ny, c.nz, a, b, ao, bo, pc, xg_of(g), ao32, bo32);
  check_launch("restrict2");
}

enum : int {
  oPC, oPN, oPS, oPE, oPW, oPNE, oPNW, oPSE, oPSW,
  uPC, uPN, uPS, uPE, uPW, uPNE, uPNW, uPSE, uPSW,
  dPC, dPN, dPS, dPE, dPW, dPNE, dPNW, dPSE, dPSW
};

// One thread per coarse cell (ii, jj, kk) in 1..nc-1: the 8 fine points of
// the cell, as VinterpPMG2.  ADD: xf += xdamp * value on fine interior
// points (Vxaxpy over the interior); otherwise xf = value (the fine
// boundary is zeroed separately).
template <class T, bool ADD>
__device__ __forceinline__ void interp_body(int nxc, int nyc, int nzc,
                                            int nxf, int nyf, int nzf,
                                            const T *__restrict__ xc,
                                            const T *__restrict__ pc,
                                            T *__restrict__ xf,
                                            const double *xdamp,
                                            const double *part,
                                            unsigned nblk, XG xg,
                                            unsigned bid) {
  int bx, by, bz;
  const bool tok = xcd_tile_at(xg, bid, bx, by, bz);
  const int ii = 1 + bx * BX + threadIdx.x;
  const int jj = 1 + by * BY + threadIdx.y;
  const int kk = 1 + bz;
  // xdamp from the (x.w0, x.y) partials of the coarse matvec, summed by
  // every block in the same fixed order (mgcsd.cc:429-452)
  __shared__ double red[8];
  __shared__ double dsh;
  if (ADD && part) {
    const int tid = threadIdx.x + threadIdx.y * BX;
    double a = 0.0, c = 0.0;
    for (unsigned q = tid; q < nblk; q += BX * BY) {
      a += part[2 * (size_t)q];
      c += part[2 * (size_t)q + 1];
    }
    a = block_sum<BX * BY>(a, red);
    c = block_sum<BX * BY>(c, red);
    if (tid == 0) dsh = c != 0.0 ? a / c : 0.0;
    __syncthreads();
  }
  if (!tok || ii > nxc - 1 || jj > nyc - 1) return;
  const size_t nc = (size_t)nxc * nyc * nzc;
  const size_t c0 = ix(nxc, nyc, ii, jj, kk);
  const size_t cx = 1, cy = nxc, cz = (size_t)nxc * nyc;
#define PW(pl, o) pc[(size_t)(pl)*nc + c0 + (o)]
#define XC(o) xc[c0 + (o)]
  const T x000 = XC(0), x100 = XC(cx), x010 = XC(cy), x110 = XC(cx + cy),
          x001 = XC(cz), x101 = XC(cx + cz), x011 = XC(cy + cz),
          x111 = XC(cx + cy + cz);
  T v[8];
  v[0] = x000;
  v[1] = PW(oPE, 0) * x000 + PW(oPW, cx) * x100;
  v[2] = PW(oPN, 0) * x000 + PW(oPS, cy) * x010;
  v[4] = PW(uPC, 0) * x000 + PW(dPC, cz) * x001;
  v[3] = PW(oPNE, 0) * x000 + PW(oPNW, cx) * x100 + PW(oPSE, cy) * x010 +
         PW(oPSW, cx + cy) * x110;
  v[5] = PW(uPE, 0) * x000 + PW(uPW, cx) * x100 + PW(dPE, cz) * x001 +
         PW(dPW, cx + cz) * x101;
  v[6] = PW(uPN, 0) * x000 + PW(uPS, cy) * x010 + PW(dPN, cz) * x001 +
         PW(dPS, cy + cz) * x011;
  v[7] = +PW(uPNE, 0) * x000 + PW(uPNW, cx) * x100 + PW(uPSE, cy) * x010 +
         PW(uPSW, cx + cy) * x110 + PW(dPNE, cz) * x001 +
         PW(dPNW, cx + cz) * x101 + PW(dPSE, cy + cz) * x011 +
         PW(dPSW, cx + cy + cz) * x111;
#undef PW
#undef XC
  const int i = 2 * ii - 1, j = 2 * jj - 1, k = 2 * kk - 1;
  T damp = (T)0;
  if (ADD) damp = (T)(part ? dsh : *xdamp);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int fi = i + (q & 1), fj = j + ((q >> 1) & 1), fk = k + (q >> 2);
    const size_t p = ix(nxf, nyf, fi, fj, fk);
    if (ADD) {
      if (fi >= 2 && fi <= nxf - 1 && fj >= 2 && fj <= nyf - 1 && fk >= 2 &&
          fk <= nzf - 1)
        xf[p] += damp * v[q];
    } else {
      xf[p] = v[q];
    }
  }
}

template <class T, bool ADD>
__global__ __launch_bounds__(256) void k_interp(int nxc, int nyc, int nzc,
                                                int nxf, int nyf, int nzf,
                                                const T *__restrict__ xc,
                                                const T *__restrict__ pc,
                                                T *__restrict__ xf,
                                                const double *xdamp,
                                                const double *part,
                                                unsigned nblk, XG xg) {
  interp_body<T, ADD>(nxc, nyc, nzc, nxf, nyf, nzf, xc, pc, xf, xdamp, part,
                      nblk, xg, blockIdx.x);
}

template <class T>
void interp(const Dim &c, const Dim &f, const T *xc, const T *pc, T *xf,
            const double *xdamp, hipStream_t s) {
  if (f.nx != 2 * c.nx - 1 || f.ny != 2 * c.ny - 1 || f.nz != 2 * c.nz - 1) {
    fprintf(stderr, "pmg: interp shape mismatch\n");
    return;
  }
  dim3 g(cdiv(c.nx - 1, BX), cdiv(c.ny - 1, BY), c.nz - 1);
  if (xdamp) {
    k_interp<T, true><<<xcd_blocks(g), dim3(BX, BY), 0, s>>>(c.nx, c.ny, c.nz, f.nx, f.ny,
                                                 f.nz, xc, pc, xf, xdamp,
                                                 nullptr, 0u, xg_of(g));
  } else {
    k_interp<T, false><<<xcd_blocks(g), dim3(BX, BY), 0, s>>>(c.nx, c.ny, c.nz, f.nx,
                                                  f.ny, f.nz, xc, pc, xf,
                                                  nullptr, nullptr, 0u, xg_of(g));
    if constexpr (sizeof(T) == 8) fbound00(f, (double *)xf, s);
  }
  check_launch("interp");
}

template <class T>
void interp_damped(const Dim &c, const Dim &f, const T *xc, const T *pc,
                   T *xf, const double *part, unsigned nblk, hipStream_t s) {
  if (f.nx != 2 * c.nx - 1 || f.ny != 2 * c.ny - 1 || f.nz != 2 * c.nz - 1) {
    fprintf(stderr, "pmg: interp shape mismatch\n");
    return;
  }
  dim3 g(cdiv(c.nx - 1, BX), cdiv(c.ny - 1, BY), c.nz - 1);
  k_interp<T, true><<<xcd_blocks(g), dim3(BX, BY), 0, s>>>(c.nx, c.ny, c.nz, f.nx, f.ny,
                                               f.nz, xc, pc, xf, nullptr, part,
                                               nblk, xg_of(g));
  check_launch("interp_damped");
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_nrm1(int nx, int ny, int nz,
                                              const T *__restrict__ x,
                                              Red rd) {
  __shared__ double red[4];
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k0 = 2 + blockIdx.z * KCH;
  double acc = 0.0;
  if (i <= nx - 1 && j <= ny - 1)
    for (int kk = 0; kk < KCH; kk++) {
      const int k = k0 + kk;
      if (k > nz - 1) break;
      acc += fabs((double)x[ix(nx, ny, i, j, k)]);
    }
  const double t = block_sum<256>(acc, red);
  red_store<1>(rd, t, 0.0);
}

template <class T>
void nrm1(const Dim &d, const T *x, const Red &red, hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  k_nrm1<T><<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, x, red);
  red_final<1, 0>(red, g, s);
  check_launch("nrm1");
}

// Stopping-test terms of istop 2-5 (mgcsd.cc:126-152, 497-526) on the
// interior, d = tru - u: KIND 2 sums |d| and then sets tru = u (the
// successive-iterate test); KIND 3 sums d^2; KIND 5 stores w = d (the
// energy norm's vector, for dot_energy).
template <int KIND>
__global__ __launch_bounds__(256) void k_stopdiff(int nx, int ny, int nz,
                                                  double *__restrict__ tru,
                                                  const double *__restrict__ u,
                                                  double *__restrict__ w,
                                                  Red rd) {
  __shared__ double red[4];
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k0 = 2 + blockIdx.z * KCH;
  double acc = 0.0;
  if (i <= nx - 1 && j <= ny - 1)
    for (int kk = 0; kk < KCH; kk++) {
      const int k = k0 + kk;
      if (k > nz - 1) break;
      const size_t p = ix(nx, ny, i, j, k);
      const double dv = tru[p] - u[p];
      if (KIND == 2) {
        acc += fabs(dv);
        tru[p] = u[p];
      } else if (KIND == 3) {
        acc += dv * dv;
      } else {
        w[p] = dv;
      }
    }
  if (KIND != 5) {
    const double t = block_sum<256>(acc, red);
    red_store<1>(rd, t, 0.0);
  }
}

void stop_diff(const Dim &d, int kind, double *tru, const double *u,
               double *w, const Red &red, hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  dim3 b(BX, BY);
  if (kind == 2) {
    k_stopdiff<2><<<g, b, 0, s>>>(d.nx, d.ny, d.nz, tru, u, w, red);
    red_final<1, 0>(red, g, s);
  } else if (kind == 3) {
    k_stopdiff<3><<<g, b, 0, s>>>(d.nx, d.ny, d.nz, tru, u, w, red);
    red_final<1, 0>(red, g, s);
  } else {
    k_stopdiff<5><<<g, b, 0, s>>>(d.nx, d.ny, d.nz, tru, u, w, red);
  }
  check_launch("stop_diff");
}

template <class T>
void dot_energy(const Dim &d, int numdia, const T *ac, const T *cc,
                const T *w, T *y, const Red &red, hipStream_t s) {
  const int kc = kchunk(d);
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, kc));
  dim3 b(BX, BY);
  if (numdia == 7)
    k_matvec_dots<T, 7><<<xcd_blocks(g), b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc,
                                                    w, w, y, red, 1, kc,
                                                    xg_of(g));
  else
    k_matvec_dots<T, 27><<<xcd_blocks(g), b, 0, s>>>(d.nx, d.ny, d.nz, ac, cc,
                                                     w, w, y, red, 1, kc,
                                                     xg_of(g));
  red_final<2, 2>(red, dim3(xcd_blocks(g)), s);
  check_launch("dot_energy");
}

// Vextrac (matvecd.cc:997-1008): coarse interior by injection.
__global__ __launch_bounds__(256) void k_extrac(int nxf, int nyf, int nxc,
                                                int nyc, int nzc,
                                                const double *__restrict__ xf,
                                                double *__restrict__ xc) {
  const int ii = 2 + blockIdx.x * BX + threadIdx.x;
  const int jj = 2 + blockIdx.y * BY + threadIdx.y;
  const int kk = 2 + blockIdx.z;
  if (ii > nxc - 1 || jj > nyc - 1) return;
  xc[ix(nxc, nyc, ii, jj, kk)] =
      xf[ix(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1)];
}

void extrac(const Dim &f, const Dim &c, const double *xf, double *xc,
            hipStream_t s) {
  if (c.nx < 3 || c.ny < 3 || c.nz < 3) return;
  dim3 g(cdiv(c.nx - 2, BX), cdiv(c.ny - 2, BY), c.nz - 2);
  k_extrac<<<g, dim3(BX, BY), 0, s>>>(f.nx, f.ny, c.nx, c.ny, c.nz, xf, xc);
  check_launch("extrac");
}

// ---------------------------------------------------------------------------
// Coarsest level: Vcghs (cgd.cc:60-124) in one 1024-thread workgroup.
// The search direction p lives in LDS over the whole level grid; x, r and
// A p live in registers (NPT interior points per thread).  FP64 throughout.
// ---------------------------------------------------------------------------
constexpr int CG_NT = 1024;
constexpr size_t CG_LDS_MAX = 60 * 1024;

bool cg_fits(const Dim &d) {
  return d.n() * sizeof(double) <= CG_LDS_MAX &&
         d.interior() <= (size_t)CG_NT * 2;
}

// Sum over the workgroup with a single barrier: every thread reads the NWM
// per-wavefront partial slots and adds them in the same pairwise tree, so
// all get the same value.  Slots of wavefronts the workgroup does not have
// hold zero (cg_zero_slots at the start of the kernel).  Consecutive calls
// must alternate buf (see the barrier argument in the CG loop below).
template <int NWM>
__device__ inline double cg_allsum(double v, double *buf) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) buf[threadIdx.x >> 6] = v;
  __syncthreads();
  // all per-wavefront values are requested before any is added
  double p[NWM];
#pragma unroll
  for (int w = 0; w < NWM; w++) p[w] = buf[w];
#pragma unroll
  for (int h = 1; h < NWM; h *= 2) {
#pragma unroll
    for (int w = 0; w + h < NWM; w += 2 * h) p[w] += p[w + h];
  }
  return p[0];
}

// Zero both partial buffers (before the kernel's first barrier).
template <int NWM>
__device__ inline void cg_zero_slots(double (*rbuf)[NWM]) {
  if (threadIdx.x < 2 * NWM) rbuf[threadIdx.x / NWM][threadIdx.x % NWM] = 0.0;
}

// 26 neighbour offsets in the reference's term order (in-plane 8, up 9,
// down 9; matvecd.cc:137-170) and the plane / position holding each
// coupling (upper directions at p, lower ones at the neighbour).
template <int ND>
__device__ inline void cg_load_row(const double *__restrict__ ac, size_t n,
                                   long long sy, long long sz, size_t p,
                                   double *c, long long *off) {
  const long long o[26] = {sy,           -sy,          1,
                           -1,           1 + sy,       -1 + sy,
                           1 - sy,       -1 - sy,      sz,
                           sy + sz,      -sy + sz,     1 + sz,
                           -1 + sz,      1 + sy + sz,  -1 + sy + sz,
                           1 - sy + sz,  -1 - sy + sz, -sz,
                           sy - sz,      -sy - sz,     1 - sz,
                           -1 - sz,      1 + sy - sz,  -1 + sy - sz,
                           1 - sy - sz,  -1 - sy - sz};
  // plane of each term, and whether it is stored at the neighbour
  const int pl[26] = {2, 2, 1, 1, 4, 5, 5, 4, 3, 8, 9, 6, 7,
                      10, 11, 12, 13, 3, 9, 8, 7, 6, 13, 12, 11, 10};
  const bool atn[26] = {0, 1, 0, 1, 0, 0, 1, 1, 0, 0, 0, 0, 0,
                        0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
#pragma unroll
  for (int q = 0; q < 26; q++) {
    off[q] = o[q];
    const bool use = ND == 27 || pl[q] <= 3;
    c[q] = use ? ac[(size_t)pl[q] * n +
                    (atn[q] ? (size_t)((long long)p + o[q]) : p)]
               : 0.0;
  }
}

template <class T, int NPT, int ND>
__global__ __launch_bounds__(CG_NT) void k_cg_single(
    int nx, int ny, int nz, const double *__restrict__ ac,
    const double *__restrict__ cc, const T *__restrict__ f,
    T *__restrict__ xout, int itmax, double errtol) {
  extern __shared__ double P[];
  __shared__ double rbuf[2][CG_NT / 64];
  cg_zero_slots<CG_NT / 64>(rbuf);
  const int tid = threadIdx.x, NTB = blockDim.x;
  const size_t n = (size_t)nx * ny * nz;
  const long long sy = nx, sz = (long long)nx * ny;
  const int ni = nx - 2, nj = ny - 2;
  const int nint = ni * nj * (nz - 2);
  for (size_t q = tid; q < n; q += NTB) P[q] = 0.0;
  // per point: LDS indices of the point and of its 26 neighbours (the
  // level fits the LDS, so 32-bit), coefficients in registers
  int pidx[NPT], nb[NPT == 1 ? 26 : 1][26];  // indices kept for NPT 1
  bool val[NPT];
  double xr[NPT], rr[NPT], apv[NPT], dg[NPT];
  double cf[NPT][26];
  long long off[26];
#pragma unroll
  for (int m = 0; m < NPT; m++) {
    const int t = tid + m * NTB;
    val[m] = t < nint;
    const int tt = val[m] ? t : 0;
    const int i = 2 + tt % ni, j = 2 + (tt / ni) % nj, k = 2 + tt / (ni * nj);
    const size_t p = ix(nx, ny, i, j, k);
    pidx[m] = (int)p;
    xr[m] = 0.0;
    rr[m] = val[m] ? (double)f[p] : 0.0;  // r = f - A*0
    apv[m] = 0.0;
    dg[m] = ac[p] + cc[p];
    cg_load_row<ND>(ac, n, sy, sz, p, cf[m], off);
    if constexpr (NPT == 1) {
#pragma unroll
      for (int q = 0; q < 26; q++) nb[0][q] = (int)((long long)p + off[q]);
    }
  }
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < NPT; m++) s += rr[m] * rr[m];
  double rhok2 = cg_allsum<CG_NT / 64>(s, rbuf[0]);
  int b = 1;
  const double denom = sqrt(rhok2);
  int iters = 0;
  double rhok1 = 0.0;
  if (denom != 0.0) {
    for (;;) {
      // stopping test sqrt(rhok2)/denom <= errtol; for errtol = 0 (the
      // multigrid coarse solve) that is rhok2 == 0, without the divide
      if (errtol == 0.0 ? rhok2 == 0.0 : sqrt(rhok2) / denom <= errtol)
        break;
      if (iters >= itmax) break;
      // the previous reduction's barrier ordered every read of P before
      // these writes
      if (iters == 0) {
#pragma unroll
        for (int m = 0; m < NPT; m++)
          if (val[m]) P[pidx[m]] = rr[m];
      } else {
        const double beta = rhok2 / rhok1;
        const double ib = 1.0 / beta;
#pragma unroll
        for (int m = 0; m < NPT; m++)
          if (val[m]) P[pidx[m]] = (P[pidx[m]] + ib * rr[m]) * beta;
      }
      __syncthreads();
      s = 0.0;
#pragma unroll
      for (int m = 0; m < NPT; m++) {
        auto NB = [&](int q) -> int {
          if constexpr (NPT == 1) return nb[0][q];
          else return pidx[m] + (int)off[q];
        };
        double o = 0.0, u = 0.0, dd = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) o += cf[m][q] * P[NB(q)];
#pragma unroll
        for (int q = 8; q < 17; q++) u += cf[m][q] * P[NB(q)];
#pragma unroll
        for (int q = 17; q < 26; q++) dd += cf[m][q] * P[NB(q)];
        const double pp = P[pidx[m]];
        const double v = -o + -u + -dd + dg[m] * pp;
        apv[m] = val[m] ? v : 0.0;
        if (val[m]) s += pp * v;
      }
      const double pAp = cg_allsum<CG_NT / 64>(s, rbuf[b]);
      b ^= 1;
      const double alpha = rhok2 / pAp;
      rhok1 = rhok2;
      s = 0.0;
#pragma unroll
      for (int m = 0; m < NPT; m++) {
        if (!val[m]) continue;
        xr[m] += alpha * P[pidx[m]];
        rr[m] += -alpha * apv[m];
        s += rr[m] * rr[m];
      }
      iters++;
      rhok2 = cg_allsum<CG_NT / 64>(s, rbuf[b]);
      b ^= 1;
    }
  }
  // x = 0 on the boundary (Vazeros before the solve), solution inside
  for (size_t q = tid; q < n; q += NTB) {
    const int i = 1 + (int)(q % nx), j = 1 + (int)((q / nx) % ny),
              k = 1 + (int)(q / ((size_t)nx * ny));
    if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz)
      xout[q] = (T)0;
  }
#pragma unroll
  for (int m = 0; m < NPT; m++)
    if (val[m]) xout[pidx[m]] = (T)xr[m];
}

// Row-segment variant for 27-point levels: each thread owns CG_R
// consecutive points of an i-row.  Per iteration it reads the R+2 values of
// each of the 9 neighbouring rows once from LDS (15 reads per point instead
// of 27) and keeps the 26 couplings of its points in registers; fewer
// wavefronts make the per-iteration reductions and barriers cheaper.  Terms
// are summed in the reference's three groups (in-plane, up, down), each
// group row by row.
constexpr int CG_R = 3, CG_RNT = 512;
// WARM: start from xout (else 0); ABS: the class solver's absolute
// criterion (else Vcghs's relative one); ZT: errtol is 0, so Vcghs's test
// sqrt(r.r)/sqrt(r0.r0) <= errtol is r.r == 0 (no square root or divide in
// the loop).  The PMG cycle uses <T, 0, 0, 1> (errtol = epsiln = 0).
extern __shared__ double cg_lds[];
template <class T, bool WARM, bool ABS, bool ZT>
__device__ __forceinline__ void cg_rows_body(
    int nx, int ny, int nz, const double *__restrict__ ac,
    const double *__restrict__ cc, const T *__restrict__ f,
    T *__restrict__ xout, int itmax, double errtol) {
  constexpr int R = CG_R;
  double *P = cg_lds;
  __shared__ double rbuf[2][CG_RNT / 64];
  cg_zero_slots<CG_RNT / 64>(rbuf);
  const int tid = threadIdx.x, NTB = blockDim.x;
  const size_t n = (size_t)nx * ny * nz;
  const long long sy = nx, sz = (long long)nx * ny;
  const int ni = nx - 2, nj = ny - 2, nk = nz - 2;
  const int nseg = (ni + R - 1) / R;
  const bool act = tid < nseg * nj * nk;
  const int tt = act ? tid : 0;
  const int j = 2 + (tt / nseg) % nj, k = 2 + tt / (nseg * nj);
  const int i0 = 2 + (tt % nseg) * R;
  // n + R + 2 entries: a tail segment may read a few values past the grid
  for (size_t q = tid; q < n + R + 2; q += NTB) P[q] = 0.0;
  int pidx[R];
  bool val[R];
  double xr[R], rr[R], apv[R], dg[R];
  double cf[R][26];
  long long off[26];
#pragma unroll
  for (int r = 0; r < R; r++) {
    val[r] = act && i0 + r <= nx - 1;
    const size_t p = ix(nx, ny, val[r] ? i0 + r : 2, j, k);
    pidx[r] = (int)p;
    xr[r] = WARM && val[r] ? (double)xout[p] : 0.0;
    rr[r] = val[r] ? (double)f[p] : 0.0;  // r = f - A x (x = 0 unless warm)
    apv[r] = 0.0;
    dg[r] = val[r] ? ac[p] + cc[p] : 0.0;
    cg_load_row<27>(ac, n, sy, sz, p, cf[r], off);
    if (!val[r]) {
#pragma unroll
      for (int q = 0; q < 26; q++) cf[r][q] = 0.0;
    }
  }
  // LDS index of (i0-1, j+dj, k+dk), rows in (dk, dj) order
  int rb[9];
#pragma unroll
  for (int a = 0; a < 9; a++)
    rb[a] = (int)ix(nx, ny, i0 - 1, j + (a % 3) - 1, k + (a / 3) - 1);
  // offset q of cg_load_row -> (di, dj, dk)
  constexpr signed char DI[26] = {0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 0, 1, -1,
                                  1, -1, 1, -1, 0, 0, 0, 1, -1, 1, -1, 1, -1};
  constexpr signed char DJ[26] = {1, -1, 0, 0, 1, 1, -1, -1, 0, 1, -1, 0, 0,
                                  1, 1, -1, -1, 0, 1, -1, 0, 0, 1, 1, -1, -1};
  constexpr signed char DK[26] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1,
                                  1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  __syncthreads();
  // With WARM the first pass through the loop computes r = f - A x0 with
  // the loop's own stencil code (P = x0), then the CG proper starts.
  bool warm_pass = WARM;
  int b = 0;
  double rhok2 = 0.0, denom = 0.0;
  double s = 0.0;
  if constexpr (!WARM) {
#pragma unroll
    for (int r = 0; r < R; r++) s += rr[r] * rr[r];
    rhok2 = cg_allsum<CG_RNT / 64>(s, rbuf[0]);
    b = 1;
    denom = sqrt(rhok2);
  }
  int iters = 0;
  double rhok1 = 0.0;
  {
    for (;;) {
      if (!warm_pass) {
        if (denom == 0.0) break;
        // ABS: the class solver's criterion, sqrt(r.r) > errtol, its first
        // test made on sqrt(sqrt(r0.r0)) (CG_Solver.h:125-133)
        if (ABS ? (iters == 0 ? sqrt(sqrt(rhok2)) : sqrt(rhok2)) <= errtol
                : (ZT ? rhok2 == 0.0 : sqrt(rhok2) / denom <= errtol))
          break;
        if (iters >= itmax) break;
        if (iters == 0) {
#pragma unroll
          for (int r = 0; r < R; r++)
            if (val[r]) P[pidx[r]] = rr[r];
        } else {
          const double beta = rhok2 / rhok1;
          const double ib = 1.0 / beta;
#pragma unroll
          for (int r = 0; r < R; r++)
            if (val[r]) P[pidx[r]] = (P[pidx[r]] + ib * rr[r]) * beta;
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; r++)
          if (val[r]) P[pidx[r]] = xr[r];
      }
      __syncthreads();
      double o[R], u[R], dd[R], pp[R];
#pragma unroll
      for (int r = 0; r < R; r++) o[r] = u[r] = dd[r] = 0.0;
#pragma unroll
      for (int a = 0; a < 9; a++) {
        const int dj = (a % 3) - 1, dk = (a / 3) - 1;
        double w[R + 2];
#pragma unroll
        for (int c = 0; c < R + 2; c++) w[c] = P[rb[a] + c];
        if (dj == 0 && dk == 0) {
#pragma unroll
          for (int r = 0; r < R; r++) pp[r] = w[r + 1];
        }
#pragma unroll
        for (int q = 0; q < 26; q++) {
          if (DJ[q] != dj || DK[q] != dk) continue;
#pragma unroll
          for (int r = 0; r < R; r++) {
            const double t = cf[r][q] * w[r + 1 + DI[q]];
            if (DK[q] == 0) o[r] += t;
            else if (DK[q] > 0) u[r] += t;
            else dd[r] += t;
          }
        }
      }
      s = 0.0;
      if (warm_pass) {  // r = f - A x0
#pragma unroll
        for (int r = 0; r < R; r++) {
          const double v = -o[r] + -u[r] + -dd[r] + dg[r] * pp[r];
          if (val[r]) {
            rr[r] -= v;
            s += rr[r] * rr[r];
          }
        }
        rhok2 = cg_allsum<CG_RNT / 64>(s, rbuf[b]);
        b ^= 1;
        denom = sqrt(rhok2);
        warm_pass = false;
        continue;
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const double v = -o[r] + -u[r] + -dd[r] + dg[r] * pp[r];
        apv[r] = val[r] ? v : 0.0;
        if (val[r]) s += pp[r] * v;
      }
      const double pAp = cg_allsum<CG_RNT / 64>(s, rbuf[b]);
      b ^= 1;
      const double alpha = rhok2 / pAp;
      rhok1 = rhok2;
      s = 0.0;
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (!val[r]) continue;
        xr[r] += alpha * pp[r];
        rr[r] += -alpha * apv[r];
        s += rr[r] * rr[r];
      }
      iters++;
      rhok2 = cg_allsum<CG_RNT / 64>(s, rbuf[b]);
      b ^= 1;
    }
  }
  for (size_t q = tid; q < n; q += NTB) {
    const int i = 1 + (int)(q % nx), jj = 1 + (int)((q / nx) % ny),
              kk = 1 + (int)(q / ((size_t)nx * ny));
    if (i == 1 || i == nx || jj == 1 || jj == ny || kk == 1 || kk == nz)
      xout[q] = (T)0;
  }
#pragma unroll
  for (int r = 0; r < R; r++)
    if (val[r]) xout[pidx[r]] = (T)xr[r];
}

template <class T, bool WARM, bool ABS, bool ZT>
__global__ __launch_bounds__(CG_RNT) void k_cg_rows(
    int nx, int ny, int nz, const double *__restrict__ ac,
    const double *__restrict__ cc, const T *__restrict__ f,
    T *__restrict__ xout, int itmax, double errtol) {
  cg_rows_body<T, WARM, ABS, ZT>(nx, ny, nz, ac, cc, f, xout, itmax, errtol);
}

template <class T, int ND>
static void cg_launch(const Dim &d, const double *ac, const double *cc,
                      const T *f, T *x, int itmax, double errtol,
                      hipStream_t s) {
  const size_t lds = d.n() * sizeof(double);
  const size_t ni = d.interior();
  const int nrow = (int)((d.nx - 2 + CG_R - 1) / CG_R) * (d.ny - 2) * (d.nz - 2);
  if (ND == 27 && nrow <= CG_RNT) {
    const int nt = (nrow + 63) / 64 * 64;
    if (errtol == 0.0)
      k_cg_rows<T, false, false, true>
          <<<1, nt, lds + (CG_R + 2) * sizeof(double), s>>>(
              d.nx, d.ny, d.nz, ac, cc, f, x, itmax, errtol);
    else
      k_cg_rows<T, false, false, false>
          <<<1, nt, lds + (CG_R + 2) * sizeof(double), s>>>(
              d.nx, d.ny, d.nz, ac, cc, f, x, itmax, errtol);
    return;
  }
  // one interior point per thread when they fit: as few wavefronts as the
  // level needs (the loop is LDS- and barrier-bound, idle waves cost)
  if (ni <= CG_NT) {
    const int nt = (int)((ni + 63) / 64) * 64;
    k_cg_single<T, 1, ND><<<1, nt, lds, s>>>(d.nx, d.ny, d.nz, ac, cc, f, x,
                                             itmax, errtol);
  } else {
    k_cg_single<T, 2, ND><<<1, CG_NT, lds, s>>>(d.nx, d.ny, d.nz, ac, cc, f,
                                                x, itmax, errtol);
  }
}

template <class T>
void cg_single(const Dim &d, int numdia, const double *ac, const double *cc,
               const T *f, T *x, int itmax, double errtol, hipStream_t s) {
  if (!cg_fits(d)) {
    fprintf(stderr, "pmg: coarsest level too large for cg_single\n");
    return;
  }
  if (numdia == 7)
    cg_launch<T, 7>(d, ac, cc, f, x, itmax, errtol, s);
  else
    cg_launch<T, 27>(d, ac, cc, f, x, itmax, errtol, s);
  check_launch("cg_single");
}

bool cg_class_fits(const Dim &d, int numdia) {
  const int nrow = (int)((d.nx - 2 + CG_R - 1) / CG_R) * (d.ny - 2) * (d.nz - 2);
  return numdia == 27 && d.n() * sizeof(double) <= CG_LDS_MAX &&
         nrow <= CG_RNT;
}

template <class T>
void cg_class(const Dim &d, int numdia, const double *ac, const double *cc,
              const T *f, T *x, int itmax, double thresh, hipStream_t s) {
  if (!cg_class_fits(d, numdia)) {
    fprintf(stderr, "pmg: coarsest level too large for cg_class\n");
    return;
  }
  const int nrow = (int)((d.nx - 2 + CG_R - 1) / CG_R) * (d.ny - 2) * (d.nz - 2);
  const int nt = (nrow + 63) / 64 * 64;
  k_cg_rows<T, true, true, false>
      <<<1, nt, d.n() * sizeof(double) + (CG_R + 2) * sizeof(double), s>>>(
          d.nx, d.ny, d.nz, ac, cc, f, x, itmax, thresh);
  check_launch("cg_class");
}

// ---------------------------------------------------------------------------
// Class-style PCG (CG_Solver.h:152-237): dot product and the two vector
// updates, their coefficients read from device scalars so the loop needs
// no host round trip.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dot(int nx, int ny, int nz,
                                             const double *__restrict__ x,
                                             const double *__restrict__ y,
                                             Red rd) {
  __shared__ double red[4];
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k0 = 2 + blockIdx.z * KCH;
  double acc = 0.0;
  if (i <= nx - 1 && j <= ny - 1)
    for (int kk = 0; kk < KCH; kk++) {
      const int k = k0 + kk;
      if (k > nz - 1) break;
      const size_t p = ix(nx, ny, i, j, k);
      acc += x[p] * y[p];
    }
  const double t = block_sum<256>(acc, red);
  red_store<1>(rd, t, 0.0);
}

void dot_interior(const Dim &d, const double *x, const double *y,
                  const Red &red, hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  k_dot<<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, x, y, red);
  red_final<1, 0>(red, g, s);
  check_launch("dot_interior");
}

// x += a p, r -= a q with a = rs / pap (0 when pap == 0), interior
__global__ __launch_bounds__(256) void k_pcg_xr(
    int nx, int ny, int nz, double *__restrict__ x, double *__restrict__ r,
    const double *__restrict__ p, const double *__restrict__ q,
    const double *__restrict__ rs, const double *__restrict__ pap) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > nx - 1 || j > ny - 1) return;
  const double a = pap[0] == 0.0 ? 0.0 : rs[0] / pap[0];
  const size_t o = ix(nx, ny, i, j, k);
  x[o] += a * p[o];
  r[o] -= a * q[o];
}

// p = s + b p with b = rsn / rs (0 when rs == 0), interior
__global__ __launch_bounds__(256) void k_pcg_p(
    int nx, int ny, int nz, double *__restrict__ p,
    const double *__restrict__ sv, const double *__restrict__ rsn,
    const double *__restrict__ rs) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > nx - 1 || j > ny - 1) return;
  const double b = rs[0] == 0.0 ? 0.0 : rsn[0] / rs[0];
  const size_t o = ix(nx, ny, i, j, k);
  p[o] = sv[o] + b * p[o];
}

void pcg_xr(const Dim &d, double *x, double *r, const double *p,
            const double *q, const double *rs, const double *pap,
            hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), d.nz - 2);
  k_pcg_xr<<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, x, r, p, q, rs, pap);
  check_launch("pcg_xr");
}

void pcg_p(const Dim &d, double *p, const double *sv, const double *rsn,
           const double *rs, hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), d.nz - 2);
  k_pcg_p<<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, p, sv, rsn, rs);
  check_launch("pcg_p");
}

// ---------------------------------------------------------------------------
// BLAS-1 helpers
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_diag(const T *__restrict__ ac, const T *__restrict__ cc,
                       T *__restrict__ dg, size_t n) {
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x)
    dg[p] = ac[p] + cc[p];
}
template <class T>
void diag_sum(const T *ac, const T *cc, T *dg, size_t n, hipStream_t s) {
  const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  k_diag<T><<<nb, 256, 0, s>>>(ac, cc, dg, n);
  check_launch("diag_sum");
}

template <class T>
__global__ void k_fill(T *x, size_t n, T v) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    x[t] = v;
}
template <class T>
void fill(T *x, size_t n, T v, hipStream_t s) {
  if (!n) return;
  int nb = (int)((n + 255) / 256);
  if (nb > 16384) nb = 16384;
  k_fill<T><<<nb, 256, 0, s>>>(x, n, v);
  check_launch("fill");
}

template <class TA, class TB>
__global__ void k_convert(const TA *a, TB *b, size_t n) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    b[t] = (TB)a[t];
}
template <class TA, class TB>
void convert(const TA *a, TB *b, size_t n, hipStream_t s) {
  if (!n) return;
  int nb = (int)((n + 255) / 256);
  if (nb > 16384) nb = 16384;
  k_convert<TA, TB><<<nb, 256, 0, s>>>(a, b, n);
  check_launch("convert");
}

template <class T>
__global__ __launch_bounds__(256) void k_axpy(int nx, int ny, int nz,
                                              double alpha,
                                              const T *__restrict__ x,
                                              T *__restrict__ y) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k0 = 2 + blockIdx.z * KCH;
  if (i > nx - 1 || j > ny - 1) return;
  for (int kk = 0; kk < KCH; kk++) {
    const int k = k0 + kk;
    if (k > nz - 1) break;
    const size_t p = ix(nx, ny, i, j, k);
    y[p] += (T)alpha * x[p];
  }
}
template <class T>
void axpy_interior(const Dim &d, double alpha, const T *x, T *y,
                   hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  k_axpy<T><<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, alpha, x, y);
  check_launch("axpy");
}

template <class T>
__device__ __forceinline__ void addcorr_body(int nx, int ny, int nz,
                                             const T *__restrict__ e,
                                             double *__restrict__ u, int bx,
                                             int by, int bz) {
  const int i = 2 + bx * BX + threadIdx.x;
  const int j = 2 + by * BY + threadIdx.y;
  const int k0 = 2 + bz * KCH;
  if (i > nx - 1 || j > ny - 1) return;
  for (int kk = 0; kk < KCH; kk++) {
    const int k = k0 + kk;
    if (k > nz - 1) break;
    const size_t p = ix(nx, ny, i, j, k);
    u[p] += (double)e[p];
  }
}
template <class T>
__global__ __launch_bounds__(256) void k_addcorr(int nx, int ny, int nz,
                                                 const T *__restrict__ e,
                                                 double *__restrict__ u) {
  addcorr_body<T>(nx, ny, nz, e, u, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <class T>
void add_correction(const Dim &d, const T *e, double *u, hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  k_addcorr<T><<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, e, u);
  check_launch("add_correction");
}

__global__ __launch_bounds__(256) void k_xpay(int nx, int ny, int nz,
                                              const double *__restrict__ x,
                                              double damp,
                                              const double *__restrict__ dx,
                                              double *__restrict__ out) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k0 = 2 + blockIdx.z * KCH;
  if (i > nx - 1 || j > ny - 1) return;
  for (int kk = 0; kk < KCH; kk++) {
    const int k = k0 + kk;
    if (k > nz - 1) break;
    const size_t p = ix(nx, ny, i, j, k);
    double v = x[p];
    v += damp * dx[p];
    out[p] = v;
  }
}
void xpay_interior(const Dim &d, const double *x, double damp,
                   const double *dx, double *out, hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  k_xpay<<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, x, damp, dx, out);
  check_launch("xpay");
}

__global__ __launch_bounds__(256) void k_copyint(int nx, int ny, int nz,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y) {
  const int i = 2 + blockIdx.x * BX + threadIdx.x;
  const int j = 2 + blockIdx.y * BY + threadIdx.y;
  const int k0 = 2 + blockIdx.z * KCH;
  if (i > nx - 1 || j > ny - 1) return;
  for (int kk = 0; kk < KCH; kk++) {
    const int k = k0 + kk;
    if (k > nz - 1) break;
    const size_t p = ix(nx, ny, i, j, k);
    y[p] = x[p];
  }
}
void copy_interior(const Dim &d, const double *x, double *y, hipStream_t s) {
  dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, KCH));
  k_copyint<<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, x, y);
  check_launch("copy_interior");
}

// Faces written i, then j, then k (VfboundPMG, mikpckd.cc:266-336, on the
// faces it intends);
// three launches keep the edge precedence of the reference.
__global__ void k_face(int nx, int ny, int nz, int axis, double *x,
                       const double *g) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  int na = axis == 0 ? ny : nx;
  if (a >= na) return;
  for (int side = 0; side < 2; side++) {
    int i, j, k;
    if (axis == 0) {
      i = side ? nx : 1, j = a + 1, k = b + 1;
    } else if (axis == 1) {
      i = a + 1, j = side ? ny : 1, k = b + 1;
    } else {
      i = a + 1, j = b + 1, k = side ? nz : 1;
    }
    size_t plane = axis == 0 ? (size_t)ny * nz
                             : (axis == 1 ? (size_t)nx * nz : (size_t)nx * ny);
    size_t gi = (size_t)a + (size_t)b * na + (size_t)side * plane;
    x[ix(nx, ny, i, j, k)] = g ? g[gi] : 0.0;
  }
}
void fbound_dirichlet(const Dim &d, double *x, const double *gx,
                      const double *gy, const double *gz, hipStream_t s) {
  k_face<<<dim3(cdiv(d.ny, 64), d.nz), 64, 0, s>>>(d.nx, d.ny, d.nz, 0, x, gx);
  k_face<<<dim3(cdiv(d.nx, 64), d.nz), 64, 0, s>>>(d.nx, d.ny, d.nz, 1, x, gy);
  k_face<<<dim3(cdiv(d.nx, 64), d.ny), 64, 0, s>>>(d.nx, d.ny, d.nz, 2, x, gz);
  check_launch("fbound");
}
void fbound00(const Dim &d, double *x, hipStream_t s) {
  fbound_dirichlet(d, x, nullptr, nullptr, nullptr, s);
}

// ---------------------------------------------------------------------------
// Setup: VbuildA_fv (buildAd.cc:62-212)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_buildA(
    int nx, int ny, int nz, double *__restrict__ ac, double *__restrict__ cc,
    double *__restrict__ fc, const double *__restrict__ xf,
    const double *__restrict__ yf, const double *__restrict__ zf,
    const double *__restrict__ gx, const double *__restrict__ gy,
    const double *__restrict__ gz, const double *__restrict__ a1,
    const double *__restrict__ a2, const double *__restrict__ a3,
    const double *__restrict__ ccf, const double *__restrict__ fcf,
    double *__restrict__ dg, float *__restrict__ ac32,
    float *__restrict__ cc32, float *__restrict__ dg32) {
  // no multiply-add contraction: the operator is the oracle's to the bit
#pragma clang fp contract(off)
  // BX*BY consecutive interior points of plane k, i fastest across row ends
  const int q = blockIdx.x * (BX * BY) + threadIdx.y * BX + threadIdx.x;
  if (q >= (nx - 2) * (ny - 2)) return;
  const int i = 2 + q % (nx - 2);
  const int j = 2 + q / (nx - 2);
  const int k = 2 + blockIdx.z;
  const size_t n = (size_t)nx * ny * nz;
  double *oC = ac, *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  const size_t p = ix(nx, ny, i, j, k);
  const double hzm1 = zf[k - 1] - zf[k - 2], hz = zf[k] - zf[k - 1];
  const double hym1 = yf[j - 1] - yf[j - 2], hy = yf[j] - yf[j - 1];
  const double hxm1 = xf[i - 1] - xf[i - 2], hx = xf[i] - xf[i - 1];
  const double cE = (hym1 + hy) * (hzm1 + hz) / (4.0 * hx);
  const double cEm = (hym1 + hy) * (hzm1 + hz) / (4.0 * hxm1);
  const double cN = (hxm1 + hx) * (hzm1 + hz) / (4.0 * hy);
  const double cNm = (hxm1 + hx) * (hzm1 + hz) / (4.0 * hym1);
  const double cU = (hxm1 + hx) * (hym1 + hy) / (4.0 * hz);
  const double cUm = (hxm1 + hx) * (hym1 + hy) / (4.0 * hzm1);
  const double cF = (hxm1 + hx) * (hym1 + hy) * (hzm1 + hz) / 8.0;
  const size_t pxm = p - 1, pym = p - nx, pzm = p - (size_t)nx * ny;
  double f = cF * fcf[p];
  const double vcc = cF * ccf[p];
  const double vC = cE * a1[p] + cEm * a1[pxm] + cN * a2[p] + cNm * a2[pym] +
                    cU * a3[p] + cUm * a3[pzm];
  cc[p] = vcc;
  oC[p] = vC;
  dg[p] = vC + vcc;  // the sweeps' and residuals' diagonal oC + cc
  const int ike = (i == nx - 1) ? 0 : 1;
  const double vE = ike * cE * a1[p];
  oE[p] = vE;
  f += (1 - ike) * cE * a1[p] * gx[(j - 1) + (size_t)(k - 1) * ny +
                                   (size_t)ny * nz];
  const int jke = (j == ny - 1) ? 0 : 1;
  const double vN = jke * cN * a2[p];
  oN[p] = vN;
  f += (1 - jke) * cN * a2[p] * gy[(i - 1) + (size_t)(k - 1) * nx +
                                   (size_t)nx * nz];
  const int kke = (k == nz - 1) ? 0 : 1;
  const double vU = kke * cU * a3[p];
  uC[p] = vU;
  f += (1 - kke) * cU * a3[p] * gz[(i - 1) + (size_t)(j - 1) * nx +
                                   (size_t)nx * ny];
  const int ikw = (i == 2) ? 0 : 1;
  f += (1 - ikw) * cEm * a1[pxm] * gx[(j - 1) + (size_t)(k - 1) * ny];
  const int jks = (j == 2) ? 0 : 1;
  f += (1 - jks) * cNm * a2[pym] * gy[(i - 1) + (size_t)(k - 1) * nx];
  const int kkd = (k == 2) ? 0 : 1;
  f += (1 - kkd) * cUm * a3[pzm] * gz[(i - 1) + (size_t)(j - 1) * nx];
  fc[p] = f;
  if (ac32) {  // the FP32 copies, rounded as a separate conversion would
    cc32[p] = (float)vcc;
    ac32[p] = (float)vC;
    dg32[p] = (float)vC + (float)vcc;  // the sum of the FP32 copies
    ac32[n + p] = (float)vE;
    ac32[2 * n + p] = (float)vN;
    ac32[3 * n + p] = (float)vU;
  }
}

void buildA(const Dim &d, double *ac, double *cc, double *fc,
            const double *xf, const double *yf, const double *zf,
            const double *gx, const double *gy, const double *gz,
            const double *a1, const double *a2, const double *a3,
            const double *ccf, const double *fcf, hipStream_t s,
            double *dg, float *ac32, float *cc32, float *dg32) {
  dim3 g(cdiv((d.nx - 2) * (d.ny - 2), BX * BY), 1, d.nz - 2);
  k_buildA<<<g, dim3(BX, BY), 0, s>>>(d.nx, d.ny, d.nz, ac, cc, fc, xf, yf,
                                      zf, gx, gy, gz, a1, a2, a3, ccf, fcf,
                                      dg, ac32, cc32, dg32);
  check_launch("buildA");
}

// ---------------------------------------------------------------------------
// Setup: operator-based prolongation (VbuildPb_op7 / op27 / trilin)
// ---------------------------------------------------------------------------
// Symmetric 27-point storage: plane holding the coupling between p and p+e
// for an upper direction e; -1 otherwise (plane order matvecd.cc:185-189).
__host__ __device__ constexpr int upper_plane(int ex, int ey, int ez) {
  return ez == 1
             ? (ey == -1 ? (ex == -1 ? 13 : (ex == 0 ? 9 : 12))
                         : (ey == 0 ? (ex == -1 ? 7 : (ex == 0 ? 3 : 6))
                                    : (ex == -1 ? 11 : (ex == 0 ? 8 : 10))))
         : (ez == 0 && ey == 1) ? (ex == 1 ? 4 : (ex == 0 ? 2 : 5))
         : (ez == 0 && ey == 0 && ex == 1) ? 1
                                           : -1;
}

template <int ND>
__device__ inline double coupling(const double *__restrict__ ac, size_t n,
                                  size_t p, long long off, int ex, int ey,
                                  int ez) {
  int pl = upper_plane(ex, ey, ez);
  size_t at = p;
  if (pl < 0) {
    pl = upper_plane(-ex, -ey, -ez);
    at = (size_t)((long long)p + off);
  }
  if (ND == 7 && pl > 3) return 0.0;
  return ac[(size_t)pl * n + at];
}

__host__ __device__ constexpr int pplane(int dx, int dy, int dz) {
  return 9 * (dz == 0 ? 0 : (dz == 1 ? 1 : 2)) +
         (dy == -1 ? (dx == -1 ? 8 : (dx == 0 ? 2 : 7))
                   : (dy == 0 ? (dx == -1 ? 4 : (dx == 0 ? 0 : 3))
                              : (dx == -1 ? 6 : (dx == 0 ? 1 : 5))));
}

// Weights of the collapsed-operator rule: for F = C_f + d, axes with d_a = 0
// are collapsed into the diagonal; neighbours towards C carry the known
// weight of their own offset, neighbours away from C weight 0.
template <int ND>
__global__ __launch_bounds__(128) void k_buildP(int nxf, int nyf, int nzf,
                                                int nxc, int nyc, int nzc,
                                                const double *__restrict__ ac,
                                                double *__restrict__ pc,
                                                float *__restrict__ pc32,
                                                int mgprol, XG xg) {
  // no multiply-add contraction: the weights divide differences of
  // couplings that cancel where the operator's row sums vanish (no ions),
  // so they match the oracle only with the same roundings
#pragma clang fp contract(off)
  int bx, by, bz;
  if (!xcd_tile(xg, bx, by, bz)) return;
  // 128 consecutive interior points of coarse plane kk, i fastest across
  // row ends, so that short coarse rows leave no lanes idle
  const int q = bx * 128 + threadIdx.y * 64 + threadIdx.x;
  if (q >= (nxc - 2) * (nyc - 2)) return;
  const int ii = 2 + q % (nxc - 2);
  const int jj = 2 + q / (nxc - 2);
  const int kk = 2 + bz;
  const size_t nf = (size_t)nxf * nyf * nzf, nc = (size_t)nxc * nyc * nzc;
  const size_t pC = ix(nxc, nyc, ii, jj, kk);
  const size_t pCf = ix(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1);
  const long long sx = 1, sy = nxf, sz = (long long)nxf * nyf;
  double w[27];
  w[13] = 1.0;  // index (dz+1)*9 + (dy+1)*3 + (dx+1)
  if (mgprol == 0) {
#pragma unroll
    for (int a = 0; a < 27; a++) {
      const int dx = a % 3 - 1, dy = (a / 3) % 3 - 1, dz = a / 9 - 1;
      const int m = (dx != 0) + (dy != 0) + (dz != 0);
      w[a] = 1.0 / (double)(1 << m);
    }
  } else {
#pragma unroll
    for (int nnz = 1; nnz <= 3; nnz++) {
#pragma unroll
      for (int a = 0; a < 27; a++) {
        const int dx = a % 3 - 1, dy = (a / 3) % 3 - 1, dz = a / 9 - 1;
        if ((dx != 0) + (dy != 0) + (dz != 0) != nnz) continue;
        const size_t pF = (size_t)((long long)pCf + dx * sx + dy * sy + dz * sz);
        double den = ac[pF], num = 0.0;
#pragma unroll
        for (int b = 0; b < 27; b++) {
          const int ex = b % 3 - 1, ey = (b / 3) % 3 - 1, ez = b / 9 - 1;
          if (b == 13) continue;
          if (ND == 7 && (ex != 0) + (ey != 0) + (ez != 0) != 1) continue;
          const int d3[3] = {dx, dy, dz}, e3[3] = {ex, ey, ez};
          bool ok = true, toward = false;
          int t3[3];
#pragma unroll
          for (int q = 0; q < 3; q++) {
            if (d3[q] == 0) {
              t3[q] = 0;
            } else if (e3[q] == 0) {
              t3[q] = d3[q];
            } else if (e3[q] == -d3[q]) {
              t3[q] = 0;
              toward = true;
            } else {
              ok = false;
            }
          }
          if (!ok) continue;
          const double cpl = coupling<ND>(ac, nf, pF, ex * sx + ey * sy + ez * sz,
                                          ex, ey, ez);
          if (!toward)
            den -= cpl;
          else
            num += cpl * w[(t3[2] + 1) * 9 + (t3[1] + 1) * 3 + (t3[0] + 1)];
        }
        w[a] = num / den;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 27; a++) {
    const int dx = a % 3 - 1, dy = (a / 3) % 3 - 1, dz = a / 9 - 1;
    pc[(size_t)pplane(dx, dy, dz) * nc + pC] = w[a];
    if (pc32) pc32[(size_t)pplane(dx, dy, dz) * nc + pC] = (float)w[a];
  }
}

void buildP(const Dim &f, const Dim &c, int mgprol, int numdia,
            const double *acf, double *pc, hipStream_t s, float *pc32) {
  if (c.nx < 3 || c.ny < 3 || c.nz < 3) return;
  dim3 g(cdiv((c.nx - 2) * (c.ny - 2), 128), 1, c.nz - 2);
  if (numdia == 7)
    k_buildP<7><<<xcd_blocks(g), dim3(64, 2), 0, s>>>(f.nx, f.ny, f.nz, c.nx, c.ny, c.nz,
                                          acf, pc, pc32, mgprol, xg_of(g));
  else
    k_buildP<27><<<xcd_blocks(g), dim3(64, 2), 0, s>>>(f.nx, f.ny, f.nz, c.nx, c.ny, c.nz,
                                           acf, pc, pc32, mgprol, xg_of(g));
  check_launch("buildP");
}

// ---------------------------------------------------------------------------
// Setup: Galerkin coarse operator XA = P^T A P (VbuildG_7 / VbuildG_27)
// ---------------------------------------------------------------------------
// Stored direction delta of each of the 14 output planes.
__host__ __device__ constexpr int gdelta(int s, int a) {
  return s == 0   ? 0
         : s == 1 ? (a == 0 ? 1 : 0)
         : s == 2 ? (a == 1 ? 1 : 0)
         : s == 3 ? (a == 2 ? 1 : 0)
         : s == 4 ? (a == 2 ? 0 : 1)
         : s == 5 ? (a == 0 ? -1 : (a == 1 ? 1 : 0))
         : s == 6 ? (a == 1 ? 0 : 1)
         : s == 7 ? (a == 0 ? -1 : (a == 1 ? 0 : 1))
         : s == 8 ? (a == 0 ? 0 : 1)
         : s == 9 ? (a == 0 ? 0 : (a == 1 ? -1 : 1))
         : s == 10 ? 1
         : s == 11 ? (a == 0 ? -1 : 1)
         : s == 12 ? (a == 1 ? -1 : 1)
                   : (a == 2 ? 1 : -1);
}

// XA(C, C + delta) = P_C^T A P_{C+delta} is computed through w = A P_C:
//   XA(C, C + delta) = sum over q in [-1, 1]^3 of P_{C+delta}(q) w(2 delta + q),
//   w(r) = sum over d in [-1, 1]^3 of A(C_f + r, C_f + d) P_C(d),
// offsets taken from C_f, the fine point of C.  w is supported on the stencil
// neighbours of the box, r in [-2, 2]^3, and the stored (upper) directions
// have delta_z in {0, 1}, so only the fine planes r_z = -1 .. 2 contribute.
// Per coarse point this loads each fine coupling about once (135 values for
// a 7-point fine operator, 729 for a 27-point one) and 27 + 151 (185) values
// of P, against 1011 (2725) loads when each entry is summed over its (d, e)
// pairs on its own.

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1, so that every
// index-derived offset, plane and test below is a constant expression (a
// loop under #pragma unroll is not expanded when the body is this large).
template <class F, int... I>
__device__ __forceinline__ void static_for_(F &&f,
                                            std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_(f, std::make_integer_sequence<int, N>{});
}

__host__ __device__ constexpr bool in_box(int x, int y, int z) {
  return x >= -1 && x <= 1 && y >= -1 && y <= 1 && z >= -1 && z <= 1;
}
// Fine offset r = (RX, RY, RZ) is 2 delta + q (q in the box) for some stored
// direction delta.
__host__ __device__ constexpr bool rap_used(int rx, int ry, int rz) {
  bool u = false;
  for (int s = 0; s < 14; s++)
    u = u || in_box(rx - 2 * gdelta(s, 0), ry - 2 * gdelta(s, 1),
                    rz - 2 * gdelta(s, 2));
  return u;
}
// e is a stencil direction of an ND-point operator (or 0).
__host__ __device__ constexpr bool stencil_dir(int nd, int ex, int ey, int ez) {
  return in_box(ex, ey, ez) &&
         (nd == 27 || (ex != 0) + (ey != 0) + (ez != 0) <= 1);
}
// w(r) has a term: r is a stencil neighbour of some d in the box.
__host__ __device__ constexpr bool rap_any(int nd, int rx, int ry, int rz) {
  bool a = false;
  for (int d = 0; d < 27; d++)
    a = a || stencil_dir(nd, d % 3 - 1 - rx, (d / 3) % 3 - 1 - ry,
                         d / 9 - 1 - rz);
  return a;
}

// A(F, F + e) at fine point F = pF: the diagonal, or the negated stored
// coupling, held at F for an upper direction and at F + e otherwise.
template <int EX, int EY, int EZ>
__device__ __forceinline__ double afine(const double *__restrict__ ac,
                                        size_t nf, size_t pF, long long sy,
                                        long long sz) {
  if constexpr (EX == 0 && EY == 0 && EZ == 0) {
    return ac[pF];
  } else if constexpr (upper_plane(EX, EY, EZ) >= 0) {
    return -ac[(size_t)upper_plane(EX, EY, EZ) * nf + pF];
  } else {
    return -ac[(size_t)upper_plane(-EX, -EY, -EZ) * nf +
               (size_t)((long long)pF + EX + EY * sy + EZ * sz)];
  }
}

// Contributions of fine plane r_z = RZ to the 14 stored entries.
template <int ND, int RZ>
__device__ __forceinline__ void rap_plane(const double *__restrict__ pc,
                                          const double *__restrict__ ac,
                                          size_t nf, size_t nc, size_t pC,
                                          size_t pCf, long long sy,
                                          long long sz, long long cy,
                                          long long cz, double (&acc)[14]) {
  // no multiply-add contraction: the next level's prolongation divides
  // differences of these entries that cancel where the row sums vanish, so
  // they match the oracle only with the same roundings
#pragma clang fp contract(off)
  static_for<25>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    constexpr int rx = t % 5 - 2, ry = t / 5 - 2;
    if constexpr (rap_used(rx, ry, RZ) && rap_any(ND, rx, ry, RZ)) {
      const size_t pF = (size_t)((long long)pCf + rx + ry * sy + RZ * sz);
      double w = 0.0;
      static_for<27>([&](auto ac_) {
        constexpr int a = decltype(ac_)::value;
        constexpr int dx = a % 3 - 1, dy = (a / 3) % 3 - 1, dz = a / 9 - 1;
        if constexpr (stencil_dir(ND, dx - rx, dy - ry, dz - RZ)) {
#pragma clang fp contract(off)
          w += afine<dx - rx, dy - ry, dz - RZ>(ac, nf, pF, sy, sz) *
               pc[(size_t)pplane(dx, dy, dz) * nc + pC];
        }
      });
      static_for<14>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int qx = rx - 2 * gdelta(s, 0), qy = ry - 2 * gdelta(s, 1),
                      qz = RZ - 2 * gdelta(s, 2);
        if constexpr (in_box(qx, qy, qz)) {
#pragma clang fp contract(off)
          const size_t pC2 = (size_t)((long long)pC + gdelta(s, 0) +
                                      gdelta(s, 1) * cy + gdelta(s, 2) * cz);
          acc[s] += pc[(size_t)pplane(qx, qy, qz) * nc + pC2] * w;
        }
      });
    }
  });
}

// Workgroup: 64 consecutive interior points of a coarse plane (i fastest
// across row ends, so short rows leave no lanes idle) x 4 wavefronts, wave
// w summing the fine plane r_z = w - 1; the four partial sums of each entry
// are added through LDS in plane order.  Tiles are XCD-contiguous (see
// xcd_tile), so the fine planes that neighbouring coarse rows share are
// fetched into one L2.
template <int ND>
__global__ __launch_bounds__(256) void k_buildG(int nxf, int nyf, int nzf,
                                                int nxc, int nyc, int nzc,
                                                const double *__restrict__ pc,
                                                const double *__restrict__ ac,
                                                double *__restrict__ acc,
                                                float *__restrict__ acc32,
                                                XG xg) {
  __shared__ double part[4][14][64];
  int bx, by, bz;
  if (!xcd_tile(xg, bx, by, bz)) return;  // uniform over the workgroup
  const int lane = threadIdx.x, w = threadIdx.y;
  const int q = bx * 64 + lane;
  const bool live = q < (nxc - 2) * (nyc - 2);
  const int ii = 2 + q % (nxc - 2);
  const int jj = 2 + q / (nxc - 2);
  const int kk = 2 + bz;
  const size_t nf = (size_t)nxf * nyf * nzf, nc = (size_t)nxc * nyc * nzc;
  const size_t pC = ix(nxc, nyc, ii, jj, kk);
  const size_t pCf = ix(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1);
  const long long sy = nxf, sz = (long long)nxf * nyf;
  const long long cy = nxc, cz = (long long)nxc * nyc;
  double a[14];
#pragma unroll
  for (int s = 0; s < 14; s++) a[s] = 0.0;
  if (live) {
    switch (w) {
      case 0: rap_plane<ND, -1>(pc, ac, nf, nc, pC, pCf, sy, sz, cy, cz, a); break;
      case 1: rap_plane<ND, 0>(pc, ac, nf, nc, pC, pCf, sy, sz, cy, cz, a); break;
      case 2: rap_plane<ND, 1>(pc, ac, nf, nc, pC, pCf, sy, sz, cy, cz, a); break;
      default: rap_plane<ND, 2>(pc, ac, nf, nc, pC, pCf, sy, sz, cy, cz, a); break;
    }
  }
#pragma unroll
  for (int s = 0; s < 14; s++) part[w][s][lane] = a[s];
  __syncthreads();
  if (!live) return;
  for (int s = w; s < 14; s += 4) {
    const double v = ((part[0][s][lane] + part[1][s][lane]) + part[2][s][lane]) +
                     part[3][s][lane];
    const double out = s == 0 ? v : -v;
    acc[(size_t)s * nc + pC] = out;
    if (acc32) acc32[(size_t)s * nc + pC] = (float)out;
  }
}

void buildG(const Dim &f, const Dim &c, int numdia, const double *pc,
            const double *acf, double *acc, hipStream_t s, float *acc32) {
  if (c.nx < 3 || c.ny < 3 || c.nz < 3) return;
  dim3 g(cdiv((c.nx - 2) * (c.ny - 2), 64), 1, c.nz - 2);
  if (numdia == 7)
    k_buildG<7><<<xcd_blocks(g), dim3(64, 4), 0, s>>>(
        f.nx, f.ny, f.nz, c.nx, c.ny, c.nz, pc, acf, acc, acc32, xg_of(g));
  else
    k_buildG<27><<<xcd_blocks(g), dim3(64, 4), 0, s>>>(
        f.nx, f.ny, f.nz, c.nx, c.ny, c.nz, pc, acf, acc, acc32, xg_of(g));
  check_launch("buildG");
}

// ---------------------------------------------------------------------------
// Grouped launches (serving): one launch runs the same step for up to GMAX
// independent problems.  Problem q's arguments sit in slot q of a table
// passed by value, blockIdx.y selects the slot, and the slot's workgroups
// are those with blockIdx.x below its own launch's workgroup count: every
// problem runs exactly the workgroups, with the same indices and the same
// arithmetic, as its own launch would, so its results are bit-identical to
// solving it alone.  The launchers split the problems by kernel variant
// and into launches of at most GMAX.
// ---------------------------------------------------------------------------
template <class A>
struct GTab {
  A a[GMAX];
};

// Launch f(tab, count) for each run of at most GMAX consecutive entries of
// the problems whose key equals key (entries built by mk(q) into the tab).
template <class A, class KeyF, class MkF, class LaunchF>
static void g_split(int n, KeyF key, MkF mk, LaunchF launch) {
  std::vector<char> done(n, 0);
  for (int q0 = 0; q0 < n; q0++) {
    if (done[q0]) continue;
    const auto k0 = key(q0);
    GTab<A> t{};
    int m = 0;
    for (int q = q0; q < n; q++) {
      if (done[q] || !(key(q) == k0)) continue;
      done[q] = 1;
      t.a[m++] = mk(q);
      if (m == GMAX) {
        launch(t, m, k0);
        m = 0;
      }
    }
    if (m) launch(t, m, k0);
  }
}

template <class T>
struct A_gs7 {
  int nx, ny, nz;
  const T *ac, *cc, *fc, *xin;
  T *xout;
  int c1, kc, gx, gy, gz;
  unsigned nblk;
  const T *dg;
};
template <class T, bool ZX, bool DG>
__global__ __launch_bounds__(SW_TX *SW_TY) void kb_gs7_sweep(
    const GTab<A_gs7<T>> t) {
  const A_gs7<T> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  gs7_sweep_body<T, ZX, DG>(a.nx, a.ny, a.nz, a.ac, a.cc, a.fc, a.xin, a.xout,
                            a.c1, a.kc, a.gx, a.gy, a.gz, blockIdx.x, a.dg);
}

template <class T>
void g_gs7_fused(const GSmooth<T> *j, int n, int nsweeps, int iadjoint,
                 bool zero_init, hipStream_t s) {
  if (nsweeps <= 0 || n <= 0) return;
  struct G {
    int kc, gx, gy, gz;
    unsigned nblk;
    T *src, *dst;
  };
  std::vector<G> g(n);
  const int cus = device_cus();
  for (int q = 0; q < n; q++) {  // the launch shape of gs7_fused
    const Dim &d = j[q].d;
    G &e = g[q];
    e.gx = cdiv(d.nx - 2, SW_TX - 4), e.gy = cdiv(d.ny - 2, SW_TY - 4);
    const int zb = std::max(1, cus / (e.gx * e.gy));
    e.kc = std::max(4, cdiv(d.nz - 2, zb));
    e.gz = cdiv(d.nz - 2, e.kc);
    e.nblk = 8 * cdiv(e.gx * e.gy * e.gz, 8);
    e.src = j[q].x, e.dst = j[q].tmp;
  }
  const int c1 = iadjoint ? 1 : 0;
  for (int it = 0; it < nsweeps; it++) {
    const bool zx = zero_init && it == 0;
    g_split<A_gs7<T>>(
        n, [&](int q) { return j[q].dg ? 1 : 0; },
        [&](int q) {
          const Dim &d = j[q].d;
          return A_gs7<T>{d.nx,    d.ny,     d.nz,     j[q].ac,   j[q].cc,
                          j[q].f,  g[q].src, g[q].dst, c1,        g[q].kc,
                          g[q].gx, g[q].gy,  g[q].gz,  g[q].nblk, j[q].dg};
        },
        [&](const GTab<A_gs7<T>> &t, int m, int dgk) {
          unsigned mx = 0;
          for (int q = 0; q < m; q++) mx = std::max(mx, t.a[q].nblk);
          dim3 grid(mx, m), b(SW_TX, SW_TY);
          if (zx && dgk)
            kb_gs7_sweep<T, true, true><<<grid, b, 0, s>>>(t);
          else if (zx)
            kb_gs7_sweep<T, true, false><<<grid, b, 0, s>>>(t);
          else if (dgk)
            kb_gs7_sweep<T, false, true><<<grid, b, 0, s>>>(t);
          else
            kb_gs7_sweep<T, false, false><<<grid, b, 0, s>>>(t);
        });
    for (int q = 0; q < n; q++) std::swap(g[q].src, g[q].dst);
  }
  for (int q = 0; q < n; q++)
    if (g[q].src != j[q].x)
      hipMemcpyAsync(j[q].x, g[q].src, j[q].d.n() * sizeof(T),
                     hipMemcpyDeviceToDevice, s);
  check_launch("g_gs7_fused");
}

template <class T>
struct A_g27 {
  int nx, ny, nz;
  const T *ac, *cc, *fc, *cur;
  T *nxt;
  const T *po;
  int pz, bp, nplanes, nstrips, zmask;
  unsigned nblk;
};
template <class T, int ADJ, int PPL, int RS2>
__global__ __launch_bounds__(64 * (RS2 + 1)) void kb_gs27_pair(
    const GTab<A_g27<T>> t) {
  const A_g27<T> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  gs27_pair_body<T, ADJ, PPL, RS2>(a.nx, a.ny, a.nz, a.ac, a.cc, a.fc, a.cur,
                                   a.nxt, a.po, a.pz, a.bp, a.nplanes,
                                   a.nstrips, a.zmask, blockIdx.x);
}

// Problems the paired launches of gs27 take (an even number of sweeps,
// rows of at most ROWS_MAXN points, a second buffer).
template <class T>
static bool g27_pairs_ok(const GSmooth<T> &j, int nsweeps) {
  const Dim &d = j.d;
  return d.nx >= 3 && d.ny >= 4 && d.nz >= 3 && d.nx <= ROWS_MAXN && j.tmp &&
         j.tmp != j.x && nsweeps % 2 == 0;
}

template <class T>
void g_gs27(const GSmooth<T> *j, int n, int nsweeps, int iadjoint,
            bool zero_init, hipStream_t s) {
  if (nsweeps <= 0 || n <= 0) return;
  std::vector<int> idx;  // problems taken here; the others run alone
  for (int q = 0; q < n; q++) {
    if (g27_pairs_ok(j[q], nsweeps))
      idx.push_back(q);
    else
      gs27<T>(j[q].d, j[q].ac, j[q].cc, j[q].f, j[q].x, nsweeps, iadjoint, s,
              zero_init, j[q].tmp);
  }
  const int m = (int)idx.size();
  if (!m) return;
  const int bp = iadjoint ? 0 : 1;
  struct G {
    int ppl;
    bool big;
    int nstrips;
    T *cur, *nxt;
  };
  std::vector<G> g(m);
  for (int r = 0; r < m; r++) {  // the launch shape of gs27's paired sweeps
    const GSmooth<T> &J = j[idx[r]];
    g[r].ppl = J.d.nx <= 129 ? 1 : (J.d.nx <= 257 ? 2 : 4);
    g[r].big = J.d.nx >= 65 && J.d.nx <= 129;  // as gs27
    g[r].nstrips = cdiv(count_parity(J.d.ny, bp), g[r].big ? 7 : 3);
    g[r].cur = J.x, g[r].nxt = J.tmp;
  }
  for (int done = 0; done < nsweeps; done++) {
    int zmask = (zero_init && done == 0) ? 0xF : 0;
    for (int h = 0; h < 2; h++) {
      const int pz = iadjoint ? 1 - h : h;
      std::vector<int> sub;  // problems with planes of this parity
      for (int r = 0; r < m; r++)
        if (count_parity(j[idx[r]].d.nz, pz) > 0) sub.push_back(r);
      g_split<A_g27<T>>(
          (int)sub.size(),
          [&](int q) { return g[sub[q]].ppl * 2 + (g[sub[q]].big ? 1 : 0); },
          [&](int q) {
            const int r = sub[q];
            const GSmooth<T> &J = j[idx[r]];
            const int npl = count_parity(J.d.nz, pz);
            const T *po = h == 0 ? g[r].cur : g[r].nxt;
            return A_g27<T>{J.d.nx,     J.d.ny,     J.d.nz,  J.ac, J.cc,
                            J.f,        g[r].cur,   g[r].nxt, po,  pz,
                            bp,         npl,        g[r].nstrips, zmask,
                            (unsigned)(8 * cdiv(npl * g[r].nstrips, 8))};
          },
          [&](const GTab<A_g27<T>> &t, int cnt, int key) {
            unsigned mx = 0;
            for (int q = 0; q < cnt; q++) mx = std::max(mx, t.a[q].nblk);
            const int ppl = key / 2;
            const bool big = key & 1;
            const dim3 grid(mx, cnt);
#define GPAIR(A, Q)                                                         \
  do {                                                                      \
    if (big)                                                                \
      kb_gs27_pair<T, A, Q, 7><<<grid, 64 * 8, 0, s>>>(t);                  \
    else                                                                    \
      kb_gs27_pair<T, A, Q, 3><<<grid, 64 * 4, 0, s>>>(t);                  \
  } while (0)
            if (iadjoint) {
              if (ppl == 1) GPAIR(1, 1);
              else if (ppl == 2) GPAIR(1, 2);
              else GPAIR(1, 4);
            } else {
              if (ppl == 1) GPAIR(0, 1);
              else if (ppl == 2) GPAIR(0, 2);
              else GPAIR(0, 4);
            }
#undef GPAIR
          });
      zmask &= ~(0x3 << (2 * pz));
    }
    for (int r = 0; r < m; r++) std::swap(g[r].cur, g[r].nxt);
  }
  check_launch("g_gs27");
}

template <class T, class TO>
struct A_res {
  int nx, ny, nz;
  const T *ac, *cc, *fc, *x;
  TO *r;
  Red rd;
  int do_red, kc;
  XG xg;
  unsigned nblk;
  const T *dg;
};
template <class T, class TO, int ND>
__global__ __launch_bounds__(256) void kb_resid(const GTab<A_res<T, TO>> t) {
  const A_res<T, TO> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  resid_body<T, TO, ND>(a.nx, a.ny, a.nz, a.ac, a.cc, a.fc, a.x, a.r, a.rd,
                        a.do_red, a.kc, a.xg, blockIdx.x);
}
// 7-point levels with the precomputed diagonal: k_resid7m's march (xg holds
// its tile counts)
template <class T, class TO>
__global__ __launch_bounds__(RM_TX *RM_TY) void kb_resid7m(
    const GTab<A_res<T, TO>> t) {
  const A_res<T, TO> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  resid7m_body<T, TO>(a.nx, a.ny, a.nz, a.ac, a.fc, a.x, a.r, a.dg, a.rd,
                      a.do_red, a.kc, a.xg.gx, a.xg.gy, a.xg.gz, blockIdx.x);
}

struct A_rf {
  const double *part;
  unsigned nblk;
  double *out, *out2;
};
template <int NV, int MODE>
__global__ __launch_bounds__(1024) void kb_red_final(const GTab<A_rf> t) {
  const A_rf &a = t.a[blockIdx.y];
  red_final_body<NV, MODE>(a.part, a.nblk, a.out);
  if (threadIdx.x == 0 && a.out2) a.out2[0] = a.out[0];
}

template <class T, class TO>
void g_resid(const GResid<T, TO> *j, int n, hipStream_t s) {
  if (n <= 0) return;
  std::vector<A_res<T, TO>> a(n);
  for (int q = 0; q < n; q++) {  // the launch shape of resid
    const Dim &d = j[q].d;
    const bool red = j[q].red != nullptr;
    const T *dg = j[q].numdia == 7 && resid7m_fits(d) ? j[q].dg : nullptr;
    int kc;
    XG xg;
    unsigned nblk;
    if (dg) {
      nblk = resid7m_shape(d, kc, xg.gx, xg.gy, xg.gz);
    } else {
      kc = kchunk(d, red ? 40 : 80);
      const dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY),
                   cdiv(d.nz - 2, kc));
      xg = xg_of(g), nblk = xcd_blocks(g);
    }
    a[q] = A_res<T, TO>{d.nx,    d.ny,    d.nz,   j[q].ac,
                        j[q].cc, j[q].fc, j[q].x, j[q].r,
                        red ? *j[q].red : Red{}, red ? 1 : 0, kc,
                        xg,      nblk,    dg};
  }
  // variant key: 27 (27-point), 7 (7-point), 8 (7-point with dg)
  g_split<A_res<T, TO>>(
      n, [&](int q) { return j[q].numdia == 7 && a[q].dg ? 8 : j[q].numdia; },
      [&](int q) { return a[q]; },
      [&](const GTab<A_res<T, TO>> &t, int m, int key) {
        unsigned mx = 0;
        for (int q = 0; q < m; q++) mx = std::max(mx, t.a[q].nblk);
        const dim3 grid(mx, m), b(BX, BY);
        if (key == 8)
          kb_resid7m<T, TO><<<grid, dim3(RM_TX, RM_TY), 0, s>>>(t);
        else if (key == 7)
          kb_resid<T, TO, 7><<<grid, b, 0, s>>>(t);
        else
          kb_resid<T, TO, 27><<<grid, b, 0, s>>>(t);
      });
  std::vector<int> rq;
  for (int q = 0; q < n; q++)
    if (j[q].red) rq.push_back(q);
  g_split<A_rf>(
      (int)rq.size(), [&](int) { return 0; },
      [&](int q) {
        const int r = rq[q];
        return A_rf{j[r].red->part, a[r].nblk, j[r].red->out, j[r].out2};
      },
      [&](const GTab<A_rf> &t, int m, int) {
        kb_red_final<1, 0><<<dim3(1, m), 1024, 0, s>>>(t);
      });
  check_launch("g_resid");
}

template <class T>
struct A_rs {
  int nxf, nyf, nxc, nyc, nzc;
  const T *xin;
  T *xout;
  const T *pc;
  XG xg;
  unsigned nblk;
};
template <class T>
__global__ __launch_bounds__(256) void kb_restrict(const GTab<A_rs<T>> t) {
  const A_rs<T> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  restrict_body<T>(a.nxf, a.nyf, a.nxc, a.nyc, a.nzc, a.xin, a.xout, a.pc,
                   a.xg, nullptr, blockIdx.x);
}

template <class T>
void g_restrict(const GRestrict<T> *j, int n, hipStream_t s) {
  std::vector<int> ok;
  for (int q = 0; q < n; q++) {
    const Dim &f = j[q].f, &c = j[q].c;
    if (c.nx < 3 || c.ny < 3 || c.nz < 3) continue;
    if (f.nx != 2 * c.nx - 1 || f.ny != 2 * c.ny - 1 || f.nz != 2 * c.nz - 1) {
      fprintf(stderr, "pmg: restrict shape mismatch\n");
      continue;
    }
    ok.push_back(q);
  }
  g_split<A_rs<T>>(
      (int)ok.size(), [&](int) { return 0; },
      [&](int q) {
        const GRestrict<T> &J = j[ok[q]];
        const dim3 g(cdiv(J.c.nx - 2, BX), cdiv(J.c.ny - 2, BY), J.c.nz - 2);
        return A_rs<T>{J.f.nx, J.f.ny, J.c.nx,   J.c.ny,   J.c.nz,
                       J.xin,  J.xout, J.pc,     xg_of(g), xcd_blocks(g)};
      },
      [&](const GTab<A_rs<T>> &t, int m, int) {
        unsigned mx = 0;
        for (int q = 0; q < m; q++) mx = std::max(mx, t.a[q].nblk);
        kb_restrict<T><<<dim3(mx, m), dim3(BX, BY), 0, s>>>(t);
      });
  check_launch("g_restrict");
}

template <class T>
struct A_mv {
  int nx, ny, nz;
  const T *ac, *cc, *x, *w0;
  T *y;
  Red rd;
  int kc;
  XG xg;
  unsigned nblk;
};
template <class T, int ND>
__global__ __launch_bounds__(256) void kb_matvec_dots(const GTab<A_mv<T>> t) {
  const A_mv<T> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  matvec_dots_body<T, ND>(a.nx, a.ny, a.nz, a.ac, a.cc, a.x, a.w0, a.y, a.rd,
                          1, a.kc, a.xg, blockIdx.x);
}

template <class T>
void g_matvec_dots_part(const GMatvec<T> *j, int n, unsigned *nb,
                        hipStream_t s) {
  std::vector<A_mv<T>> a(n);
  for (int q = 0; q < n; q++) {  // the launch shape of matvec_dots_part
    const Dim &d = j[q].d;
    const int kc = kchunk(d);
    const dim3 g(cdiv(d.nx - 2, BX), cdiv(d.ny - 2, BY), cdiv(d.nz - 2, kc));
    Red rd;
    rd.part = j[q].part;
    a[q] = A_mv<T>{d.nx,   d.ny,    d.nz, j[q].ac,  j[q].cc,      j[q].x,
                   j[q].w0, j[q].y, rd,   kc,       xg_of(g),     xcd_blocks(g)};
    nb[q] = a[q].nblk;
  }
  g_split<A_mv<T>>(
      n, [&](int q) { return j[q].numdia; }, [&](int q) { return a[q]; },
      [&](const GTab<A_mv<T>> &t, int m, int nd) {
        unsigned mx = 0;
        for (int q = 0; q < m; q++) mx = std::max(mx, t.a[q].nblk);
        const dim3 grid(mx, m), b(BX, BY);
        if (nd == 7)
          kb_matvec_dots<T, 7><<<grid, b, 0, s>>>(t);
        else
          kb_matvec_dots<T, 27><<<grid, b, 0, s>>>(t);
      });
  check_launch("g_matvec_dots");
}

template <class T>
struct A_ip {
  int nxc, nyc, nzc, nxf, nyf, nzf;
  const T *xc, *pc;
  T *xf;
  const double *part;
  unsigned npart;
  XG xg;
  unsigned nblk;
};
template <class T>
__global__ __launch_bounds__(256) void kb_interp_damped(const GTab<A_ip<T>> t) {
  const A_ip<T> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  interp_body<T, true>(a.nxc, a.nyc, a.nzc, a.nxf, a.nyf, a.nzf, a.xc, a.pc,
                       a.xf, nullptr, a.part, a.npart, a.xg, blockIdx.x);
}

template <class T>
void g_interp_damped(const GInterp<T> *j, int n, hipStream_t s) {
  std::vector<int> ok;
  for (int q = 0; q < n; q++) {
    const Dim &c = j[q].c, &f = j[q].f;
    if (f.nx != 2 * c.nx - 1 || f.ny != 2 * c.ny - 1 || f.nz != 2 * c.nz - 1) {
      fprintf(stderr, "pmg: interp shape mismatch\n");
      continue;
    }
    ok.push_back(q);
  }
  g_split<A_ip<T>>(
      (int)ok.size(), [&](int) { return 0; },
      [&](int q) {
        const GInterp<T> &J = j[ok[q]];
        const dim3 g(cdiv(J.c.nx - 1, BX), cdiv(J.c.ny - 1, BY), J.c.nz - 1);
        return A_ip<T>{J.c.nx, J.c.ny, J.c.nz, J.f.nx, J.f.ny,
                       J.f.nz, J.xc,   J.pc,   J.xf,   J.part,
                       J.nblk, xg_of(g), xcd_blocks(g)};
      },
      [&](const GTab<A_ip<T>> &t, int m, int) {
        unsigned mx = 0;
        for (int q = 0; q < m; q++) mx = std::max(mx, t.a[q].nblk);
        kb_interp_damped<T><<<dim3(mx, m), dim3(BX, BY), 0, s>>>(t);
      });
  check_launch("g_interp_damped");
}

template <class T>
struct A_cg {
  int nx, ny, nz;
  const double *ac, *cc;
  const T *f;
  T *x;
  int itmax;
  double errtol;
};
template <class T, bool ZT>
__global__ __launch_bounds__(CG_RNT) void kb_cg_rows(const GTab<A_cg<T>> t) {
  const A_cg<T> &a = t.a[blockIdx.y];
  cg_rows_body<T, false, false, ZT>(a.nx, a.ny, a.nz, a.ac, a.cc, a.f, a.x,
                                    a.itmax, a.errtol);
}

// The 27-point coarsest levels k_cg_rows takes (cg_launch's first case).
static bool g_cg_rows_ok(const Dim &d, int numdia) {
  const int nrow = (int)((d.nx - 2 + CG_R - 1) / CG_R) * (d.ny - 2) * (d.nz - 2);
  return numdia == 27 && nrow <= CG_RNT && cg_fits(d);
}

template <class T>
void g_cg(const GCoarse<T> *j, int n, int itmax, double errtol,
          hipStream_t s) {
  std::vector<int> ok;
  for (int q = 0; q < n; q++) {
    if (g_cg_rows_ok(j[q].d, j[q].numdia))
      ok.push_back(q);
    else
      cg_single<T>(j[q].d, j[q].numdia, j[q].ac, j[q].cc, j[q].f, j[q].x,
                   itmax, errtol, s);
  }
  // One launch for every coarsest shape: the workgroup is as wide, and its
  // LDS as large, as the largest shape needs.  k_cg_rows gives the same
  // bits at any width: threads past a shape's rows hold no unknowns and add
  // exact zeros to wavefront slots the reductions sum in a fixed order
  // (cg_allsum), whether those wavefronts exist or not.
  g_split<A_cg<T>>(
      (int)ok.size(), [&](int) { return 0; },
      [&](int q) {
        const GCoarse<T> &J = j[ok[q]];
        return A_cg<T>{J.d.nx, J.d.ny, J.d.nz, J.ac, J.cc, J.f, J.x, itmax,
                       errtol};
      },
      [&](const GTab<A_cg<T>> &t, int m, int) {
        int nt = 64;
        size_t lds = 0;
        for (int q = 0; q < m; q++) {
          const Dim d{t.a[q].nx, t.a[q].ny, t.a[q].nz};
          const int nrow =
              (int)((d.nx - 2 + CG_R - 1) / CG_R) * (d.ny - 2) * (d.nz - 2);
          nt = std::max(nt, (nrow + 63) / 64 * 64);
          lds = std::max(lds, d.n() * sizeof(double) +
                                  (CG_R + 2) * sizeof(double));
        }
        if (errtol == 0.0)
          kb_cg_rows<T, true><<<dim3(1, m), nt, lds, s>>>(t);
        else
          kb_cg_rows<T, false><<<dim3(1, m), nt, lds, s>>>(t);
      });
  check_launch("g_cg");
}

template <class T>
struct A_acr {
  int nx, ny, nz;
  const T *e;
  double *u;
  int gx, gy;
  unsigned nblk;
};
template <class T>
__global__ __launch_bounds__(256) void kb_addcorr(const GTab<A_acr<T>> t) {
  const A_acr<T> &a = t.a[blockIdx.y];
  if (blockIdx.x >= a.nblk) return;
  const int b = (int)blockIdx.x;
  addcorr_body<T>(a.nx, a.ny, a.nz, a.e, a.u, b % a.gx, (b / a.gx) % a.gy,
                  b / (a.gx * a.gy));
}

template <class T>
void g_addcorr(const GAddcorr<T> *j, int n, hipStream_t s) {
  g_split<A_acr<T>>(
      n, [&](int) { return 0; },
      [&](int q) {
        const Dim &d = j[q].d;
        const int gx = cdiv(d.nx - 2, BX), gy = cdiv(d.ny - 2, BY),
                  gz = cdiv(d.nz - 2, KCH);
        return A_acr<T>{d.nx, d.ny, d.nz, j[q].e, j[q].u, gx, gy,
                        (unsigned)(gx * gy * gz)};
      },
      [&](const GTab<A_acr<T>> &t, int m, int) {
        unsigned mx = 0;
        for (int q = 0; q < m; q++) mx = std::max(mx, t.a[q].nblk);
        kb_addcorr<T><<<dim3(mx, m), dim3(BX, BY), 0, s>>>(t);
      });
  check_launch("g_addcorr");
}

// ---------------------------------------------------------------------------
// explicit instantiations
// ---------------------------------------------------------------------------
template void dot_energy<double>(const Dim &, int, const double *,
                                 const double *, const double *, double *,
                                 const Red &, hipStream_t);

#define INST(T)                                                               \
  template void gs7<T>(const Dim &, const T *, const T *, const T *, T *,     \
                       int, int, hipStream_t);                                \
  template bool gs7_fused<T>(const Dim &, const T *, const T *, const T *,    \
                             T *, T *, int, int, hipStream_t, bool,           \
                             const T *, double *);                            \
  template void diag_sum<T>(const T *, const T *, T *, size_t, hipStream_t);  \
  template void gs27<T>(const Dim &, const T *, const T *, const T *, T *,    \
                        int, int, hipStream_t, bool, T *);                    \
  template void matvec_dots<T>(const Dim &, int, const T *, const T *,        \
                               const T *, const T *, T *, const Red *,        \
                               hipStream_t);                                  \
  template void restrict_<T>(const Dim &, const Dim &, const T *, T *,        \
                             const T *, hipStream_t, float *);                \
  template void interp<T>(const Dim &, const Dim &, const T *, const T *,     \
                          T *, const double *, hipStream_t);                  \
  template void interp_damped<T>(const Dim &, const Dim &, const T *,         \
                                 const T *, T *, const double *, unsigned,    \
                                 hipStream_t);                                \
  template unsigned matvec_dots_part<T>(const Dim &, int, const T *,          \
                                        const T *, const T *, const T *, T *, \
                                        double *, hipStream_t);               \
  template void nrm1<T>(const Dim &, const T *, const Red &, hipStream_t);     \
  template void cg_single<T>(const Dim &, int, const double *,                \
                             const double *, const T *, T *, int, double,     \
                             hipStream_t);                                    \
  template void cg_class<T>(const Dim &, int, const double *, const double *, \
                            const T *, T *, int, double, hipStream_t);        \
  template void fill<T>(T *, size_t, T, hipStream_t);                         \
  template void axpy_interior<T>(const Dim &, double, const T *, T *,         \
                                 hipStream_t);                                \
  template void add_correction<T>(const Dim &, const T *, double *,           \
                                  hipStream_t);
INST(double)
INST(float)
#undef INST
template void resid<double, double>(const Dim &, int, const double *,
                                    const double *, const double *,
                                    const double *, double *, const Red *,
                                    hipStream_t, const double *);
template void resid<double, float>(const Dim &, int, const double *,
                                   const double *, const double *,
                                   const double *, float *, const Red *,
                                   hipStream_t, const double *);
template void resid<float, float>(const Dim &, int, const float *,
                                  const float *, const float *, const float *,
                                  float *, const Red *, hipStream_t,
                                  const float *);
template void convert<double, float>(const double *, float *, size_t,
                                     hipStream_t);
template void convert<float, double>(const float *, double *, size_t,
                                     hipStream_t);
template void fill<int>(int *, size_t, int, hipStream_t);

#define GINST(T)                                                              \
  template void g_gs7_fused<T>(const GSmooth<T> *, int, int, int, bool,       \
                               hipStream_t);                                  \
  template void g_gs27<T>(const GSmooth<T> *, int, int, int, bool,            \
                          hipStream_t);                                       \
  template void g_restrict<T>(const GRestrict<T> *, int, hipStream_t);        \
  template void g_matvec_dots_part<T>(const GMatvec<T> *, int, unsigned *,    \
                                      hipStream_t);                           \
  template void g_interp_damped<T>(const GInterp<T> *, int, hipStream_t);     \
  template void g_cg<T>(const GCoarse<T> *, int, int, double, hipStream_t);   \
  template void g_addcorr<T>(const GAddcorr<T> *, int, hipStream_t);
GINST(double)
GINST(float)
#undef GINST
template void g_resid<double, double>(const GResid<double, double> *, int,
                                      hipStream_t);
template void g_resid<double, float>(const GResid<double, float> *, int,
                                     hipStream_t);
template void g_resid<float, float>(const GResid<float, float> *, int,
                                    hipStream_t);

}  // namespace pmg
