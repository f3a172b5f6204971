// pmg_pbe.hip -- Poisson-Boltzmann input assembly on the GPU.
//
// Turns an atom list {x, y, z, radius, charge} into the fine-level
// arguments of the PMG driver (a1cf, a2cf, a3cf, ccf, fcf, the six
// Debye-Hueckel faces and the node coordinates), as the reference does in
// PBE_pmgc_class.h:191-344 with the benchmark's molecule conventions:
//   - nodes x_i = c0 + i*h with c0 = -(n-1)h/2;
//   - a node is inside an atom when |x-a|^2 <= r^2 and inside its ion
//     exclusion shell when |x-a|^2 <= (r+2)^2 (the inclusive sphere test of
//     dot_finder.h:24-87; shell as PBE_pmgc_class.h:276-281);
//   - eps = 2 inside, 78.54 outside; a1cf(i) = harmonic_average(eps(i),
//     eps(i+1)) (scientific_quantities.h:30-32), the node value on the last
//     plane;
//   - ccf = 0 inside the shell, eps_w*kappa^2 outside (PBE_pmgc_class.h:293-297);
//   - fcf: trilinear spreading of 4*pi*lambda_B*q over the atom's cell
//     corners, summed over atoms in list order
//     (create_charge_distribution.h:16-61, PBE_pmgc_class.h:331-339);
//   - faces: Debye-Hueckel spheres summed in list order
//     (PBE_pmgc_class.h:173-189, Debye_Hueckel_functions.h:10-14).
//
// Kernels.  The reference marks spheres per atom and spreads charges in a
// serial single_task.  Here:
//   k_pb_init   clears the per-node cell-list heads and the two marks;
//   k_pb_mark   one workgroup per atom writes the marks of the nodes in its
//               shell's bounding box (idempotent byte stores, no atomics)
//               and pushes the atom onto the list of its cell's lower corner;
//   k_pb_nodes  one thread per node writes a1cf..fcf in one coalesced pass;
//               the charge of a node is gathered from the lists of its (up to)
//               8 cells, taken in increasing atom index, so the sum has the
//               reference's serial order and the result is deterministic;
//   k_pb_faces  four wavefronts per 64 low/high boundary node pairs split
//               the atoms, staged through LDS in chunks.
// Floating-point contraction is off in these kernels so that every
// coordinate, weight and mark is the same double as the serial restatement
// computes; only the boundary values (exp, reordered sums) differ from it
// by rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "pmg_internal.h"

namespace pmg {

constexpr double EPS_IN = 2.0, EPS_OUT = 78.54, ION_SHELL = 2.0;  // defaults

namespace {
constexpr double PI_D = 3.14159265358979323846;

inline int cdivp(long long a, int b) { return (int)((a + b - 1) / b); }

__device__ inline double harmonic_average(double x, double y) {
#pragma clang fp contract(off)
  return 2 * x * y / (x + y);
}

__global__ void k_pb_init(size_t N, int *head, uint8_t *inside,
                          uint8_t *excl) {
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < N;
       p += (size_t)gridDim.x * blockDim.x) {
    head[p] = -1;
    inside[p] = 0;
    excl[p] = 0;
  }
}

__global__ __launch_bounds__(256) void k_pb_mark(
    int n, double c0, double h, double shell, int natoms,
    const double *__restrict__ atoms, int *head, int *next, uint8_t *inside,
    uint8_t *excl) {
#pragma clang fp contract(off)
  const int a = blockIdx.x;
  if (a >= natoms) return;
  const double *A = atoms + 5 * (size_t)a;
  const double ax = A[0], ay = A[1], az = A[2], r = A[3];
  const double re = r + shell;
  int lo[3], hi[3];
  const double P[3] = {ax, ay, az};
#pragma unroll
  for (int d = 0; d < 3; d++) {  // a superset of the shell's nodes
    double l = floor((P[d] - re - c0) / h) - 1.0;
    double u = ceil((P[d] + re - c0) / h) + 1.0;
    if (!(l >= 0.0)) l = 0.0;
    if (!(u <= n - 1.0)) u = n - 1.0;
    lo[d] = (int)l, hi[d] = (int)u;
  }
  if (threadIdx.x == 0) {  // push onto the list of the cell's lower corner
    int c[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double fl = floor((P[d] - c0) / h);
      if (!(fl >= 0.0 && fl <= (double)(n - 2))) ok = false;
      c[d] = ok ? (int)fl : 0;
    }
    if (ok) {
      size_t q = ((size_t)c[2] * n + c[1]) * n + c[0];
      next[a] = atomicExch(&head[q], a);
    }
  }
  const int bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1,
            bz = hi[2] - lo[2] + 1;
  if (bx <= 0 || by <= 0 || bz <= 0) return;
  const int vol = bx * by * bz;
  const double re2 = re * re, r2 = r * r;
  for (int t = threadIdx.x; t < vol; t += blockDim.x) {
    const int i = lo[0] + t % bx, rem = t / bx;
    const int j = lo[1] + rem % by, k = lo[2] + rem / by;
    const double dx = (c0 + i * h) - ax, dy = (c0 + j * h) - ay,
                 dz = (c0 + k * h) - az;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const size_t p = ((size_t)k * n + j) * n + i;
    if (d2 <= re2) excl[p] = 1;
    if (d2 <= r2) inside[p] = 1;
  }
}

// Charge atom a puts on corner (di,dj,dk) of its cell.
__device__ inline double corner_charge(const double *__restrict__ atoms, int a,
                                       int corner, double c0, double h,
                                       double q4, double h3) {
#pragma clang fp contract(off)
  const double *A = atoms + 5 * (size_t)a;
  double w = 1.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double g = (A[d] - c0) / h;
    const int c = (int)floor(g);
    const double wl = (c + 1) - g, wu = g - c;
    w *= ((corner >> d) & 1) ? wu : wl;
  }
  const double cq = q4 * A[4];
  return cq * w / h3;
}

__global__ __launch_bounds__(256) void k_pb_nodes(
    int n, double c0, double h, double q4, double h3, double kap2e,
    double eps_in, double eps_out,
    const double *__restrict__ atoms, const int *__restrict__ head,
    const int *__restrict__ next, const uint8_t *__restrict__ inside,
    const uint8_t *__restrict__ excl, double *xf, double *yf, double *zf,
    double *a1, double *a2, double *a3, double *ccf, double *fcf) {
#pragma clang fp contract(off)
  // 256 consecutive nodes of plane k, i fastest across row ends
  const int q = blockIdx.x * 256 + threadIdx.y * 64 + threadIdx.x;
  if (q >= n * n) return;
  const int i = q % n, j = q / n;
  const int k = blockIdx.z;
  const size_t nn = (size_t)n * n;
  const size_t p = (size_t)k * nn + (size_t)j * n + i;
  if (j == 0 && k == 0) xf[i] = yf[i] = zf[i] = c0 + i * h;
  const double e0 = inside[p] ? eps_in : eps_out;
  a1[p] = i + 1 < n ? harmonic_average(e0, inside[p + 1] ? eps_in : eps_out)
                    : e0;
  a2[p] = j + 1 < n ? harmonic_average(e0, inside[p + n] ? eps_in : eps_out)
                    : e0;
  a3[p] = k + 1 < n ? harmonic_average(e0, inside[p + nn] ? eps_in : eps_out)
                    : e0;
  ccf[p] = excl[p] ? 0.0 : kap2e;
  // charge: atoms of the cells whose corner (di,dj,dk) is this node, in
  // increasing index (selection over the unordered lists)
  double acc = 0.0;
  int last = -1;
  for (;;) {
    int best = INT_MAX, bc = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int qi = i - (c & 1), qj = j - ((c >> 1) & 1),
                qk = k - ((c >> 2) & 1);
      if (qi < 0 || qj < 0 || qk < 0 || qi > n - 2 || qj > n - 2 ||
          qk > n - 2)
        continue;
      const size_t q = (size_t)qk * nn + (size_t)qj * n + qi;
      for (int a = head[q]; a >= 0; a = next[a])
        if (a > last && a < best) best = a, bc = c;
    }
    if (best == INT_MAX) break;
    acc += corner_charge(atoms, best, bc, c0, h, q4, h3);
    last = best;
  }
  fcf[p] = acc;
}

// exp(x) for -708 <= x <= 0 by Tang's table-driven method (P. T. P. Tang,
// "Table-driven implementation of the exponential function in IEEE
// floating-point arithmetic", ACM TOMS 15(2):144-157, 1989):
// x = (64 m + j) ln2/64 + R with |R| <= ln2/128, exp(x) = 2^m 2^(j/64) e^R,
// e^R - 1 by its Taylor polynomial of degree 5 (truncation error below
// 3.5e-17), ln2/64 split as L1 + L2 with N L1 exact, 2^(j/64) as lead +
// trail (tables in LDS).  Within one ulp of exp (checked against the C
// library on 2e6 arguments in [-60, 0], as with degree 6); about half the
// instructions of the general routine.
constexpr double EXP_INV_L = 0x1.71547652b82fep+6;  // 64 / ln2
constexpr double EXP_L1 = 0x1.62e42fef00000p-7;      // ln2/64, low 20 bits zero
constexpr double EXP_L2 = 0x1.473de6af278edp-40;      // ln2/64 - L1
// 2^(j/64), j = 0..63: the double nearest to it (lead) and the double
// nearest to the remainder (trail), computed from the definition
__constant__ double c_exp2_lead[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0,
    0x1.0874518759bc8p+0, 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0,
    0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, 0x1.172b83c7d517bp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0,
    0x1.2d285a6e4030bp+0, 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0,
    0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, 0x1.3dea64c123422p+0,
    0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0,
    0x1.56f4736b527dap+0, 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0,
    0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, 0x1.6a09e667f3bcdp+0,
    0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0,
    0x1.868d99b4492edp+0, 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0,
    0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, 0x1.9c49182a3f090p+0,
    0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0,
    0x1.bcc1e904bc1d2p+0, 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0,
    0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, 0x1.d5818dcfba487p+0,
    0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0,
    0x1.fa7c1819e90d8p+0,
};
__constant__ double c_exp2_trail[64] = {
    0x0.0p+0, -0x1.19083535b085dp-56, 0x1.d73e2a475b465p-55,
    0x1.186be4bb284ffp-57, 0x1.8a62e4adc610bp-54, 0x1.03a1727c57b53p-59,
    -0x1.6c51039449b3ap-54, -0x1.32fbf9af1369ep-54, -0x1.19041b9d78a76p-55,
    0x1.e5b4c7b4968e4p-55, 0x1.e016e00a2643cp-54, 0x1.dc775814a8495p-55,
    0x1.9b07eb6c70573p-54, 0x1.2bd339940e9d9p-55, 0x1.612e8afad1255p-55,
    0x1.0024754db41d5p-54, 0x1.6f46ad23182e4p-55, 0x1.32721843659a6p-54,
    -0x1.63aeabf42eae2p-54, -0x1.5e436d661f5e3p-56, 0x1.ada0911f09ebcp-55,
    -0x1.ef3691c309278p-58, 0x1.89b7a04ef80d0p-59, 0x1.3c1a3b69062f0p-56,
    0x1.d4397afec42e2p-56, -0x1.4b309d25957e3p-54, -0x1.07abe1db13cadp-55,
    0x1.9bb2c011d93adp-54, 0x1.6324c054647adp-54, 0x1.ba6f93080e65ep-54,
    -0x1.383c17e40b497p-54, -0x1.bb60987591c34p-54, -0x1.bdd3413b26456p-54,
    -0x1.bbe3a683c88abp-57, -0x1.16e4786887a99p-55, -0x1.0245957316dd3p-54,
    -0x1.41577ee04992fp-55, 0x1.05d02ba15797ep-56, -0x1.d4c1dd41532d8p-54,
    -0x1.fc6f89bd4f6bap-54, 0x1.6e9f156864b27p-54, 0x1.5cc13a2e3976cp-55,
    -0x1.75fc781b57ebcp-57, -0x1.d185b7c1b85d1p-54, 0x1.c7c46b071f2bep-56,
    -0x1.359495d1cd533p-54, -0x1.d2f6edb8d41e1p-54, 0x1.0fac90ef7fd31p-54,
    0x1.7a1cd345dcc81p-54, -0x1.2805e3084d708p-57, -0x1.5584f7e54ac3bp-56,
    0x1.23dd07a2d9e84p-55, 0x1.11065895048ddp-55, 0x1.2884dff483cadp-54,
    0x1.503cbd1e949dbp-56, -0x1.cbc3743797a9cp-54, 0x1.2ed02d75b3707p-55,
    0x1.c2300696db532p-54, -0x1.1a5cd4f184b5cp-54, 0x1.39e8980a9cc8fp-55,
    -0x1.e9c23179c2893p-54, 0x1.dc7f486a4b6b0p-54, 0x1.9d3e12dd8a18bp-54,
    0x1.74853f3a5931ep-55,
};

// lead, trail: 2^(j/64) as lead + trail; sum: lead + trail, added once.
__device__ __forceinline__ double exp_tab(double x, const double *lead,
                                          const double *trail,
                                          const double *sum) {
  const double N = rint(x * EXP_INV_L);
  const int n = (int)N;
  const double R = fma(-N, EXP_L2, fma(-N, EXP_L1, x));
  // degree-5 Taylor polynomial: on |R| <= ln2/128 the next term is below
  // 3.5e-17 relative
  double q = fma(R, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(R, q, 1.0 / 6.0);
  q = fma(R, q, 0.5);
  const double p = fma(R * R, q, R);  // e^R - 1
  const int j = n & 63, m = n >> 6;   // n = 64 m + j, 0 <= j < 64
  const double sl = lead[j], st = trail[j];
  return ldexp(sl + fma(sum[j], p, st), m);
}

// 1/sqrt(x) for a positive, finite x: the hardware estimate y0 with one
// second-order (Householder) refinement y0 (1 + t/2 + 3t^2/8),
// t = 1 - x y0^2.  The library's rsqrt takes the same steps and then
// re-selects y0 for zero, infinite and NaN estimates, a test a squared
// distance to an atom never needs (three instructions per term saved).
__device__ __forceinline__ double rsqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double t = fma(y * -x, y, 1.0);
  return fma(y * t, fma(t, 0.375, 0.5), y);
}

constexpr int FACE_NT = 256, FACE_NODES = 64, FACE_SPLIT = FACE_NT / FACE_NODES;

// Atoms [t0, t1) of the staged chunk for the node pair at in-plane
// coordinates (va, vb) of axis AX: su/sw hold the atoms' in-plane
// coordinates, sn the normal one.  The squared distance is summed in x, y, z
// order, as the serial restatement does.
template <int AX>
__device__ __forceinline__ void face_terms(int t0, int t1, double va, double vb,
                                           double lo, double hi, double kap,
                                           const double *su, const double *sw,
                                           const double *sn, const double *sc,
                                           const double *tl, const double *tt,
                                           const double *ts, double &s0,
                                           double &s1) {
#pragma clang fp contract(off)
  for (int t = t0; t < t1; t++) {
    const double du = va - su[t], dw = vb - sw[t];
    const double d0 = lo - sn[t], d1 = hi - sn[t];
    const double d0s = d0 * d0, d1s = d1 * d1, dus = du * du, dws = dw * dw;
    const double r0 = AX == 0 ? d0s + dus + dws
                              : (AX == 1 ? dus + d0s + dws : dus + dws + d0s);
    const double r1 = AX == 0 ? d1s + dus + dws
                              : (AX == 1 ? dus + d1s + dws : dus + dws + d1s);
    const double i0 = rsqrt_pos(r0), i1 = rsqrt_pos(r1);
    s0 = fma(sc[t] * exp_tab(-kap * (r0 * i0), tl, tt, ts), i0, s0);
    s1 = fma(sc[t] * exp_tab(-kap * (r1 * i1), tl, tt, ts), i1, s1);
  }
}

// Debye-Hueckel faces.  A workgroup takes 64 boundary node pairs of one
// axis: each lane owns the node (a, b) on the low face and on the high face
// of that axis, which share the two in-plane coordinate differences to
// every atom.  Its four wavefronts each sum a quarter of every staged chunk
// of atoms (all lanes of a wavefront read the same atom: an LDS broadcast),
// and the four partial sums are added in a fixed order.  Per term the atom
// factor lambda_B q exp(kappa r) / (eps_w (1 + kappa r)) is staged once,
// leaving one exp (exp_tab) and one reciprocal square root per (node,
// atom).  The sum is deterministic; its order differs from the serial one
// by rounding only.
__global__ __launch_bounds__(FACE_NT) void k_pb_faces(
    int n, double c0, double h, double lb, double kap, double eps_out,
    int natoms,
    const double *__restrict__ atoms, double *gx, double *gy, double *gz) {
#pragma clang fp contract(off)
  __shared__ double sx[FACE_NT], sy[FACE_NT], sz[FACE_NT], sc[FACE_NT];
  __shared__ double part[2][FACE_SPLIT][FACE_NODES];
  // exp_tab's tables (read after a barrier)
  __shared__ double tl[64], tt[64], ts[64];
  if (threadIdx.x < 64) {
    tl[threadIdx.x] = c_exp2_lead[threadIdx.x];
    tt[threadIdx.x] = c_exp2_trail[threadIdx.x];
    ts[threadIdx.x] = c_exp2_lead[threadIdx.x] + c_exp2_trail[threadIdx.x];
  }
  const int axis = blockIdx.y;
  const int lane = threadIdx.x % FACE_NODES, split = threadIdx.x / FACE_NODES;
  const long long nn = (long long)n * n;
  const long long idx = (long long)blockIdx.x * FACE_NODES + lane;
  const bool valid = idx < nn;
  const int a = valid ? (int)(idx % n) : 0, b = valid ? (int)(idx / n) : 0;
  const double lo = c0, hi = c0 + (n - 1) * h;
  const double va = c0 + a * h, vb = c0 + b * h;
  double s0 = 0.0, s1 = 0.0;
  for (int base = 0; base < natoms; base += FACE_NT) {
    const int m = min(FACE_NT, natoms - base);
    __syncthreads();
    if ((int)threadIdx.x < m) {
      const double *A = atoms + 5 * (size_t)(base + threadIdx.x);
      sx[threadIdx.x] = A[0];
      sy[threadIdx.x] = A[1];
      sz[threadIdx.x] = A[2];
      sc[threadIdx.x] = lb * A[4] / (eps_out * (1.0 + kap * A[3])) *
                        exp(kap * A[3]);
    }
    __syncthreads();
    const int per = (m + FACE_SPLIT - 1) / FACE_SPLIT;
    const int t0 = split * per, t1 = min(m, t0 + per);
    if (axis == 0)
      face_terms<0>(t0, t1, va, vb, lo, hi, kap, sy, sz, sx, sc, tl, tt, ts, s0,
                    s1);
    else if (axis == 1)
      face_terms<1>(t0, t1, va, vb, lo, hi, kap, sx, sz, sy, sc, tl, tt, ts, s0,
                    s1);
    else
      face_terms<2>(t0, t1, va, vb, lo, hi, kap, sx, sy, sz, sc, tl, tt, ts, s0,
                    s1);
  }
  part[0][split][lane] = s0;
  part[1][split][lane] = s1;
  __syncthreads();
  if (split != 0 || !valid) return;
  double *dst = axis == 0 ? gx : (axis == 1 ? gy : gz);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    double tot = part[side][0][lane];
#pragma unroll
    for (int q = 1; q < FACE_SPLIT; q++) tot += part[side][q][lane];
    dst[(size_t)side * nn + (size_t)b * n + a] = tot;
  }
}

void check(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    fprintf(stderr, "pmg: launch of %s failed: %s\n", what,
            hipGetErrorString(e));
}

}  // namespace

void pbe_constants(double ionic, double eps_out, double *lb, double *kap2) {
  const double kB = 1.380649e-23, qe = 1.602176634e-19,
               eps0 = 8.8541878128e-12, NA = 6.02214076e23, Tk = 298.15;
  *lb = qe * qe / (4.0 * PI_D * eps0 * kB * Tk) * 1e10;
  *kap2 = 2.0 * ionic * 1000.0 * NA * qe * qe / (eps0 * eps_out * kB * Tk) *
          1e-20;
}

size_t pbe_scratch_bytes(int n) {
  const size_t N = (size_t)n * n * n;
  return N * sizeof(int) + 2 * N;
}

int pbe_assemble(int n, double h, double ionic, int natoms,
                 const double *atoms, int *next, void *scratch, double *xf,
                 double *yf, double *zf, double *gx, double *gy, double *gz,
                 double *a1, double *a2, double *a3, double *ccf, double *fcf,
                 hipStream_t s, double eps_in, double eps_out, double shell) {
  if (n < 3 || !(h > 0.0) || !(ionic >= 0.0) || natoms < 0 ||
      !(eps_in > 0.0) || !(eps_out > 0.0) || !(shell >= 0.0) ||
      (long long)n * n * n > (1LL << 31) - 1)
    return 2;
  double lb, kap2;
  pbe_constants(ionic, eps_out, &lb, &kap2);
  const double kap = sqrt(kap2), kap2e = eps_out * kap2;
  const double c0 = -(n - 1) * h / 2.0, h3 = h * h * h;
  const double q4 = 4.0 * PI_D * lb;
  const size_t N = (size_t)n * n * n;
  int *head = (int *)scratch;
  uint8_t *inside = (uint8_t *)scratch + N * sizeof(int);
  uint8_t *excl = inside + N;
  k_pb_init<<<(unsigned)std::min<size_t>((N + 255) / 256, 8192), 256, 0, s>>>(
      N, head, inside, excl);
  check("k_pb_init");
  if (natoms > 0) {
    k_pb_mark<<<natoms, 256, 0, s>>>(n, c0, h, shell, natoms, atoms, head,
                                     next, inside, excl);
    check("k_pb_mark");
  }
  dim3 gn(cdivp((long long)n * n, 256), 1, n);
  k_pb_nodes<<<gn, dim3(64, 4), 0, s>>>(n, c0, h, q4, h3, kap2e, eps_in,
                                        eps_out, atoms, head,
                                        next, inside, excl, xf, yf, zf, a1,
                                        a2, a3, ccf, fcf);
  check("k_pb_nodes");
  dim3 gf(cdivp((long long)n * n, FACE_NODES), 3);
  k_pb_faces<<<gf, FACE_NT, 0, s>>>(n, c0, h, lb, kap, eps_out, natoms, atoms,
                                    gx, gy,
                                    gz);
  check("k_pb_faces");
  return 0;
}

}  // namespace pmg

// ---- C ABI on host arrays ----
#include "../../include/pmg_mi355x.h"

extern "C" int pmg_pbe_assemble_ex(
    int n, double h, double ionic_strength, double eps_in, double eps_out,
    double ion_shell, int natoms, const double *atoms, double *xf, double *yf,
    double *zf, double *gxcf, double *gycf, double *gzcf, double *a1cf,
    double *a2cf, double *a3cf, double *ccf, double *fcf, double *tcf) {
  if (n < 3 || !(h > 0.0) || !(ionic_strength >= 0.0) || natoms < 0 ||
      (natoms > 0 && !atoms) || (long long)n * n * n > (1LL << 31) - 1)
    return PMG_EBADPARM;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PMG_EDEVICE;
  const size_t N = (size_t)n * n * n, nn = (size_t)n * n;
  // device layout: 5 N-arrays, 3 coordinate arrays, 3 face pairs, atoms,
  // cell-list links, assembly scratch
  const size_t nd = 5 * N + 3 * (size_t)n + 6 * nn + 5 * (size_t)natoms;
  double *d = nullptr;
  void *scr = nullptr;
  int *nxt = nullptr;
  hipStream_t s = nullptr;
  int rc = PMG_OK;
  if (hipMalloc(&d, nd * sizeof(double)) != hipSuccess ||
      hipMalloc(&scr, pmg::pbe_scratch_bytes(n)) != hipSuccess ||
      hipMalloc(&nxt, sizeof(int) * (natoms > 0 ? natoms : 1)) != hipSuccess)
    rc = PMG_ENOMEM;
  if (!rc && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    rc = PMG_EDEVICE;
  if (!rc) {
    double *A1 = d, *A2 = A1 + N, *A3 = A2 + N, *CC = A3 + N, *FC = CC + N;
    double *X = FC + N, *Y = X + n, *Z = Y + n;
    double *GX = Z + n, *GY = GX + 2 * nn, *GZ = GY + 2 * nn;
    double *AT = GZ + 2 * nn;
    if (natoms > 0)
      hipMemcpyAsync(AT, atoms, sizeof(double) * 5 * natoms,
                     hipMemcpyHostToDevice, s);
    rc = pmg::pbe_assemble(n, h, ionic_strength, natoms, AT, nxt, scr, X, Y,
                           Z, GX, GY, GZ, A1, A2, A3, CC, FC, s, eps_in,
                           eps_out, ion_shell);
    if (rc) rc = PMG_EBADPARM;
    auto get = [&](double *dst, const double *src, size_t cnt) {
      if (dst)
        hipMemcpyAsync(dst, src, cnt * sizeof(double), hipMemcpyDeviceToHost,
                       s);
    };
    get(a1cf, A1, N), get(a2cf, A2, N), get(a3cf, A3, N), get(ccf, CC, N);
    get(fcf, FC, N), get(xf, X, n), get(yf, Y, n), get(zf, Z, n);
    get(gxcf, GX, 2 * nn), get(gycf, GY, 2 * nn), get(gzcf, GZ, 2 * nn);
    if (hipStreamSynchronize(s) != hipSuccess) rc = PMG_EDEVICE;
  }
  if (!rc && tcf)
    for (size_t p = 0; p < N; p++) tcf[p] = 0.0;
  if (s) hipStreamDestroy(s);
  hipFree(d);
  hipFree(scr);
  hipFree(nxt);
  return rc;
}

extern "C" int pmg_pbe_assemble(int n, double h, double ionic_strength,
                                int natoms, const double *atoms, double *xf,
                                double *yf, double *zf, double *gxcf,
                                double *gycf, double *gzcf, double *a1cf,
                                double *a2cf, double *a3cf, double *ccf,
                                double *fcf, double *tcf) {
  return pmg_pbe_assemble_ex(n, h, ionic_strength, pmg::EPS_IN, pmg::EPS_OUT,
                             pmg::ION_SHELL, natoms, atoms, xf, yf, zf, gxcf,
                             gycf, gzcf, a1cf, a2cf, a3cf, ccf, fcf, tcf);
}
