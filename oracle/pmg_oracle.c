/*
 * pmg_oracle.c -- CPU oracle for the PMG (pmgc) hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pmg_oracle.h).  A fresh C restatement of
 * the reference's pmgc algorithm with APBS semantics; every function cites
 * the reference lines it follows.  Deviations from the literal reference
 * text are applied where the reference's driver files would otherwise not
 * run (each is described in a comment where it applies).
 *
 * Reductions are deterministic: partial sums per k-plane (i fastest, then
 * j), then a sequential sum over planes, independent of the thread count.
 */
#include "pmg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define IX(nx, ny, i, j, k)                                                   \
  ((size_t)((k)-1) * (size_t)(nx) * (size_t)(ny) + (size_t)((j)-1) * (nx) +   \
   (size_t)((i)-1))

/* ------------------------------------------------------------------------ */
/* Stencil direction tables                                                  */
/* ------------------------------------------------------------------------ */

/* Plane of the symmetric 27-point storage holding the coupling between p and
 * p+e for an "upper" direction e (ez>0, or ez==0&&ey>0, or ez==ey==0&&ex>0);
 * -1 for lower directions and e == 0.  Plane order as in
 * pmgc/matvecd.cc:185-189; the coupling meaning of each plane follows the
 * neighbour each plane multiplies in Vmatvec27_1s (matvecd.cc:137-175). */
static int upper_plane(int ex, int ey, int ez) {
  if (ez == 1) {
    static const int t[3][3] = {/* ey=-1 */ {13, 9, 12},
                                /* ey= 0 */ {7, 3, 6},
                                /* ey=+1 */ {11, 8, 10}};
    return t[ey + 1][ex + 1];
  }
  if (ez == 0 && ey == 1) return ex == 1 ? 4 : (ex == 0 ? 2 : 5);
  if (ez == 0 && ey == 0 && ex == 1) return 1;
  return -1;
}

/* Coupling a(p, p+e) >= 0 convention (the matrix entry is -a).  For a lower
 * direction it is stored at the neighbour under the mirrored direction. */
static inline double coupling(const double *ac, size_t n, int numdia,
                              size_t p, long long poff, int ex, int ey,
                              int ez) {
  int pl = upper_plane(ex, ey, ez);
  if (pl >= 0) {
    if (numdia == 7 && pl > 3) return 0.0;
    return ac[(size_t)pl * n + p];
  }
  pl = upper_plane(-ex, -ey, -ez);
  if (numdia == 7 && pl > 3) return 0.0;
  return ac[(size_t)pl * n + (size_t)((long long)p + poff)];
}

/* P plane index of coarse-to-fine offset d (pmgc/matvecd.cc:809-816). */
static int pplane(int dx, int dy, int dz) {
  static const int ip[3][3] = {/* dy=-1 */ {8, 2, 7},
                               /* dy= 0 */ {4, 0, 3},
                               /* dy=+1 */ {6, 1, 5}};
  int zb = dz == 0 ? 0 : (dz == 1 ? 1 : 2);
  return 9 * zb + ip[dy + 1][dx + 1];
}

/* The 26 neighbour terms of the 27-point operator in the order of
 * Vmatvec27_1s / Vgsrb27x (matvecd.cc:137-170, gsd.cc:174-205):
 * 8 in-plane (tmpO), 9 up (tmpU), 9 down (tmpD). */
static const int nb27[26][3] = {
    {0, 1, 0},   {0, -1, 0},  {1, 0, 0},   {-1, 0, 0},  {1, 1, 0},
    {-1, 1, 0},  {1, -1, 0},  {-1, -1, 0}, {0, 0, 1},   {0, 1, 1},
    {0, -1, 1},  {1, 0, 1},   {-1, 0, 1},  {1, 1, 1},   {-1, 1, 1},
    {1, -1, 1},  {-1, -1, 1}, {0, 0, -1},  {0, 1, -1},  {0, -1, -1},
    {1, 0, -1},  {-1, 0, -1}, {1, 1, -1},  {-1, 1, -1}, {1, -1, -1},
    {-1, -1, -1}};

/* Sum of the 26 couplings times x, split into the three reference partial
 * sums (tmpO, tmpU, tmpD). */
static inline void sum27(const double *ac, size_t n, int nx, int ny,
                         size_t p, const double *x, double *sO, double *sU,
                         double *sD) {
  double t[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < 26; q++) {
    int ex = nb27[q][0], ey = nb27[q][1], ez = nb27[q][2];
    long long off = (long long)ez * nx * ny + (long long)ey * nx + ex;
    double a = coupling(ac, n, 27, p, off, ex, ey, ez);
    t[q < 8 ? 0 : (q < 17 ? 1 : 2)] += a * x[(size_t)((long long)p + off)];
  }
  *sO = t[0];
  *sU = t[1];
  *sD = t[2];
}

/* ------------------------------------------------------------------------ */
/* Sizes and structure                                                       */
/* ------------------------------------------------------------------------ */

static int corsr(int n) { return (n - 1) / 2 + 1; } /* Vcorsr mgsubd.cc */

int pmgo_maxlev(int n1, int n2, int n3) { /* Vmaxlev mgsubd.cc:456-484 */
  int lev = 0, done = 0;
  do {
    lev++;
    int iden = 1 << (lev - 1);
    int c1 = (n1 - 1) / iden + 1, c2 = (n2 - 1) / iden + 1,
        c3 = (n3 - 1) / iden + 1;
    if ((c1 - 1) * iden != n1 - 1 || c1 <= 2) done = 1;
    if ((c2 - 1) * iden != n2 - 1 || c2 <= 2) done = 1;
    if ((c3 - 1) * iden != n3 - 1 || c3 <= 2) done = 1;
  } while (!done);
  return lev - 1;
}

int pmgo_mgsz(int mgcoar, int mgdisc, int mgsolv, int nx, int ny, int nz,
              int nlev, int *out) { /* Vmgsz mgdrvd.cc:505-601 */
  int nf = nx * ny * nz, narr = nf;
  int nxc = nx, nyc = ny, nzc = nz;
  for (int l = 2; l <= nlev; l++) {
    nxc = corsr(nxc);
    nyc = corsr(nyc);
    nzc = corsr(nzc);
    narr += nxc * nyc * nzc;
  }
  int nc = nxc * nyc * nzc, narrc = narr - nf;
  int nfop, ncop, nband = 0;
  if (mgdisc == 0) nfop = 4;
  else if (mgdisc == 1) nfop = 14;
  else return 2;
  if ((mgcoar == 0 || mgcoar == 1) && mgdisc == 0) ncop = 4;
  else if (mgcoar == 2) ncop = 14;
  else return 2;
  if (mgsolv == 1) {
    int nb = ((mgcoar == 0 || mgcoar == 1) && mgdisc == 0)
                 ? 1 + (nxc - 2) * (nyc - 2)
                 : 1 + (nxc - 2) * (nyc - 2) + (nxc - 2) + 1;
    nband = (nxc - 2) * (nyc - 2) * (nzc - 2) * nb;
  } else if (mgsolv != 0) {
    return 2;
  }
  int nrpc = 100 * (nlev + 1);
  out[0] = nxc;
  out[1] = nyc;
  out[2] = nzc;
  out[3] = nf;
  out[4] = nc;
  out[5] = narr;
  out[6] = narrc;
  out[7] = nrpc;
  out[8] = 50 * (nlev + 1);
  out[9] = 100 * (nlev + 1);
  out[10] = 2 * narr + nfop * nf + (27 + ncop) * narrc + nband + nrpc;
  out[11] = out[8] + out[9];
  return 0;
}

#define IZ(iz, a, lev) (iz)[((lev)-1) * 50 + ((a)-1)]

void pmgo_buildstr(int nx, int ny, int nz, int nlev, int *iz) {
  /* Vbuildstr mgsubd.cc:266-347 */
  int n = nx * ny * nz;
  for (int a = 1; a <= 11; a++) IZ(iz, a, 1) = 1;
  IZ(iz, 1, 2) = IZ(iz, 1, 1) + n;
  IZ(iz, 2, 2) = IZ(iz, 2, 1) + 4 * ny * nz;
  IZ(iz, 3, 2) = IZ(iz, 3, 1) + 4 * nx * nz;
  IZ(iz, 4, 2) = IZ(iz, 4, 1) + 4 * nx * ny;
  IZ(iz, 5, 2) = IZ(iz, 5, 1) + 100;
  IZ(iz, 6, 2) = IZ(iz, 6, 1) + 100;
  IZ(iz, 8, 2) = IZ(iz, 8, 1) + nx;
  IZ(iz, 9, 2) = IZ(iz, 9, 1) + ny;
  IZ(iz, 10, 2) = IZ(iz, 10, 1) + nz;
  for (int lev = 2; lev <= nlev; lev++) {
    nx = corsr(nx);
    ny = corsr(ny);
    nz = corsr(nz);
    n = nx * ny * nz;
    IZ(iz, 1, lev + 1) = IZ(iz, 1, lev) + n;
    IZ(iz, 2, lev + 1) = IZ(iz, 2, lev) + 4 * ny * nz;
    IZ(iz, 3, lev + 1) = IZ(iz, 3, lev) + 4 * nx * nz;
    IZ(iz, 4, lev + 1) = IZ(iz, 4, lev) + 4 * nx * ny;
    IZ(iz, 5, lev + 1) = IZ(iz, 5, lev) + 100;
    IZ(iz, 6, lev + 1) = IZ(iz, 6, lev) + 100;
    IZ(iz, 7, lev + 1) = IZ(iz, 7, lev) + 4 * n;
    IZ(iz, 8, lev + 1) = IZ(iz, 8, lev) + nx;
    IZ(iz, 9, lev + 1) = IZ(iz, 9, lev) + ny;
    IZ(iz, 10, lev + 1) = IZ(iz, 10, lev) + nz;
    IZ(iz, 11, lev) = IZ(iz, 11, lev - 1) + 27 * n;
  }
}

/* ------------------------------------------------------------------------ */
/* Fine operator: VbuildA_fv (buildAd.cc:62-212)                             */
/* ------------------------------------------------------------------------ */

void pmgo_buildA_fv(int nx, int ny, int nz, int ipkey, int *ipc, double *ac,
                    double *cc, double *fc, const double *xf, const double *yf,
                    const double *zf, const double *gxcf, const double *gycf,
                    const double *gzcf, const double *a1cf, const double *a2cf,
                    const double *a3cf, const double *ccf, const double *fcf) {
  size_t n = (size_t)nx * ny * nz;
  double *oC = ac, *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  ipc[9] = ipkey; /* VAT(ipc,10) */
  ipc[10] = 7;    /* VAT(ipc,11) */
  ipc[11] = 1;    /* VAT(ipc,12) */
  /* the reference's z-range uses ny (buildAd.cc:124); nz is used. */
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t p = IX(nx, ny, i, j, k);
        double hzm1 = zf[k - 1] - zf[k - 2], hz = zf[k] - zf[k - 1];
        double hym1 = yf[j - 1] - yf[j - 2], hy = yf[j] - yf[j - 1];
        double hxm1 = xf[i - 1] - xf[i - 2], hx = xf[i] - xf[i - 1];
        double cE = (hym1 + hy) * (hzm1 + hz) / (4.0 * hx);
        double cEm = (hym1 + hy) * (hzm1 + hz) / (4.0 * hxm1);
        double cN = (hxm1 + hx) * (hzm1 + hz) / (4.0 * hy);
        double cNm = (hxm1 + hx) * (hzm1 + hz) / (4.0 * hym1);
        double cU = (hxm1 + hx) * (hym1 + hy) / (4.0 * hz);
        double cUm = (hxm1 + hx) * (hym1 + hy) / (4.0 * hzm1);
        double cF = (hxm1 + hx) * (hym1 + hy) * (hzm1 + hz) / 8.0;
        size_t pxm = IX(nx, ny, i - 1, j, k), pym = IX(nx, ny, i, j - 1, k),
               pzm = IX(nx, ny, i, j, k - 1);
        double f = cF * fcf[p];
        cc[p] = cF * ccf[p];
        oC[p] = cE * a1cf[p] + cEm * a1cf[pxm] + cN * a2cf[p] +
                cNm * a2cf[pym] + cU * a3cf[p] + cUm * a3cf[pzm];
        int ike = (i == nx - 1) ? 0 : 1;
        oE[p] = ike * cE * a1cf[p];
        f += (1 - ike) * cE * a1cf[p] * gxcf[(j - 1) + (size_t)(k - 1) * ny +
                                             (size_t)ny * nz];
        int jke = (j == ny - 1) ? 0 : 1;
        oN[p] = jke * cN * a2cf[p];
        f += (1 - jke) * cN * a2cf[p] * gycf[(i - 1) + (size_t)(k - 1) * nx +
                                             (size_t)nx * nz];
        int kke = (k == nz - 1) ? 0 : 1;
        uC[p] = kke * cU * a3cf[p];
        f += (1 - kke) * cU * a3cf[p] * gzcf[(i - 1) + (size_t)(j - 1) * nx +
                                             (size_t)nx * ny];
        int ikw = (i == 2) ? 0 : 1;
        f += (1 - ikw) * cEm * a1cf[pxm] * gxcf[(j - 1) + (size_t)(k - 1) * ny];
        int jks = (j == 2) ? 0 : 1;
        f += (1 - jks) * cNm * a2cf[pym] * gycf[(i - 1) + (size_t)(k - 1) * nx];
        int kkd = (k == 2) ? 0 : 1;
        f += (1 - kkd) * cUm * a3cf[pzm] * gzcf[(i - 1) + (size_t)(j - 1) * nx];
        fc[p] = f;
      }
}

/* ------------------------------------------------------------------------ */
/* Prolongation: VbuildP (buildPd.cc:1118-1147)                              */
/* ------------------------------------------------------------------------ */

/* Operator-based weights (VbuildPb_op7 buildPd.cc:172-554, VbuildPb_op27
 * :578-1090).  The reference spells out 26 formulas per stencil; all of them
 * are one rule, applied here by enumeration.  For the fine point F = C_f + d
 * (C_f the fine image of coarse point C, d in {-1,0,1}^3):
 *   - axes with d_a == 0 are collapsed: couplings of F to neighbours that
 *     differ from F only along collapsed axes are subtracted from oC(F);
 *   - every neighbour G = F + e with e_a in {0, -d_a} on the non-collapsed
 *     axes (at least one nonzero) lies towards C, at offset d + e' of C
 *     (e' = e restricted to the non-collapsed axes); the coupling a(F,G),
 *     summed over the collapsed axes, is weighted by the already-known
 *     weight w(d + e') (w(0) = 1);
 *   - neighbours away from C (e_a == +d_a) carry weight 0.
 * w(d) = numerator / reduced diagonal.  Weights are computed in order of the
 * number of nonzero components of d, so every w(d + e') is known. */
static void opP_point(const double *ac, size_t nf, int numdia, int nxf,
                      int nyf, size_t pC, double w[3][3][3]) {
  for (int a = 0; a < 27; a++) (&w[0][0][0])[a] = 0.0;
  w[1][1][1] = 1.0;
  for (int nnz = 1; nnz <= 3; nnz++)
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int d[3] = {dx, dy, dz};
          if ((dx != 0) + (dy != 0) + (dz != 0) != nnz) continue;
          long long sx = 1, sy = nxf, sz = (long long)nxf * nyf;
          size_t pF = (size_t)((long long)pC + dx * sx + dy * sy + dz * sz);
          double den = ac[pF];
          double num = 0.0;
          for (int ez = -1; ez <= 1; ez++)
            for (int ey = -1; ey <= 1; ey++)
              for (int ex = -1; ex <= 1; ex++) {
                if (!ex && !ey && !ez) continue;
                int e[3] = {ex, ey, ez};
                int ok = 1, toward = 0;
                int t[3];
                for (int a = 0; a < 3; a++) {
                  if (d[a] == 0) {
                    t[a] = 0;
                  } else if (e[a] == 0) {
                    t[a] = d[a];
                  } else if (e[a] == -d[a]) {
                    t[a] = 0;
                    toward = 1;
                  } else {
                    ok = 0;
                  }
                }
                if (!ok) continue;
                long long off = ex * sx + ey * sy + ez * sz;
                double a = coupling(ac, nf, numdia, pF, off, ex, ey, ez);
                if (!toward) {
                  /* only collapsed axes move: part of the reduced diagonal */
                  den -= a;
                } else {
                  num += a * w[t[2] + 1][t[1] + 1][t[0] + 1];
                }
              }
          w[dz + 1][dy + 1][dx + 1] = num / den;
        }
}

void pmgo_buildP(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 int mgprol, int numdia, const double *acf, double *pc) {
  size_t nf = (size_t)nxf * nyf * nzf, nc = (size_t)nxc * nyc * nzc;
  (void)nzf;
#pragma omp parallel for schedule(static)
  for (int kk = 2; kk <= nzc - 1; kk++)
    for (int jj = 2; jj <= nyc - 1; jj++)
      for (int ii = 2; ii <= nxc - 1; ii++) {
        size_t pc_ = IX(nxc, nyc, ii, jj, kk);
        double w[3][3][3];
        if (mgprol == 0) { /* VbuildPb_trilin buildPd.cc:116-151 */
          for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dx = -1; dx <= 1; dx++) {
                int m = (dx != 0) + (dy != 0) + (dz != 0);
                w[dz + 1][dy + 1][dx + 1] = 1.0 / (double)(1 << m);
              }
        } else {
          size_t pF = IX(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1);
          opP_point(acf, nf, numdia, nxf, nyf, pF, w);
        }
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++)
              pc[(size_t)pplane(dx, dy, dz) * nc + pc_] =
                  w[dz + 1][dy + 1][dx + 1];
      }
}

/* ------------------------------------------------------------------------ */
/* Galerkin coarse operator: VbuildG (buildGd.cc:3276-3368)                  */
/* ------------------------------------------------------------------------ */

/* The reference's VbuildG_7 / VbuildG_27 are expanded (Maple-generated)
 * forms of XA(C, C+delta) = sum_p P_C(p) (A P_{C+delta})(p) over the 27-point
 * supports of P; stored off-diagonals are -XA (buildGd.cc:527-1360,
 * 1449-3273).  Here the same sum is taken through w = A P_C (A symmetric):
 *   w(r) = sum_d A(C_f + r, C_f + d) P_C(d)      (d in [-1,1]^3, dz, dy, dx
 *                                                 order, dx fastest),
 *   XA(C, C+delta) = sum_q P_{C+delta}(q) w(2 delta + q),
 * offsets from C_f, the fine point of C.  The second sum is grouped by the
 * fine plane r_z = 2 delta_z + q_z (-1 .. 2 for the stored directions, which
 * have delta_z in {0, 1}): within a plane over (r_y, r_x) ascending, then the
 * four plane sums added in plane order.  The Maple forms fix no summation
 * order of their own that a restatement could follow term for term; this
 * grouping is the one the GPU kernel uses (one wavefront per plane), so the
 * two agree to the last bit where neither contracts multiply-adds.  The
 * prolongation of the next level divides differences of these entries that
 * cancel where the row sums vanish, so any other grouping moves the deeper
 * levels' P and Galerkin entries by ~1e-11 relative. */
void pmgo_buildG(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 int numdia, const double *pc, const double *acf,
                 double *acc) {
  static const int g_delta[14][3] = {
      {0, 0, 0},  {1, 0, 0},  {0, 1, 0},  {0, 0, 1},   {1, 1, 0},
      {-1, 1, 0}, {1, 0, 1},  {-1, 0, 1}, {0, 1, 1},   {0, -1, 1},
      {1, 1, 1},  {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}};
  size_t nf = (size_t)nxf * nyf * nzf, nc = (size_t)nxc * nyc * nzc;
  (void)nzf;
  long long sx = 1, sy = nxf, sz = (long long)nxf * nyf;
  long long cx = 1, cy = nxc, cz = (long long)nxc * nyc;
#pragma omp parallel for schedule(static)
  for (int kk = 2; kk <= nzc - 1; kk++)
    for (int jj = 2; jj <= nyc - 1; jj++)
      for (int ii = 2; ii <= nxc - 1; ii++) {
        size_t pC = IX(nxc, nyc, ii, jj, kk);
        size_t pCf = IX(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1);
        double w[5][5][5];
        int has[5][5][5];
        for (int rz = -2; rz <= 2; rz++)
          for (int ry = -2; ry <= 2; ry++)
            for (int rx = -2; rx <= 2; rx++) {
              size_t pF = (size_t)((long long)pCf + rx * sx + ry * sy + rz * sz);
              double v = 0.0;
              int any = 0;
              for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                  for (int dx = -1; dx <= 1; dx++) {
                    int ex = dx - rx, ey = dy - ry, ez = dz - rz;
                    if (ex < -1 || ex > 1 || ey < -1 || ey > 1 || ez < -1 ||
                        ez > 1)
                      continue;
                    int ne = (ex != 0) + (ey != 0) + (ez != 0);
                    if (numdia == 7 && ne > 1) continue;
                    any = 1;
                    double A = ne == 0 ? acf[pF]
                                       : -coupling(acf, nf, numdia, pF,
                                                   ex * sx + ey * sy + ez * sz,
                                                   ex, ey, ez);
                    v += A * pc[(size_t)pplane(dx, dy, dz) * nc + pC];
                  }
              w[rz + 2][ry + 2][rx + 2] = v;
              has[rz + 2][ry + 2][rx + 2] = any;
            }
        for (int s = 0; s < 14; s++) {
          const int *dl = g_delta[s];
          size_t pC2 = (size_t)((long long)pC + dl[0] * cx + dl[1] * cy +
                                dl[2] * cz);
          double plane[4];
          for (int rz = -1; rz <= 2; rz++) {
            double t = 0.0;
            for (int ry = -2; ry <= 2; ry++)
              for (int rx = -2; rx <= 2; rx++) {
                int qx = rx - 2 * dl[0], qy = ry - 2 * dl[1],
                    qz = rz - 2 * dl[2];
                if (qx < -1 || qx > 1 || qy < -1 || qy > 1 || qz < -1 || qz > 1)
                  continue;
                if (!has[rz + 2][ry + 2][rx + 2]) continue;
                t += pc[(size_t)pplane(qx, qy, qz) * nc + pC2] *
                     w[rz + 2][ry + 2][rx + 2];
              }
            plane[rz + 1] = t;
          }
          double v = ((plane[0] + plane[1]) + plane[2]) + plane[3];
          acc[(size_t)s * nc + pC] = (s == 0) ? v : -v;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Matvec / residual (matvecd.cc)                                           */
/* ------------------------------------------------------------------------ */

void pmgo_matvec(int nx, int ny, int nz, int numdia, const double *ac,
                 const double *cc, const double *x, double *y) {
  size_t n = (size_t)nx * ny * nz;
  const double *oC = ac, *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  size_t sy = nx, sz = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t p = IX(nx, ny, i, j, k);
        if (numdia == 7) { /* Vmatvec7_1s matvecd.cc:76-89 */
          y[p] = -oN[p] * x[p + sy] - oN[p - sy] * x[p - sy] -
                 oE[p] * x[p + 1] - oE[p - 1] * x[p - 1] -
                 uC[p - sz] * x[p - sz] - uC[p] * x[p + sz] +
                 (oC[p] + cc[p]) * x[p];
        } else { /* Vmatvec27_1s matvecd.cc:137-175 */
          double sO, sU, sD;
          sum27(ac, n, nx, ny, p, x, &sO, &sU, &sD);
          y[p] = -sO + -sU + -sD + (oC[p] + cc[p]) * x[p];
        }
      }
}

void pmgo_mresid(int nx, int ny, int nz, int numdia, const double *ac,
                 const double *cc, const double *fc, const double *x,
                 double *r) {
  size_t n = (size_t)nx * ny * nz;
  const double *oC = ac, *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  size_t sy = nx, sz = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t p = IX(nx, ny, i, j, k);
        if (numdia == 7) { /* Vmresid7_1s matvecd.cc:397-410 */
          r[p] = fc[p] + oN[p] * x[p + sy] + oN[p - sy] * x[p - sy] +
                 oE[p] * x[p + 1] + oE[p - 1] * x[p - 1] +
                 uC[p - sz] * x[p - sz] + uC[p] * x[p + sz] -
                 (oC[p] + cc[p]) * x[p];
        } else { /* Vmresid27_1s matvecd.cc:456-494 */
          double sO, sU, sD;
          sum27(ac, n, nx, ny, p, x, &sO, &sU, &sD);
          r[p] = fc[p] + sO + sU + sD - (oC[p] + cc[p]) * x[p];
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Ion terms (mypdec.cc)                                                     */
/* ------------------------------------------------------------------------ */

#define MAXIONS 50
static int g_nion = 0;
static double g_charge[MAXIONS], g_sconc[MAXIONS];

void pmgo_mypdefinit(int nion, const double *charge, const double *sconc) {
  /* Vmypdefinitlpbe / npbe mypdec.cc:72-106 */
  if (nion > MAXIONS) nion = MAXIONS;
  g_nion = nion;
  for (int i = 0; i < nion; i++) {
    g_charge[i] = charge[i];
    g_sconc[i] = sconc[i];
  }
}

/* Vc_vecpmg (mypdec.cc:136-212) and Vdc_vecpmg (:348-414), ipkey 0.
 * Evaluated on whole arrays exactly as the reference host loops. */
static void ion_term(const double *coef, const double *uin, double *uout,
                     size_t n, int deriv) {
  const double ZSMALL = 1.0e-20, ZLARGE = 1.0e20, SMIN = -85.0, SMAX = 85.0;
  for (size_t i = 0; i < n; i++) uout[i] = 0.0;
  for (int s = 0; s < g_nion; s++) {
    double zcf2 = deriv ? g_sconc[s] * g_charge[s] * g_charge[s]
                        : -1.0 * g_sconc[s] * g_charge[s];
    double zu2 = -1.0 * g_charge[s];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
      double az = fmin(ZSMALL, fabs(zcf2 * coef[i])) * ZLARGE;
      double an = fmax(fmin(zu2 * uin[i], 0.0), SMIN);
      double ap = fmin(fmax(zu2 * uin[i], 0.0), SMAX);
      uout[i] = uout[i] + zcf2 * coef[i] * exp(az * (an + ap));
    }
  }
}

void pmgo_c_vec(const double *coef, const double *uin, double *uout, int nx,
                int ny, int nz) {
  ion_term(coef, uin, uout, (size_t)nx * ny * nz, 0);
}
void pmgo_dc_vec(const double *coef, const double *uin, double *uout, int nx,
                 int ny, int nz) {
  ion_term(coef, uin, uout, (size_t)nx * ny * nz, 1);
}

void pmgo_nmresid(int nx, int ny, int nz, int numdia, const double *ac,
                  const double *cc, const double *fc, const double *x,
                  double *r, double *w1) {
  /* Vnmresid7_1s / Vnmresid27_1s (matvecd.cc:530-693) */
  size_t n = (size_t)nx * ny * nz;
  const double *oC = ac, *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  size_t sy = nx, sz = (size_t)nx * ny;
  pmgo_c_vec(cc, x, w1, nx, ny, nz);
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t p = IX(nx, ny, i, j, k);
        if (numdia == 7) {
          r[p] = fc[p] + oN[p] * x[p + sy] + oN[p - sy] * x[p - sy] +
                 oE[p] * x[p + 1] + oE[p - 1] * x[p - 1] +
                 uC[p - sz] * x[p - sz] + uC[p] * x[p + sz] -
                 oC[p] * x[p] - w1[p];
        } else {
          double sO, sU, sD;
          sum27(ac, n, nx, ny, p, x, &sO, &sU, &sD);
          r[p] = +sO + sU + sD + fc[p] - oC[p] * x[p] - w1[p];
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Smoothers (gsd.cc)                                                        */
/* ------------------------------------------------------------------------ */

void pmgo_gsrb(int nx, int ny, int nz, int numdia, const double *ac,
               const double *cc, const double *fc, double *x, int itmax,
               int iadjoint) {
  size_t n = (size_t)nx * ny * nz;
  const double *oC = ac, *oE = ac + n, *oN = ac + 2 * n, *uC = ac + 3 * n;
  size_t sy = nx, sz = (size_t)nx * ny;
  for (int it = 1; it <= itmax; it++) {
    if (numdia == 7) { /* Vgsrb7x gsd.cc:81-110 */
      for (int c = 0; c <= 1; c++) {
        int col = (1 - iadjoint) * c + iadjoint * (1 - c);
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 1; k++)
          for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
              if ((i + j + k) % 2 != col) continue;
              size_t p = IX(nx, ny, i, j, k);
              x[p] = (fc[p] + (oN[p] * x[p + sy] + oN[p - sy] * x[p - sy] +
                               oE[p] * x[p + 1] + oE[p - 1] * x[p - 1] +
                               uC[p - sz] * x[p - sz] + uC[p] * x[p + sz])) /
                     (oC[p] + cc[p]);
            }
      }
    } else { /* Vgsrb27x gsd.cc:160-211 (8 colours; the reference's driver
              * reaches it for 27-point levels) */
      for (int c = 0; c < 8; c++) {
        int col = c * (1 - iadjoint) + iadjoint * (7 - c);
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 1; k++)
          for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nx - 1; i++) {
              if ((i % 2) + 2 * (j % 2) + 4 * (k % 2) != col) continue;
              size_t p = IX(nx, ny, i, j, k);
              double sO, sU, sD;
              sum27(ac, n, nx, ny, p, x, &sO, &sU, &sD);
              x[p] = (fc[p] + (sO + sU + sD)) / (oC[p] + cc[p]);
            }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Grid transfer (matvecd.cc:695-1012)                                       */
/* ------------------------------------------------------------------------ */

void pmgo_restrc(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 const double *xin, double *xout, const double *pc) {
  /* Vrestrc2 matvecd.cc:755-795: xout(C) = sum_d P_C(d) xin(C_f + d),
   * summed as o-block, u-block, d-block, each in plane order. */
  size_t nc = (size_t)nxc * nyc * nzc;
  (void)nzf;
  static const int io[9][2] = {{0, 0},  {0, 1},  {0, -1}, {1, 0}, {-1, 0},
                               {1, 1},  {-1, 1}, {1, -1}, {-1, -1}};
  static const int zb[3] = {0, 1, -1};
#pragma omp parallel for schedule(static)
  for (int kk = 2; kk <= nzc - 1; kk++)
    for (int jj = 2; jj <= nyc - 1; jj++)
      for (int ii = 2; ii <= nxc - 1; ii++) {
        size_t pC = IX(nxc, nyc, ii, jj, kk);
        int i = 2 * ii - 1, j = 2 * jj - 1, k = 2 * kk - 1;
        double t[3];
        for (int b = 0; b < 3; b++) {
          double s = 0.0;
          for (int q = 0; q < 9; q++)
            s += pc[(size_t)(9 * b + q) * nc + pC] *
                 xin[IX(nxf, nyf, i + io[q][0], j + io[q][1], k + zb[b])];
          t[b] = s;
        }
        xout[pC] = t[0] + t[1] + t[2];
      }
}

void pmgo_interp(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                 const double *xin, double *xout, const double *pc) {
  /* VinterpPMG2 matvecd.cc:878-960, coarse cells ii=1..nxc-1. */
  size_t nc = (size_t)nxc * nyc * nzc;
  (void)nzf;
#define PW(pl, a, b, c) pc[(size_t)(pl)*nc + IX(nxc, nyc, a, b, c)]
#define XC(a, b, c) xin[IX(nxc, nyc, a, b, c)]
#define XF(a, b, c) xout[IX(nxf, nyf, a, b, c)]
  enum { oPC, oPN, oPS, oPE, oPW, oPNE, oPNW, oPSE, oPSW, uPC, uPN, uPS, uPE,
         uPW, uPNE, uPNW, uPSE, uPSW, dPC, dPN, dPS, dPE, dPW, dPNE, dPNW,
         dPSE, dPSW };
#pragma omp parallel for schedule(static)
  for (int kk = 1; kk <= nzc - 1; kk++)
    for (int jj = 1; jj <= nyc - 1; jj++)
      for (int ii = 1; ii <= nxc - 1; ii++) {
        int i = 2 * ii - 1, j = 2 * jj - 1, k = 2 * kk - 1;
        XF(i, j, k) = XC(ii, jj, kk);
        XF(i + 1, j, k) = PW(oPE, ii, jj, kk) * XC(ii, jj, kk) +
                          PW(oPW, ii + 1, jj, kk) * XC(ii + 1, jj, kk);
        XF(i, j + 1, k) = PW(oPN, ii, jj, kk) * XC(ii, jj, kk) +
                          PW(oPS, ii, jj + 1, kk) * XC(ii, jj + 1, kk);
        XF(i, j, k + 1) = PW(uPC, ii, jj, kk) * XC(ii, jj, kk) +
                          PW(dPC, ii, jj, kk + 1) * XC(ii, jj, kk + 1);
        XF(i + 1, j + 1, k) =
            PW(oPNE, ii, jj, kk) * XC(ii, jj, kk) +
            PW(oPNW, ii + 1, jj, kk) * XC(ii + 1, jj, kk) +
            PW(oPSE, ii, jj + 1, kk) * XC(ii, jj + 1, kk) +
            PW(oPSW, ii + 1, jj + 1, kk) * XC(ii + 1, jj + 1, kk);
        XF(i + 1, j, k + 1) =
            PW(uPE, ii, jj, kk) * XC(ii, jj, kk) +
            PW(uPW, ii + 1, jj, kk) * XC(ii + 1, jj, kk) +
            PW(dPE, ii, jj, kk + 1) * XC(ii, jj, kk + 1) +
            PW(dPW, ii + 1, jj, kk + 1) * XC(ii + 1, jj, kk + 1);
        XF(i, j + 1, k + 1) =
            PW(uPN, ii, jj, kk) * XC(ii, jj, kk) +
            PW(uPS, ii, jj + 1, kk) * XC(ii, jj + 1, kk) +
            PW(dPN, ii, jj, kk + 1) * XC(ii, jj, kk + 1) +
            PW(dPS, ii, jj + 1, kk + 1) * XC(ii, jj + 1, kk + 1);
        XF(i + 1, j + 1, k + 1) =
            +PW(uPNE, ii, jj, kk) * XC(ii, jj, kk) +
            PW(uPNW, ii + 1, jj, kk) * XC(ii + 1, jj, kk) +
            PW(uPSE, ii, jj + 1, kk) * XC(ii, jj + 1, kk) +
            PW(uPSW, ii + 1, jj + 1, kk) * XC(ii + 1, jj + 1, kk) +
            PW(dPNE, ii, jj, kk + 1) * XC(ii, jj, kk + 1) +
            PW(dPNW, ii + 1, jj, kk + 1) * XC(ii + 1, jj, kk + 1) +
            PW(dPSE, ii, jj + 1, kk + 1) * XC(ii, jj + 1, kk + 1) +
            PW(dPSW, ii + 1, jj + 1, kk + 1) * XC(ii + 1, jj + 1, kk + 1);
      }
#undef PW
#undef XC
#undef XF
  /* VfboundPMG00 on the result (matvecd.cc:963 is commented out). */
  pmgo_fbound00(nxf, nyf, nzf, xout);
}

void pmgo_extrac(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 const double *xin, double *xout) {
  /* Vextrac matvecd.cc:997-1008 */
  (void)nzf;
  for (int kk = 2; kk <= nzc - 1; kk++)
    for (int jj = 2; jj <= nyc - 1; jj++)
      for (int ii = 2; ii <= nxc - 1; ii++)
        xout[IX(nxc, nyc, ii, jj, kk)] =
            xin[IX(nxf, nyf, 2 * ii - 1, 2 * jj - 1, 2 * kk - 1)];
}

/* ------------------------------------------------------------------------ */
/* BLAS-1 and boundaries (mikpckd.cc)                                        */
/* ------------------------------------------------------------------------ */

/* Deterministic interior reduction: per-plane partials summed in order. */
static double reduce_interior(int nx, int ny, int nz, const double *x,
                              const double *y, int mode) {
  int npl = nz - 2 > 0 ? nz - 2 : 0;
  double *part = (double *)calloc((size_t)npl + 1, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; k++) {
    double s = 0.0;
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t p = IX(nx, ny, i, j, k);
        s += mode == 0 ? fabs(x[p]) : (mode == 1 ? x[p] * x[p] : x[p] * y[p]);
      }
    part[k - 2] = s;
  }
  double t = 0.0;
  for (int k = 0; k < npl; k++) t += part[k];
  free(part);
  return t;
}

double pmgo_xnrm1(int nx, int ny, int nz, const double *x) {
  return reduce_interior(nx, ny, nz, x, 0, 0); /* Vxnrm1 mikpckd.cc:133 */
}
double pmgo_xnrm2(int nx, int ny, int nz, const double *x) {
  return sqrt(reduce_interior(nx, ny, nz, x, 0, 1)); /* Vxnrm2 :161 */
}
double pmgo_xdot(int nx, int ny, int nz, const double *x, const double *y) {
  /* Vxdot mikpckd.cc:192-222: x.y (the reference squares x, :216) */
  return reduce_interior(nx, ny, nz, x, y, 2);
}
void pmgo_xaxpy(int nx, int ny, int nz, double alpha, const double *x,
                double *y) {
  /* Vxaxpy mikpckd.cc:115-131 over the interior range. */
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t p = IX(nx, ny, i, j, k);
        y[p] += alpha * x[p];
      }
}
void pmgo_xcopy(int nx, int ny, int nz, const double *x, double *y) {
  /* Vxcopy mikpckd.cc:61-76 (interior) */
#pragma omp parallel for schedule(static)
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        size_t p = IX(nx, ny, i, j, k);
        y[p] = x[p];
      }
}
static void xscal(int nx, int ny, int nz, double fac, double *x) {
  /* Vxscal mikpckd.cc:339-350 */
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) x[IX(nx, ny, i, j, k)] *= fac;
}

void pmgo_fbound00(int nx, int ny, int nz, double *x) {
  /* VfboundPMG00 mikpckd.cc:234-264 on the faces it intends. */
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      x[IX(nx, ny, 1, j, k)] = 0.0;
      x[IX(nx, ny, nx, j, k)] = 0.0;
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      x[IX(nx, ny, i, 1, k)] = 0.0;
      x[IX(nx, ny, i, ny, k)] = 0.0;
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      x[IX(nx, ny, i, j, 1)] = 0.0;
      x[IX(nx, ny, i, j, nz)] = 0.0;
    }
}

void pmgo_fbound(int ibound, int nx, int ny, int nz, double *x,
                 const double *gxc, const double *gyc, const double *gzc) {
  /* VfboundPMG mikpckd.cc:266-336: tests *ibound, on the faces it intends.
   * Faces are written i, then j, then k (edges take the later face). */
  if (ibound == 0) {
    pmgo_fbound00(nx, ny, nz, x);
    return;
  }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      x[IX(nx, ny, 1, j, k)] = gxc[(j - 1) + (size_t)(k - 1) * ny];
      x[IX(nx, ny, nx, j, k)] =
          gxc[(j - 1) + (size_t)(k - 1) * ny + (size_t)ny * nz];
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      x[IX(nx, ny, i, 1, k)] = gyc[(i - 1) + (size_t)(k - 1) * nx];
      x[IX(nx, ny, i, ny, k)] =
          gyc[(i - 1) + (size_t)(k - 1) * nx + (size_t)nx * nz];
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      x[IX(nx, ny, i, j, 1)] = gzc[(i - 1) + (size_t)(j - 1) * nx];
      x[IX(nx, ny, i, j, nz)] =
          gzc[(i - 1) + (size_t)(j - 1) * nx + (size_t)nx * ny];
    }
}

/* ------------------------------------------------------------------------ */
/* Coarsest-level CG: Vcghs (cgd.cc:60-124)                                  */
/* ------------------------------------------------------------------------ */

int pmgo_cghs(int nx, int ny, int nz, int numdia, const double *ac,
              const double *cc, const double *fc, double *x, double *p,
              double *ap, double *r, int itmax, double errtol) {
  int iters = 0;
  double rhok1 = 0.0;
  if (iters >= itmax) return iters;
  pmgo_mresid(nx, ny, nz, numdia, ac, cc, fc, x, r);
  double denom = pmgo_xnrm2(nx, ny, nz, r);
  if (denom == 0.0) return iters;
  for (;;) {
    double rhok2 = pmgo_xdot(nx, ny, nz, r, r); /* r . r */
    double rsnrm = sqrt(rhok2);
    if (rsnrm / denom <= errtol) break;
    if (iters >= itmax) break;
    if (iters == 0) {
      pmgo_xcopy(nx, ny, nz, r, p);
    } else {
      double beta = rhok2 / rhok1;
      pmgo_xaxpy(nx, ny, nz, 1.0 / beta, r, p);
      xscal(nx, ny, nz, beta, p);
    }
    pmgo_matvec(nx, ny, nz, numdia, ac, cc, p, ap);
    double pAp = pmgo_xdot(nx, ny, nz, p, ap);
    double alpha = rhok2 / pAp;
    rhok1 = rhok2;
    pmgo_xaxpy(nx, ny, nz, alpha, p, x);
    pmgo_xaxpy(nx, ny, nz, -alpha, ap, r);
    iters++;
  }
  return iters;
}

void pmgo_dpbfa(double *abd, int lda, int n, int m, int *info);
void pmgo_dpbsl(const double *abd, int lda, int n, int m, double *b);

/* ------------------------------------------------------------------------ */
/* Banded coarsest solve, mgsolv = 1: Vbuildband (buildBd.cc:62-304) and the  */
/* LINPACK routines Vddot / Vdpbfa / Vdpbsl (mlinpckd.cc:112-235).           */
/* Band storage: column-major abd(lda, n), 1-based; A(i,j), i <= j, lives at  */
/* abd(m + 1 + i - j, j).                                                    */
/* ------------------------------------------------------------------------ */

#define AB(abd, lda, r, c) (abd)[(size_t)((c)-1) * (lda) + ((r)-1)]

/* Vddot (mlinpckd.cc:112-165), unit strides: the n % 5 leading products are
 * summed one by one, then groups of five are summed and added. */
static double ddot5(int n, const double *x, const double *y) {
  double t = 0.0;
  if (n <= 0) return 0.0;
  int m = n % 5;
  for (int i = 0; i < m; i++) t += x[i] * y[i];
  if (n < 5) return t;
  for (int i = m; i < n; i += 5) {
    double g = x[i] * y[i];
    g += x[i + 1] * y[i + 1];
    g += x[i + 2] * y[i + 2];
    g += x[i + 3] * y[i + 3];
    g += x[i + 4] * y[i + 4];
    t += g;
  }
  return t;
}

int pmgo_buildband(int nx, int ny, int nz, int numdia, const double *ac,
                   double *abd, int *ipcb) {
  /* Vbuildband1_7 (buildBd.cc:62-124) / Vbuildband1_27 (:126-249): the
   * diagonal is oC alone (cc is not added); column jj (lexicographic
   * interior index) holds at offset jj - ii the negated plane values the
   * reference reads at (i-1,j,k), (i,j-1,k) or (i,j,k-1).  For the
   * diagonal directions these are the mirrored planes: e.g. point
   * (i+1,j-1,k) gets oNE(i,j-1,k), where Vmatvec27 (matvecd.cc:137-175)
   * couples the two through oNW(i+1,j-1,k). */
  size_t N = (size_t)nx * ny * nz;
  int sx = nx - 2, sxy = (nx - 2) * (ny - 2);
  int n = sxy * (nz - 2);
  int m = numdia == 7 ? sxy : sxy + sx + 1;
  int lda = m + 1;
  ipcb[0] = n;
  ipcb[1] = m;
  ipcb[2] = lda;
  ipcb[3] = 0;
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++) {
        int jj = (k - 2) * sxy + (j - 2) * sx + (i - 2) + 1;
#define PUT(off, val) AB(abd, lda, m + 1 - (off), jj) = (val)
        /* in the reference's order: later writes win where offsets
         * coincide (nx - 2 == 2); planes as pmgc/matvecd.cc:185-189 */
#define PL(pl, a, b, c) ac[(size_t)(pl) * N + IX(nx, ny, a, b, c)]
        PUT(0, PL(0, i, j, k));               /* oC */
        PUT(1, -PL(1, i - 1, j, k));          /* oE */
        PUT(sx, -PL(2, i, j - 1, k));         /* oN */
        if (numdia == 27) {
          PUT(sx - 1, -PL(4, i, j - 1, k));   /* oNE */
          PUT(sx + 1, -PL(5, i, j - 1, k));   /* oNW */
        }
        PUT(sxy, -PL(3, i, j, k - 1));        /* uC */
        if (numdia == 27) {
          PUT(sxy - 1, -PL(6, i, j, k - 1));       /* uE */
          PUT(sxy + 1, -PL(7, i, j, k - 1));       /* uW */
          PUT(sxy - sx, -PL(8, i, j, k - 1));      /* uN */
          PUT(sxy + sx, -PL(9, i, j, k - 1));      /* uS */
          PUT(sxy - sx - 1, -PL(10, i, j, k - 1)); /* uNE */
          PUT(sxy - sx + 1, -PL(11, i, j, k - 1)); /* uNW */
          PUT(sxy + sx - 1, -PL(12, i, j, k - 1)); /* uSE */
          PUT(sxy + sx + 1, -PL(13, i, j, k - 1)); /* uSW */
        }
#undef PL
#undef PUT
      }
  int info = 0;
  pmgo_dpbfa(abd, lda, n, m, &info); /* Vbuildband buildBd.cc:290-303 */
  ipcb[3] = 1;
  return info;
}

void pmgo_dpbfa(double *abd, int lda, int n, int m, int *info) {
  /* Vdpbfa (mlinpckd.cc:167-206): left-looking band Cholesky A = R^T R,
   * column j of R from the already factored columns j-m .. j-1. */
  *info = 0;
  for (int j = 1; j <= n; j++) {
    double s = 0.0;
    int ik = m + 1;
    int jk = j - m > 1 ? j - m : 1;
    int mu = m + 2 - j > 1 ? m + 2 - j : 1;
    for (int k = mu; k <= m; k++) {
      double t = AB(abd, lda, k, j) -
                 ddot5(k - mu, &AB(abd, lda, ik, jk), &AB(abd, lda, mu, j));
      t /= AB(abd, lda, m + 1, jk);
      AB(abd, lda, k, j) = t;
      s += t * t;
      ik--;
      jk++;
    }
    s = AB(abd, lda, m + 1, j) - s;
    if (s <= 0.0) {
      *info = j;
      return;
    }
    AB(abd, lda, m + 1, j) = sqrt(s);
  }
}

void pmgo_dpbsl(const double *abd, int lda, int n, int m, double *b) {
  /* Vdpbsl (mlinpckd.cc:208-235): R^T y = b by inner products, then
   * R x = y by column updates (Vdaxpy, element-wise). */
  for (int k = 1; k <= n; k++) {
    int lm = k - 1 < m ? k - 1 : m;
    int la = m + 1 - lm, lb = k - lm;
    double t = ddot5(lm, &AB(abd, lda, la, k), &b[lb - 1]);
    b[k - 1] = (b[k - 1] - t) / AB(abd, lda, m + 1, k);
  }
  for (int kb = 1; kb <= n; kb++) {
    int k = n + 1 - kb;
    int lm = k - 1 < m ? k - 1 : m;
    int la = m + 1 - lm, lb = k - lm;
    b[k - 1] /= AB(abd, lda, m + 1, k);
    double t = -b[k - 1];
    if (t != 0.0) /* Vdaxpy returns early for da == 0 */
      for (int i = 0; i < lm; i++)
        b[lb - 1 + i] += t * AB(abd, lda, la + i, k);
  }
}

/* Coarsest solve with the factored band (mgcsd.cc:391-407): the interior of
 * rhs is packed (Vxcopy_small), solved, unpacked into x (Vxcopy_large) and
 * the boundary of x zeroed (VfboundPMG00).  wk holds n doubles. */
void pmgo_bandsolve(int nx, int ny, int nz, const double *abd, const int *ipcb,
                    const double *rhs, double *x, double *wk) {
  int sx = nx - 2, sxy = (nx - 2) * (ny - 2);
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++)
        wk[(k - 2) * sxy + (j - 2) * sx + (i - 2)] = rhs[IX(nx, ny, i, j, k)];
  pmgo_dpbsl(abd, ipcb[2], ipcb[0], ipcb[1], wk);
  for (int k = 2; k <= nz - 1; k++)
    for (int j = 2; j <= ny - 1; j++)
      for (int i = 2; i <= nx - 1; i++)
        x[IX(nx, ny, i, j, k)] = wk[(k - 2) * sxy + (j - 2) * sx + (i - 2)];
  pmgo_fbound00(nx, ny, nz, x);
}

/* ------------------------------------------------------------------------ */
/* Hierarchy and V-cycle: Vmvcs (mgcsd.cc:60-530)                            */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nlev;
  int nx[32], ny[32], nz[32];
  size_t n[32];
  int numdia[32];
  double *ac[32], *cc[32], *fc[32], *pc[32]; /* pc[l]: P from l+1 to l */
  double *band;                              /* mgsolv 1: factored band */
  const int *ipcb;                           /* {n, m, lda, factored} */
  double *bwk;                               /* n doubles for the solve */
} hier_t;

static void hier_dims(hier_t *h, int nx, int ny, int nz, int nlev) {
  h->nlev = nlev;
  for (int l = 0; l < nlev; l++) {
    h->nx[l] = nx;
    h->ny[l] = ny;
    h->nz[l] = nz;
    h->n[l] = (size_t)nx * ny * nz;
    nx = corsr(nx);
    ny = corsr(ny);
    nz = corsr(nz);
  }
}

typedef struct {
  int istop, itmax, nu1, nu2, mgsolv;
  double errtol, epsiln;
} cyc_t;

static int g_last_iters = 0, g_last_ierror = 0;
static double g_last_relres = 0.0;
#define HIST_CAP 4096
static double g_hist[HIST_CAP];
static int g_nhist = 0;

/* Stopping-test norm (mgcsd.cc:483-526): 0, 1 the L1 residual; 2 the L1
 * difference to tru, which then takes x (successive iterates); 3, 4 the
 * 2-norm of tru - x; 5 its energy norm sqrt((tru-x) . A (tru-x)) (Vxdot
 * as a true dot product; the reference's squares x).  w2 is scratch for istop 5. */
static double stop_norm(const hier_t *h, const cyc_t *c, double *x,
                        const double *fc1, const double *cc1, double *w1,
                        double *w2, double *tru) {
  int nx = h->nx[0], ny = h->ny[0], nz = h->nz[0];
  if (c->istop == 0 || c->istop == 1) {
    pmgo_mresid(nx, ny, nz, h->numdia[0], h->ac[0], cc1, fc1, x, w1);
    return pmgo_xnrm1(nx, ny, nz, w1);
  }
  pmgo_xcopy(nx, ny, nz, tru, w1);
  pmgo_xaxpy(nx, ny, nz, -1.0, x, w1);
  if (c->istop == 5) {
    pmgo_matvec(nx, ny, nz, h->numdia[0], h->ac[0], cc1, w1, w2);
    return sqrt(pmgo_xdot(nx, ny, nz, w1, w2));
  }
  double v = c->istop == 2 ? pmgo_xnrm1(nx, ny, nz, w1)
                           : pmgo_xnrm2(nx, ny, nz, w1);
  if (c->istop == 2) pmgo_xcopy(nx, ny, nz, x, tru);
  return v;
}

/* x[l], w0[l]: per-level vectors; cc/fc at level 0 may be overridden (Newton
 * passes cprime / rhs).  Returns ierror. */
static int mvcs(const hier_t *h, const cyc_t *c, double **x, double **w0,
                double **ccl, const double *fc1, double *w1, double *w2,
                double *w3, double *tru, int *iters_out, int record) {
  int nlev = h->nlev;
  double rsden, rsnrm;
  int nx0 = h->nx[0], ny0 = h->ny[0], nz0 = h->nz[0];
  if (c->istop == 0) rsden = 1.0; /* mgcsd.cc:133-152 */
  else if (c->istop == 1) rsden = pmgo_xnrm1(nx0, ny0, nz0, fc1);
  else if (c->istop == 2) rsden = sqrt((double)nx0 * ny0 * nz0);
  else if (c->istop == 5) {
    pmgo_matvec(nx0, ny0, nz0, h->numdia[0], h->ac[0], ccl[0], tru, w1);
    rsden = sqrt(pmgo_xdot(nx0, ny0, nz0, tru, w1));
  } else rsden = pmgo_xnrm2(nx0, ny0, nz0, tru);
  if (rsden == 0.0) rsden = 1.0; /* mgcsd.cc:154-157 */
  rsnrm = rsden;
  if (record) g_nhist = 0;

  if (nlev == 1) { /* mgcsd.cc:170-291 */
    memset(x[0], 0, h->n[0] * sizeof(double));
    if (c->mgsolv == 0)
      pmgo_cghs(nx0, ny0, nz0, h->numdia[0], h->ac[0], ccl[0], fc1, x[0], w1,
                w2, w3, 100, c->epsiln);
    else
      pmgo_bandsolve(nx0, ny0, nz0, h->band, h->ipcb, fc1, x[0], h->bwk);
    *iters_out = 1;
    rsnrm = stop_norm(h, c, x[0], fc1, ccl[0], w1, w2, tru);
    if (record) {
      g_hist[0] = rsnrm / rsden;
      g_nhist = 1;
      g_last_relres = rsnrm / rsden;
    }
    return 0;
  }

  int iters = 0;
  do {
    /* fine pre-smoothing, then the residual into w1 (the reference's Vgsrb
     * drops its iresid output) */
    pmgo_gsrb(nx0, ny0, nz0, h->numdia[0], h->ac[0], ccl[0], fc1, x[0],
              c->nu1, 0);
    pmgo_mresid(nx0, ny0, nz0, h->numdia[0], h->ac[0], ccl[0], fc1, x[0], w1);
    pmgo_xcopy(nx0, ny0, nz0, w1, w0[0]);
    for (int l = 1; l < nlev; l++) { /* down: mgcsd.cc:325-360 */
      pmgo_restrc(h->nx[l - 1], h->ny[l - 1], h->nz[l - 1], h->nx[l],
                  h->ny[l], h->nz[l], w1, w0[l], h->pc[l - 1]);
      if (l != nlev - 1) {
        memset(x[l], 0, h->n[l] * sizeof(double));
        pmgo_gsrb(h->nx[l], h->ny[l], h->nz[l], h->numdia[l], h->ac[l],
                  ccl[l], w0[l], x[l], c->nu1, 0);
        pmgo_mresid(h->nx[l], h->ny[l], h->nz[l], h->numdia[l], h->ac[l],
                    ccl[l], w0[l], x[l], w1);
      }
    }
    { /* coarsest: Vcghs, 100 iterations, errtol = epsiln = 0, or the
       * factored band (mgcsd.cc:391-407) */
      int l = nlev - 1;
      memset(x[l], 0, h->n[l] * sizeof(double));
      if (c->mgsolv == 0)
        pmgo_cghs(h->nx[l], h->ny[l], h->nz[l], h->numdia[l], h->ac[l],
                  ccl[l], w0[l], x[l], w1, w2, w3, 100, c->epsiln);
      else
        pmgo_bandsolve(h->nx[l], h->ny[l], h->nz[l], h->band, h->ipcb, w0[l],
                       x[l], h->bwk);
    }
    for (int l = nlev - 2; l >= 0; l--) { /* up: mgcsd.cc:418-473 */
      int cl = l + 1;
      pmgo_interp(h->nx[cl], h->ny[cl], h->nz[cl], h->nx[l], h->ny[l],
                  h->nz[l], x[cl], w1, h->pc[l]);
      pmgo_matvec(h->nx[cl], h->ny[cl], h->nz[cl], h->numdia[cl], h->ac[cl],
                  ccl[cl], x[cl], w2);
      double xnum = pmgo_xdot(h->nx[cl], h->ny[cl], h->nz[cl], x[cl], w0[cl]);
      double xden = pmgo_xdot(h->nx[cl], h->ny[cl], h->nz[cl], x[cl], w2);
      /* a zero coarse correction (xden == 0) gives xdamp = 0 where the
       * reference divides 0/0 (mgcsd.cc:452) and spreads NaN. */
      double xdamp = xden != 0.0 ? xnum / xden : 0.0;
      pmgo_xaxpy(h->nx[l], h->ny[l], h->nz[l], xdamp, w1, x[l]);
      pmgo_gsrb(h->nx[l], h->ny[l], h->nz[l], h->numdia[l], h->ac[l], ccl[l],
                l == 0 ? fc1 : w0[l], x[l], c->nu2, 1);
    }
    iters++;
    rsnrm = stop_norm(h, c, x[0], fc1, ccl[0], w1, w2, tru); /* every cycle */
    if (record && g_nhist < HIST_CAP) g_hist[g_nhist++] = rsnrm / rsden;
  } while (iters < c->itmax && rsnrm / rsden > c->errtol);
  *iters_out = iters;
  if (record) g_last_relres = rsnrm / rsden;
  return iters < c->itmax ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* Drivers                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
  int nx, ny, nz, nlev, nu1, nu2, mgkey, itmax, istop, ipkey, mode, mgprol,
      mgcoar, mgdisc, mgsmoo, mgsolv, iperf;
  double errtol;
} params_t;

static int decode(const int *iparm, const double *rparm, params_t *p) {
  p->nx = iparm[2];
  p->ny = iparm[3];
  p->nz = iparm[4];
  p->nlev = iparm[5];
  p->nu1 = iparm[6];
  p->nu2 = iparm[7];
  p->mgkey = iparm[8];
  p->itmax = iparm[9];
  p->istop = iparm[10];
  p->ipkey = iparm[13];
  p->mode = iparm[15];
  p->mgprol = iparm[16];
  p->mgcoar = iparm[17];
  p->mgdisc = iparm[18];
  p->mgsmoo = iparm[19];
  p->mgsolv = iparm[20];
  p->iperf = iparm[21];
  p->errtol = rparm[0];
  if (p->nlev < 1 || p->nx < 3 || p->ny < 3 || p->nz < 3 || p->nlev > 30)
    return 2;
  if (p->nlev > pmgo_maxlev(p->nx, p->ny, p->nz) && p->nlev > 1) return 2;
  if (p->mgcoar != 2 || p->mgdisc != 0 || p->mgsmoo != 1 ||
      (p->mgsolv != 0 && p->mgsolv != 1) ||
      (p->mgprol != 0 && p->mgprol != 1) || p->istop < 0 || p->istop > 5 ||
      p->iperf != 0 || p->ipkey != 0 || p->nu1 < 0 || p->nu2 < 0)
    return 2;
  return 0;
}

/* Setup shared by both drivers: carve work arrays as Vmgdriv
 * (mgdrvd.cc:145-166), Vbuildstr, Vbuildops ido=0 and ido=1
 * (mgsubd.cc:63-264). */
static int setup(const params_t *p, int *iwork, double *rwork, int newton,
                 const double *xf, const double *yf, const double *zf,
                 const double *gxcf, const double *gycf, const double *gzcf,
                 const double *a1cf, const double *a2cf, const double *a3cf,
                 const double *ccf, const double *fcf, hier_t *h, double **cc,
                 double **fc, double **w1n, double **w2n) {
  int sz[12];
  if (pmgo_mgsz(p->mgcoar, p->mgdisc, p->mgsolv, p->nx, p->ny, p->nz, p->nlev,
                sz))
    return 2;
  int nf = sz[3], narr = sz[5], narrc = sz[6], nrpc = sz[7], niz = sz[8];
  int *iz = iwork, *ipc = iwork + niz;
  memset(iwork, 0, (size_t)sz[11] * sizeof(int));
  double *rpc = rwork;
  double *ccb = rpc + nrpc, *fcb = ccb + narr;
  double *after = fcb + narr;
  if (newton) { /* Vnewdriv adds w1, w2 (newdrvd.cc:126-133) */
    *w1n = after;
    *w2n = after + nf;
    after += 2 * (size_t)nf;
  }
  double *pcb = after, *acb = pcb + 27 * (size_t)narrc;
  size_t rtot = (size_t)sz[10] + (newton ? 2 * (size_t)nf : 0);
  memset(rwork, 0, rtot * sizeof(double)); /* load-bearing zeros */
  (void)rpc;
  pmgo_buildstr(p->nx, p->ny, p->nz, p->nlev, iz);
  hier_dims(h, p->nx, p->ny, p->nz, p->nlev);
  /* fine level */
  h->ac[0] = acb + (IZ(iz, 7, 1) - 1);
  h->cc[0] = ccb + (IZ(iz, 1, 1) - 1);
  h->fc[0] = fcb + (IZ(iz, 1, 1) - 1);
  pmgo_buildA_fv(p->nx, p->ny, p->nz, p->ipkey, ipc + IZ(iz, 5, 1) - 1,
                 h->ac[0], h->cc[0], h->fc[0], xf, yf, zf, gxcf, gycf, gzcf,
                 a1cf, a2cf, a3cf, ccf, fcf);
  h->numdia[0] = 7;
  IZ(iz, 7, 2) = IZ(iz, 7, 1) + 4 * (int)h->n[0];
  for (int l = 1; l < p->nlev; l++) {
    int lev = l + 1;
    h->pc[l - 1] = pcb + (IZ(iz, 11, lev - 1) - 1);
    h->ac[l] = acb + (IZ(iz, 7, lev) - 1);
    h->cc[l] = ccb + (IZ(iz, 1, lev) - 1);
    h->fc[l] = fcb + (IZ(iz, 1, lev) - 1);
    pmgo_buildP(h->nx[l - 1], h->ny[l - 1], h->nz[l - 1], h->nx[l], h->ny[l],
                h->nz[l], p->mgprol, h->numdia[l - 1], h->ac[l - 1],
                h->pc[l - 1]);
    pmgo_buildG(h->nx[l - 1], h->ny[l - 1], h->nz[l - 1], h->nx[l], h->ny[l],
                h->nz[l], h->numdia[l - 1], h->pc[l - 1], h->ac[l - 1],
                h->ac[l]);
    int *ipcl = ipc + IZ(iz, 5, lev) - 1;
    ipcl[10] = 27;
    ipcl[9] = p->ipkey;
    h->numdia[l] = 27;
    pmgo_restrc(h->nx[l - 1], h->ny[l - 1], h->nz[l - 1], h->nx[l], h->ny[l],
                h->nz[l], h->cc[l - 1], h->cc[l], h->pc[l - 1]);
    pmgo_restrc(h->nx[l - 1], h->ny[l - 1], h->nz[l - 1], h->nx[l], h->ny[l],
                h->nz[l], h->fc[l - 1], h->fc[l], h->pc[l - 1]);
    IZ(iz, 7, lev + 1) = IZ(iz, 7, lev) + 14 * (int)h->n[l];
  }
  h->band = 0;
  h->ipcb = 0;
  h->bwk = 0;
  if (p->mgsolv == 1) { /* Vbuildops mgsubd.cc:249-262 */
    int L = p->nlev;
    /* With one interior point in x or y, offsets of buildBd.cc:192-245 fall
     * on the diagonal or outside the band column (a write past the band in
     * the reference): refused. */
    if (h->numdia[L - 1] == 27 && (h->nx[L - 1] < 4 || h->ny[L - 1] < 4))
      return 2;
    int *ipcb = ipc + IZ(iz, 5, L + 1) - 1;
    h->band = acb + (IZ(iz, 7, L + 1) - 1);
    h->ipcb = ipcb;
    if (pmgo_buildband(h->nx[L - 1], h->ny[L - 1], h->nz[L - 1],
                       h->numdia[L - 1], h->ac[L - 1], h->band, ipcb))
      return 3; /* not positive definite: the reference exits here */
  }
  *cc = ccb;
  *fc = fcb;
  return 0;
}

static void level_ptrs(const hier_t *h, double *base, double **out) {
  size_t off = 0;
  for (int l = 0; l < h->nlev; l++) {
    out[l] = base + off;
    off += h->n[l];
  }
}

int pmgo_mgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                double *u, const double *xf, const double *yf,
                const double *zf, const double *gxcf, const double *gycf,
                const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                double *ccf, double *fcf, double *tcf) {
  params_t p;
  if (decode(iparm, rparm, &p)) return 2;
  if (p.mode != 0) return 2; /* mode 1/2 need FAS (out of scope) */
  hier_t h;
  double *ccb, *fcb;
  int rc = setup(&p, iwork, rwork, 0, xf, yf, zf, gxcf, gycf, gzcf, a1cf,
                 a2cf, a3cf, ccf, fcf, &h, &ccb, &fcb, 0, 0);
  if (rc) return rc;
  size_t narr = 0;
  for (int l = 0; l < h.nlev; l++) narr += h.n[l];
  size_t nf = h.n[0];
  /* private workspaces instead of clobbering a1cf..ccf, tcf */
  double *w0b = (double *)calloc(narr, sizeof(double));
  double *w1 = (double *)calloc(nf, sizeof(double));
  double *w2 = (double *)calloc(nf, sizeof(double));
  double *w3 = (double *)calloc(nf, sizeof(double));
  double *tru = (double *)calloc(narr, sizeof(double));
  memcpy(tru, tcf, nf * sizeof(double));
  /* Only the fine level of u is the caller's; coarse-level iterates live in
   * a private buffer (the reference's u is narr long, mgdrvd.cc:445). */
  double *xcb = (double *)calloc(narr, sizeof(double));
  double *bwk = (double *)calloc(h.n[h.nlev - 1], sizeof(double));
  h.bwk = bwk;
  double *x[32], *w0[32], *ccl[32];
  level_ptrs(&h, xcb, x);
  x[0] = u;
  level_ptrs(&h, w0b, w0);
  for (int l = 0; l < h.nlev; l++) ccl[l] = h.cc[l];
  if (p.istop == 4 || p.istop == 5) {
    /* Vbuildalg (mgdrvd.cc:424-433, mgsubd.cc:891-950): the right-hand side
     * of every level becomes A tcf, tcf carried to the coarse levels by
     * injection (Vextrac in Vbuildops), so that tcf solves the system */
    double *tl[32];
    level_ptrs(&h, tru, tl);
    for (int l = 1; l < h.nlev; l++)
      pmgo_extrac(h.nx[l - 1], h.ny[l - 1], h.nz[l - 1], h.nx[l], h.ny[l],
                  h.nz[l], tl[l - 1], tl[l]);
    for (int l = 0; l < h.nlev; l++)
      pmgo_matvec(h.nx[l], h.ny[l], h.nz[l], h.numdia[l], h.ac[l], h.cc[l],
                  tl[l], h.fc[l]);
  }
  memset(u, 0, nf * sizeof(double)); /* Vazeros mgdrvd.cc:419 */
  pmgo_fbound00(h.nx[0], h.ny[0], h.nz[0], u);
  cyc_t c = {p.istop, p.itmax, p.nu1, p.nu2, p.mgsolv, p.errtol, 0.0};
  int iters = 0;
  int ierr = mvcs(&h, &c, x, w0, ccl, h.fc[0], w1, w2, w3, tru, &iters, 1);
  pmgo_fbound(1, h.nx[0], h.ny[0], h.nz[0], u, gxcf, gycf, gzcf);
  g_last_iters = iters;
  g_last_ierror = ierr;
  free(w0b);
  free(w1);
  free(w2);
  free(w3);
  free(tru);
  free(xcb);
  free(bwk);
  (void)a1cf, (void)a2cf, (void)a3cf, (void)ccf, (void)fcf;
  return ierr;
}

/* ------------------------------------------------------------------------ */
/* Class-style solver: CG_solver (CG_Solver.h:45-242) preconditioned by the   */
/* class V-cycle v_cycle<false, nlev, true> (PBE_pmgc_class.h:1070-1133),     */
/* run here on the PMG hierarchy built by setup().                            */
/* ------------------------------------------------------------------------ */

/* Coarsest solve of the class V-cycle: Solver_CG with thresh_criterion(1e-8)
 * (PBE_pmgc_class.h:852-860, CG_Solver.h:258-276), started from x as it is
 * (the previous call's coarse solution).  The criterion is absolute on
 * sqrt(r.r); its first test is made on sqrt(sqrt(r0.r0)) (CG_Solver.h:
 * 125-133).  itmax caps the loop (the reference has no cap). */
int pmgo_cg_class(int nx, int ny, int nz, int numdia, const double *ac,
                  const double *cc, const double *f, double *x, double *p,
                  double *ap, double *r, double thresh, int itmax) {
  pmgo_mresid(nx, ny, nz, numdia, ac, cc, f, x, r);
  pmgo_xcopy(nx, ny, nz, r, p);
  double rs = pmgo_xdot(nx, ny, nz, r, r);
  pmgo_matvec(nx, ny, nz, numdia, ac, cc, p, ap);
  double pap = pmgo_xdot(nx, ny, nz, p, ap);
  if (pap == 0.0) return 0;
  double alpha = rs / pap;
  double residual = sqrt(sqrt(rs));
  int it = 0;
  while (residual > thresh && it < itmax) {
    it++;
    pmgo_xaxpy(nx, ny, nz, alpha, p, x);
    pmgo_xaxpy(nx, ny, nz, -alpha, ap, r);
    double rsn = pmgo_xdot(nx, ny, nz, r, r);
    double beta = rs == 0.0 ? 0.0 : rsn / rs;
    rs = rsn;
    xscal(nx, ny, nz, beta, p); /* p = r + beta p */
    pmgo_xaxpy(nx, ny, nz, 1.0, r, p);
    pmgo_matvec(nx, ny, nz, numdia, ac, cc, p, ap);
    pap = pmgo_xdot(nx, ny, nz, p, ap);
    alpha = pap == 0.0 ? 0.0 : rs / pap;
    residual = sqrt(rs);
  }
  return it;
}

/* One application of the class V-cycle as a preconditioner: x[0] = M rhs[0].
 * Per level (PBE_pmgc_class.h:1077-1113): two red-black sweeps from zero
 * (iadjoint 0), the defect restricted to the next level's right-hand side,
 * the recursion, the Hackbusch-Reusken factor (r.x)/(x.Ax) of the coarse
 * solution (MultigridDomain.h:329-353; 0 for a zero denominator), the
 * prolongation added with that factor, two sweeps again with iadjoint 0.
 * The coarsest level is pmgo_cg_class on its warm x. */
static void class_vcycle(const hier_t *h, int l, double **x, double **rhs,
                         double **ccl, double *w1, double *w2, double *w3) {
  int nx = h->nx[l], ny = h->ny[l], nz = h->nz[l];
  if (l == h->nlev - 1) {
    int ni = (nx - 2) * (ny - 2) * (nz - 2);
    pmgo_cg_class(nx, ny, nz, h->numdia[l], h->ac[l], ccl[l], rhs[l], x[l],
                  w1, w2, w3, 1e-8, ni * 4 > 1000 ? ni * 4 : 1000);
    return;
  }
  memset(x[l], 0, h->n[l] * sizeof(double));
  pmgo_gsrb(nx, ny, nz, h->numdia[l], h->ac[l], ccl[l], rhs[l], x[l], 2, 0);
  pmgo_mresid(nx, ny, nz, h->numdia[l], h->ac[l], ccl[l], rhs[l], x[l], w1);
  pmgo_restrc(nx, ny, nz, h->nx[l + 1], h->ny[l + 1], h->nz[l + 1], w1,
              rhs[l + 1], h->pc[l]);
  class_vcycle(h, l + 1, x, rhs, ccl, w1, w2, w3);
  int cl = l + 1;
  pmgo_matvec(h->nx[cl], h->ny[cl], h->nz[cl], h->numdia[cl], h->ac[cl],
              ccl[cl], x[cl], w2);
  double up = pmgo_xdot(h->nx[cl], h->ny[cl], h->nz[cl], rhs[cl], x[cl]);
  double lo = pmgo_xdot(h->nx[cl], h->ny[cl], h->nz[cl], x[cl], w2);
  double lambda = lo != 0.0 ? up / lo : 0.0;
  pmgo_interp(h->nx[cl], h->ny[cl], h->nz[cl], nx, ny, nz, x[cl], w1,
              h->pc[l]);
  pmgo_xaxpy(nx, ny, nz, lambda, w1, x[l]);
  pmgo_gsrb(nx, ny, nz, h->numdia[l], h->ac[l], ccl[l], rhs[l], x[l], 2, 0);
}

int pmgo_pcgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                 double *u, const double *xf, const double *yf,
                 const double *zf, const double *gxcf, const double *gycf,
                 const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                 double *ccf, double *fcf, double *tcf, int num_iters) {
  /* CG_solver (CG_Solver.h:45-242) with the class V-cycle as the
   * preconditioner (PBE_pmgc_class.h:1116-1133), on the Vmgdriv hierarchy:
   * x0 = 0, r = f - A x, s = M r, p = s, then per iteration
   * x += a p, r -= a A p, s = M r, b = (r.s)_new / (r.s), p = s + b p,
   * a = (r.s) / (p.A p).  num_iters > 0: counting_criterion (exactly that
   * many iterations, as PBE_pmgc_class_with_non_zero_IS.cc:129-136);
   * num_iters == 0: iterate until ||r||_1 / ||f||_1 <= errtol or itmax
   * iterations.  Returns 0, 1 (itmax, tolerance mode) or 2 (bad iparm). */
  params_t p;
  if (decode(iparm, rparm, &p)) return 2;
  if (p.mode != 0 || p.nlev < 2) return 2;
  hier_t h;
  double *ccb, *fcb;
  int rc = setup(&p, iwork, rwork, 0, xf, yf, zf, gxcf, gycf, gzcf, a1cf,
                 a2cf, a3cf, ccf, fcf, &h, &ccb, &fcb, 0, 0);
  if (rc) return rc;
  size_t narr = 0;
  for (int l = 0; l < h.nlev; l++) narr += h.n[l];
  size_t nf = h.n[0];
  int nx = h.nx[0], ny = h.ny[0], nz = h.nz[0];
  double *xb = (double *)calloc(narr, sizeof(double));
  double *rhsb = (double *)calloc(narr, sizeof(double));
  double *w1 = (double *)calloc(nf, sizeof(double));
  double *w2 = (double *)calloc(nf, sizeof(double));
  double *w3 = (double *)calloc(nf, sizeof(double));
  double *r = (double *)calloc(nf, sizeof(double));
  double *pv = (double *)calloc(nf, sizeof(double));
  double *q = (double *)calloc(nf, sizeof(double));
  double *x[32], *rhs[32], *ccl[32];
  level_ptrs(&h, xb, x);
  level_ptrs(&h, rhsb, rhs);
  for (int l = 0; l < h.nlev; l++) ccl[l] = h.cc[l];
  const double *fc1 = h.fc[0];
  memset(u, 0, nf * sizeof(double));
  double rsden = pmgo_xnrm1(nx, ny, nz, fc1);
  if (rsden == 0.0) rsden = 1.0;
  /* r = f - A x0; s = M r (the preconditioner's rhs is r, its result x[0]) */
  pmgo_mresid(nx, ny, nz, 7, h.ac[0], h.cc[0], fc1, u, r);
  pmgo_xcopy(nx, ny, nz, r, rhs[0]);
  class_vcycle(&h, 0, x, rhs, ccl, w1, w2, w3);
  pmgo_xcopy(nx, ny, nz, x[0], pv);
  double rs = pmgo_xdot(nx, ny, nz, r, x[0]);
  pmgo_matvec(nx, ny, nz, 7, h.ac[0], h.cc[0], pv, q);
  double pap = pmgo_xdot(nx, ny, nz, pv, q);
  int iters = 0, ierr = 0;
  g_nhist = 0;
  double rel = pmgo_xnrm1(nx, ny, nz, r) / rsden;
  if (pap != 0.0) {
    double alpha = rs / pap;
    for (;;) {
      if (num_iters > 0) {
        if (iters >= num_iters) break;
      } else {
        if (rel <= p.errtol) break;
        if (iters >= p.itmax) {
          ierr = 1;
          break;
        }
      }
      iters++;
      pmgo_xaxpy(nx, ny, nz, alpha, pv, u);
      pmgo_xaxpy(nx, ny, nz, -alpha, q, r);
      pmgo_xcopy(nx, ny, nz, r, rhs[0]);
      class_vcycle(&h, 0, x, rhs, ccl, w1, w2, w3);
      double rsn = pmgo_xdot(nx, ny, nz, r, x[0]);
      double beta = rs == 0.0 ? 0.0 : rsn / rs;
      rs = rsn;
      xscal(nx, ny, nz, beta, pv); /* p = s + beta p */
      pmgo_xaxpy(nx, ny, nz, 1.0, x[0], pv);
      pmgo_matvec(nx, ny, nz, 7, h.ac[0], h.cc[0], pv, q);
      pap = pmgo_xdot(nx, ny, nz, pv, q);
      alpha = pap == 0.0 ? 0.0 : rs / pap;
      rel = pmgo_xnrm1(nx, ny, nz, r) / rsden;
      if (g_nhist < HIST_CAP) g_hist[g_nhist++] = rel;
    }
  }
  /* the stopping measure reported: the true residual of the final u */
  pmgo_mresid(nx, ny, nz, 7, h.ac[0], h.cc[0], fc1, u, w1);
  g_last_relres = pmgo_xnrm1(nx, ny, nz, w1) / rsden;
  pmgo_fbound(1, nx, ny, nz, u, gxcf, gycf, gzcf);
  g_last_iters = iters;
  g_last_ierror = ierr;
  free(xb), free(rhsb), free(w1), free(w2), free(w3), free(r), free(pv),
      free(q);
  (void)a1cf, (void)a2cf, (void)a3cf, (void)ccf, (void)fcf, (void)tcf;
  return ierr;
}

/* Vnewton (newtond.cc:140-486) with Vgetjac (:488-531). */
/* Vnewton (newtond.cc:140-486) on the levels L..nlev-1 of h0 (level L as
 * the finest), from the iterate x of level L as given: damped inexact
 * Newton with the Vmvcs inner solve to min(0.9 |r|, |r|^2) (absolute L1,
 * istop 0, at most 1000 cycles), stopping test istop 0 / 1 on the L1
 * nonlinear residual.  record: keep the per-iteration history and the
 * final statistics (the finest Newton of a driver call). */
static int newton_at(const hier_t *h0, int L, int istop, int itmax,
                     double errtol, int nu1, int nu2, int mgsolv, double *x0,
                     int record) {
  hier_t h = *h0; /* levels L.. of h0 as levels 0.. */
  h.nlev = h0->nlev - L;
  for (int l = 0; l < h.nlev; l++) {
    h.nx[l] = h0->nx[L + l], h.ny[l] = h0->ny[L + l], h.nz[l] = h0->nz[L + l];
    h.n[l] = h0->n[L + l];
    h.numdia[l] = h0->numdia[L + l];
    h.ac[l] = h0->ac[L + l], h.cc[l] = h0->cc[L + l], h.fc[l] = h0->fc[L + l];
    h.pc[l] = l + 1 < h.nlev ? h0->pc[L + l] : 0;
  }
  int nx = h.nx[0], ny = h.ny[0], nz = h.nz[0], nd = h.numdia[0];
  size_t nf = h.n[0], narr = 0;
  for (int l = 0; l < h.nlev; l++) narr += h.n[l];
  double *w0b = (double *)calloc(narr, sizeof(double));
  double *w1 = (double *)calloc(nf, sizeof(double));
  double *w2 = (double *)calloc(nf, sizeof(double));
  double *w3 = (double *)calloc(nf, sizeof(double));
  double *tru = (double *)calloc(narr, sizeof(double));
  double *cprb = (double *)calloc(narr, sizeof(double));
  double *rhsb = (double *)calloc(narr, sizeof(double));
  double *xtmpb = (double *)calloc(narr, sizeof(double));
  double *t1 = (double *)calloc(nf, sizeof(double));
  double *t2 = (double *)calloc(nf, sizeof(double));
  double *w0[32], *cpr[32], *rhs[32], *xtmp[32];
  level_ptrs(&h, w0b, w0);
  level_ptrs(&h, cprb, cpr);
  level_ptrs(&h, rhsb, rhs);
  level_ptrs(&h, xtmpb, xtmp);
  const double *cc1 = h.cc[0], *fc1 = h.fc[0];
  double rsden;
  if (istop == 0) {
    rsden = 1.0;
  } else { /* istop 1: residual at a zero guess (newtond.cc:191-202) */
    memset(w1, 0, nf * sizeof(double));
    pmgo_nmresid(nx, ny, nz, nd, h.ac[0], cc1, fc1, w1, w2, w3);
    rsden = pmgo_xnrm1(nx, ny, nz, w2);
  }
  if (rsden == 0.0) rsden = 1.0;
  double rsnrm = rsden;
  pmgo_nmresid(nx, ny, nz, nd, h.ac[0], cc1, fc1, x0, w0[0], w2);
  double xnorm_old = pmgo_xnrm1(nx, ny, nz, w0[0]);
  double xnorm_new = xnorm_old, xnorm_med = xnorm_old;
  int idamp = 1, iters = 0;
  if (record) g_nhist = 0;
  for (;;) {
    iters++;
    /* Vgetjac: rhs = r, cprime = dc_vec(cc, x), restricted to all levels */
    pmgo_xcopy(nx, ny, nz, w0[0], rhs[0]);
    pmgo_dc_vec(cc1, x0, cpr[0], nx, ny, nz);
    for (int l = 1; l < h.nlev; l++) {
      pmgo_restrc(h.nx[l - 1], h.ny[l - 1], h.nz[l - 1], h.nx[l], h.ny[l],
                  h.nz[l], rhs[l - 1], rhs[l], h.pc[l - 1]);
      pmgo_restrc(h.nx[l - 1], h.ny[l - 1], h.nz[l - 1], h.nx[l], h.ny[l],
                  h.nz[l], cpr[l - 1], cpr[l], h.pc[l - 1]);
    }
    double errtol_s = fmin(0.9 * xnorm_old, 1.0 * pow(xnorm_old, 2.0));
    memset(xtmp[0], 0, nf * sizeof(double));
    cyc_t c = {0, 1000, nu1, nu2, mgsolv, errtol_s, 0.0};
    int its = 0;
    mvcs(&h, &c, xtmp, w0, cpr, rhs[0], w1, w2, w3, tru, &its, 0);
    if (idamp == 1) { /* newtond.cc:313-376 */
      pmgo_xcopy(nx, ny, nz, x0, w1);
      pmgo_xaxpy(nx, ny, nz, 1.0, xtmp[0], w1);
      pmgo_nmresid(nx, ny, nz, nd, h.ac[0], cc1, fc1, w1, w0[0], rhs[0]);
      xnorm_new = pmgo_xnrm1(nx, ny, nz, w0[0]);
      double damp = 1.0;
      int iter_d = 0, mode = 0;
      while (iter_d < 10) {
        if (mode == 0) {
          if (xnorm_new < xnorm_old) mode = 1;
        } else if (xnorm_new > xnorm_med) {
          break;
        }
        pmgo_xcopy(nx, ny, nz, w1, t1);
        pmgo_xcopy(nx, ny, nz, w0[0], t2);
        xnorm_med = xnorm_new;
        pmgo_xcopy(nx, ny, nz, x0, w1);
        damp = damp / 2.0;
        pmgo_xaxpy(nx, ny, nz, damp, xtmp[0], w1);
        pmgo_nmresid(nx, ny, nz, nd, h.ac[0], cc1, fc1, w1, w0[0], rhs[0]);
        xnorm_new = pmgo_xnrm1(nx, ny, nz, w0[0]);
        iter_d++;
      }
      pmgo_xcopy(nx, ny, nz, t1, x0);
      pmgo_xcopy(nx, ny, nz, t2, w0[0]);
      xnorm_new = xnorm_med;
      xnorm_old = xnorm_new;
      if (iter_d - 1 == 0) idamp = 0;
    } else {
      pmgo_xaxpy(nx, ny, nz, 1.0, xtmp[0], x0);
      pmgo_nmresid(nx, ny, nz, nd, h.ac[0], cc1, fc1, x0, w0[0], rhs[0]);
      xnorm_new = pmgo_xnrm1(nx, ny, nz, w0[0]);
      xnorm_old = xnorm_new;
    }
    rsnrm = xnorm_new; /* istop 0/1 (newtond.cc:396-401), every iteration */
    if (record && g_nhist < HIST_CAP) g_hist[g_nhist++] = rsnrm / rsden;
    if (rsnrm / rsden <= errtol) break;
    if (iters >= itmax) break;
  }
  const int ierr = rsnrm / rsden <= errtol ? 0 : 1;
  if (record) {
    g_last_iters = iters;
    g_last_ierror = ierr;
    g_last_relres = rsnrm / rsden;
  }
  free(w0b), free(w1), free(w2), free(w3), free(tru), free(cprb), free(rhsb),
      free(xtmpb), free(t1), free(t2);
  return ierr;
}

/* Vnewdriv / Vnewdriv2 (newdrvd.cc:60-300): mgkey 0 runs Vnewton on the
 * fine level from the caller's u (interior as given, boundary zeroed,
 * VfboundPMG00 newdrvd.cc:263); mgkey 1 runs Vfnewton (newtond.cc:60-138,
 * its level loop run upward as its level-- intends): Newton on the coarsest
 * level, prolongated (VinterpPMG) as the
 * starting guess of the next finer level's Newton, up to the finest level.
 * The coarse-level iterates start from zero (the reference reads them from
 * the coarse part of a narr-long u, which this interface does not carry). */
int pmgo_newdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                 double *u, const double *xf, const double *yf,
                 const double *zf, const double *gxcf, const double *gycf,
                 const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                 double *ccf, double *fcf, double *tcf) {
  params_t p;
  if (decode(iparm, rparm, &p)) return 2;
  if (p.mgkey != 0 && p.mgkey != 1) return 2;
  if (p.istop > 1) return 2; /* Newton: the residual tests only */
  hier_t h;
  double *ccb, *fcb, *w1n, *w2n;
  int rc = setup(&p, iwork, rwork, 1, xf, yf, zf, gxcf, gycf, gzcf, a1cf,
                 a2cf, a3cf, ccf, fcf, &h, &ccb, &fcb, &w1n, &w2n);
  if (rc) return rc;
  (void)w1n, (void)w2n, (void)tcf;
  int nx = h.nx[0], ny = h.ny[0], nz = h.nz[0];
  double *bwk = (double *)calloc(h.n[h.nlev - 1], sizeof(double));
  h.bwk = bwk;
  pmgo_fbound00(nx, ny, nz, u);
  if (p.mgkey == 1) {
    size_t narr = 0;
    for (int l = 0; l < h.nlev; l++) narr += h.n[l];
    double *xb = (double *)calloc(narr, sizeof(double)), *xl[32];
    level_ptrs(&h, xb, xl);
    xl[0] = u;
    for (int L = h.nlev - 1; L >= 1; L--) {
      newton_at(&h, L, p.istop, 1000, p.errtol, p.nu1, p.nu2, p.mgsolv, xl[L],
                0);
      pmgo_interp(h.nx[L], h.ny[L], h.nz[L], h.nx[L - 1], h.ny[L - 1],
                  h.nz[L - 1], xl[L], xl[L - 1], h.pc[L - 1]);
    }
    free(xb);
  }
  int ierr = newton_at(&h, 0, p.istop, p.itmax, p.errtol, p.nu1, p.nu2,
                       p.mgsolv, u, 1);
  pmgo_fbound(1, nx, ny, nz, u, gxcf, gycf, gzcf);
  free(bwk);
  (void)a1cf, (void)a2cf, (void)a3cf, (void)ccf, (void)fcf;
  return ierr;
}

int pmgo_last_stats(int *iters, int *ierror, double *relres, double *hist,
                    int hist_cap) {
  *iters = g_last_iters;
  *ierror = g_last_ierror;
  *relres = g_last_relres;
  int m = g_nhist < hist_cap ? g_nhist : hist_cap;
  for (int i = 0; i < m; i++) hist[i] = g_hist[i];
  return m;
}
