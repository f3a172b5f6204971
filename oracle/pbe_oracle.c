/*
 * pbe_oracle.c -- CPU oracle of the Poisson-Boltzmann input assembly.
 *
 * TEST INFRASTRUCTURE ONLY (see pmg_oracle.h).  Restates how the reference
 * turns an atom list into the fine-level arguments of the PMG driver
 * (PBE_pmgc_class.h:191-344), with the synthetic-molecule choices of the
 * benchmark workloads:
 *   grid      n^3 nodes x_i = c0 + i*h, c0 = -(n-1)*h/2 (ball centred);
 *   marking   node inside an atom sphere when |x - a|^2 <= r^2, inside the
 *             ion-exclusion shell when |x - a|^2 <= (r+2)^2 (the inclusive
 *             sphere test of find_dots_in_sphere, dot_finder.h:24-87, and
 *             the radius + ion-radius shell, PBE_pmgc_class.h:276-281);
 *   epsilon   2 at marked nodes, 78.54 elsewhere; face coefficient
 *             a1cf(i) = harmonic_average(eps(i), eps(i+1))
 *             (scientific_quantities.h:30-32), the node value on the last
 *             plane, where the driver never reads it;
 *   kappa     ccf = 0 inside the shell, eps_w * kappa^2 outside
 *             (PBE_pmgc_class.h:293-297);
 *   source    trilinear spreading of 4*pi*lambda_B*q (the reference's
 *             charge/epsilon) over the 8 corners of the atom's cell, atoms
 *             in list order, corners in bit order x,y,z
 *             (create_charge_distribution.h:16-61, PBE_pmgc_class.h:331-339);
 *             atoms whose cell leaves the grid are dropped;
 *   boundary  Debye-Hueckel sphere sum over all atoms in list order on the six
 *             faces (PBE_pmgc_class.h:173-189, Debye_Hueckel_functions.h:10-14).
 * Physical constants are CODATA 2018 at 298.15 K; lambda_B and kappa in A.
 * Every floating-point expression is written in the order the HIP
 * assembly kernels use, so results agree bit for bit except where exp()
 * enters (the boundary values).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pmg_oracle.h"

#define KB_SI 1.380649e-23
#define QE_SI 1.602176634e-19
#define EPS0_SI 8.8541878128e-12
#define NA_SI 6.02214076e23
#define T_K 298.15
#define EPS_IN 2.0
#define EPS_OUT 78.54
#define ION_SHELL 2.0
#define PI_D 3.14159265358979323846

void pbo_constants_eps(double ionic, double eps_out, double *lb,
                       double *kap2) {
  *lb = QE_SI * QE_SI / (4.0 * PI_D * EPS0_SI * KB_SI * T_K) * 1e10;
  *kap2 = 2.0 * ionic * 1000.0 * NA_SI * QE_SI * QE_SI /
          (EPS0_SI * eps_out * KB_SI * T_K) * 1e-20;
}

void pbo_constants(double ionic, double *lb, double *kap2) {
  pbo_constants_eps(ionic, EPS_OUT, lb, kap2);
}

static double harmonic_average(double x, double y) {
  return 2 * x * y / (x + y);
}

void pbo_assemble_ex(int n, double h, double ionic, double eps_in,
                     double eps_out, double shell, int natoms,
                     const double *atoms, double *xf, double *yf, double *zf,
                     double *gxcf, double *gycf, double *gzcf, double *a1cf,
                     double *a2cf, double *a3cf, double *ccf, double *fcf,
                     double *tcf) {
  double lb, kap2;
  pbo_constants_eps(ionic, eps_out, &lb, &kap2);
  const double kap = sqrt(kap2), kap2e = eps_out * kap2;
  const double c0 = -(n - 1) * h / 2.0, h3 = h * h * h;
  const size_t nn = (size_t)n * n, N = nn * n;
  for (int i = 0; i < n; i++) xf[i] = yf[i] = zf[i] = c0 + i * h;

  unsigned char *inside = calloc(N, 1), *excl = calloc(N, 1);
  for (int a = 0; a < natoms; a++) {
    const double *A = atoms + 5 * (size_t)a;
    const double r = A[3], re = r + shell;
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) { /* a superset of the shell's nodes */
      double l = floor((A[d] - re - c0) / h) - 1.0;
      double u = ceil((A[d] + re - c0) / h) + 1.0;
      if (!(l >= 0.0)) l = 0.0;
      if (!(u <= n - 1.0)) u = n - 1.0;
      lo[d] = (int)l, hi[d] = (int)u;
    }
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          double dx = (c0 + i * h) - A[0], dy = (c0 + j * h) - A[1],
                 dz = (c0 + k * h) - A[2];
          double d2 = dx * dx + dy * dy + dz * dz;
          size_t p = (size_t)k * nn + (size_t)j * n + i;
          if (d2 <= re * re) excl[p] = 1;
          if (d2 <= r * r) inside[p] = 1;
        }
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        size_t p = (size_t)k * nn + (size_t)j * n + i;
        double e0 = inside[p] ? eps_in : eps_out;
        a1cf[p] = i + 1 < n
                      ? harmonic_average(e0, inside[p + 1] ? eps_in : eps_out)
                      : e0;
        a2cf[p] = j + 1 < n
                      ? harmonic_average(e0, inside[p + n] ? eps_in : eps_out)
                      : e0;
        a3cf[p] = k + 1 < n
                      ? harmonic_average(e0, inside[p + nn] ? eps_in : eps_out)
                      : e0;
        ccf[p] = excl[p] ? 0.0 : kap2e;
        fcf[p] = 0.0;
        tcf[p] = 0.0;
      }
  free(inside);
  free(excl);

  const double q4 = 4.0 * PI_D * lb;
  for (int a = 0; a < natoms; a++) {
    const double *A = atoms + 5 * (size_t)a;
    double wl[3], wu[3];
    int c[3], ok = 1;
    for (int d = 0; d < 3; d++) {
      double g = (A[d] - c0) / h;
      double fl = floor(g);
      if (!(fl >= 0.0 && fl <= (double)(n - 2))) ok = 0;
      c[d] = ok ? (int)fl : 0;
      wl[d] = (c[d] + 1) - g;
      wu[d] = g - c[d];
    }
    if (!ok) continue;
    const double cq = q4 * A[4];
    for (int corner = 0; corner < 8; corner++) {
      int di = corner & 1, dj = (corner >> 1) & 1, dk = (corner >> 2) & 1;
      double w = 1.0;
      w *= di ? wu[0] : wl[0];
      w *= dj ? wu[1] : wl[1];
      w *= dk ? wu[2] : wl[2];
      size_t p = (size_t)(c[2] + dk) * nn + (size_t)(c[1] + dj) * n +
                 (size_t)(c[0] + di);
      fcf[p] += cq * w / h3;
    }
  }

  /* faces: gxcf(j,k,1|2), gycf(i,k,1|2), gzcf(i,j,1|2) */
  const double lo = c0, hi = c0 + (n - 1) * h;
  /* every node's sum runs over the atoms in list order, whatever thread
   * takes the node */
#pragma omp parallel for collapse(2) schedule(static)
  for (int f = 0; f < 6; f++)
    for (int b = 0; b < n; b++)
      for (int a = 0; a < n; a++) {
        double va = c0 + a * h, vb = c0 + b * h, X, Y, Z;
        double *dst;
        switch (f >> 1) {
          case 0: X = (f & 1) ? hi : lo, Y = va, Z = vb; dst = gxcf; break;
          case 1: X = va, Y = (f & 1) ? hi : lo, Z = vb; dst = gycf; break;
          default: X = va, Y = vb, Z = (f & 1) ? hi : lo; dst = gzcf; break;
        }
        double s = 0.0;
        for (int t = 0; t < natoms; t++) {
          const double *A = atoms + 5 * (size_t)t;
          double dx = X - A[0], dy = Y - A[1], dz = Z - A[2];
          double d = sqrt(dx * dx + dy * dy + dz * dz);
          s += lb * A[4] / (eps_out * (1.0 + kap * A[3])) *
               exp(-kap * (d - A[3])) / d;
        }
        dst[(size_t)(f & 1) * nn + (size_t)b * n + a] = s;
      }
}

void pbo_assemble(int n, double h, double ionic, int natoms,
                  const double *atoms, double *xf, double *yf, double *zf,
                  double *gxcf, double *gycf, double *gzcf, double *a1cf,
                  double *a2cf, double *a3cf, double *ccf, double *fcf,
                  double *tcf) {
  pbo_assemble_ex(n, h, ionic, EPS_IN, EPS_OUT, ION_SHELL, natoms, atoms, xf,
                  yf, zf, gxcf, gycf, gzcf, a1cf, a2cf, a3cf, ccf, fcf, tcf);
}
