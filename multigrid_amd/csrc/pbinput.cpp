// pbinput.cpp -- seeded synthetic Poisson-Boltzmann inputs in the layout of
// the PMG driver's fine-level arguments (pmgc/mgdrvd.cc:60-65).
//
// The reference assembles these from PQR files (PBE_pmgc_class.h:191-344:
// Debye-Hueckel boundary :211-226, sphere marking :257-290, charge
// spreading :331-339).  Here a molecule is drawn from a seeded RNG:
//   atoms   uniform in a ball of radius R (direction = normalised Gaussian
//           triplet, radius R*U^(1/3)), radii U[1.2, 2.0] A;
//   charges mode 0: floor(N/2) of +1 and of -1 (0 for an odd remainder),
//           shuffled; mode 1: uniform over {-1,-0.5,0,0.5,1};
//   grid    n^3 nodes of spacing h centred on the ball centre;
//   eps     2 inside the union of atom spheres, 78.54 outside, at nodes;
//           face coefficient a1cf(i) = harmonic mean of eps(i), eps(i+1)
//           (the reference's harmonic_average, scientific_quantities.h:30);
//   kappa2  eps_w * kappa^2 outside radius+2 A of every atom, else 0;
//   source  4*pi*lambda_B * sum q_i * trilinear weight / h^3
//           (create_charge_distribution.h:16-61, charge/epsilon);
//   bc      Debye-Hueckel spheres summed on the six faces
//           (Debye_Hueckel_functions.h:10-14).
// RNG: SplitMix64 seeding xoshiro256**, doubles from the top 53 bits.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/pmg_mi355x.h"

namespace {

struct Xoshiro {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Xoshiro(uint64_t seed) {
    for (auto &v : s) v = splitmix(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double gauss() {  // Box-Muller, one value per call
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
  }
};

// Physical constants (CODATA 2018), SI.
constexpr double kB = 1.380649e-23, qe = 1.602176634e-19,
                 eps0 = 8.8541878128e-12, NA = 6.02214076e23, Tk = 298.15;
constexpr double EPS_IN = 2.0, EPS_OUT = 78.54;

template <class F>
void parallel_for(int n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if ((int)nt > n) nt = n > 0 ? n : 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (int i = (int)t; i < n; i += (int)nt) f(i);
    });
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" int pmg_pbe_synth(int n, double h, int natoms, double sphere_r,
                             double ionic_strength, int charge_mode,
                             uint64_t seed, double *xf, double *yf,
                             double *zf, double *gxcf, double *gycf,
                             double *gzcf, double *a1cf, double *a2cf,
                             double *a3cf, double *ccf, double *fcf,
                             double *tcf, double *atoms) {
  if (n < 3 || h <= 0.0 || natoms < 0 || sphere_r < 0.0 ||
      ionic_strength < 0.0 || (charge_mode != 0 && charge_mode != 1))
    return PMG_EBADPARM;
  Xoshiro rng(seed);
  std::vector<double> ax(natoms), ay(natoms), az(natoms), ar(natoms),
      aq(natoms);
  for (int a = 0; a < natoms; a++) {
    double gx = rng.gauss(), gy = rng.gauss(), gz = rng.gauss();
    double nrm = std::sqrt(gx * gx + gy * gy + gz * gz);
    if (nrm == 0.0) nrm = 1.0, gx = 1.0;
    double rad = sphere_r * std::cbrt(rng.uniform());
    ax[a] = gx / nrm * rad;
    ay[a] = gy / nrm * rad;
    az[a] = gz / nrm * rad;
    ar[a] = 1.2 + 0.8 * rng.uniform();
  }
  if (charge_mode == 0) {
    int half = natoms / 2;
    for (int a = 0; a < natoms; a++)
      aq[a] = a < half ? 1.0 : (a < 2 * half ? -1.0 : 0.0);
    for (int a = natoms - 1; a > 0; a--) {  // Fisher-Yates
      int b = (int)(rng.uniform() * (a + 1));
      if (b > a) b = a;
      std::swap(aq[a], aq[b]);
    }
  } else {
    static const double vals[5] = {-1.0, -0.5, 0.0, 0.5, 1.0};
    for (int a = 0; a < natoms; a++) {
      int b = (int)(rng.uniform() * 5.0);
      aq[a] = vals[b > 4 ? 4 : b];
    }
  }
  if (atoms)
    for (int a = 0; a < natoms; a++) {
      atoms[5 * a + 0] = ax[a];
      atoms[5 * a + 1] = ay[a];
      atoms[5 * a + 2] = az[a];
      atoms[5 * a + 3] = ar[a];
      atoms[5 * a + 4] = aq[a];
    }

  const double lB = qe * qe / (4.0 * M_PI * eps0 * kB * Tk) * 1e10;  // A
  // kappa^2 in A^-2 for a 1:1 salt of ionic strength I (mol/L).
  const double kap2 = 2.0 * ionic_strength * 1000.0 * NA * qe * qe /
                      (eps0 * EPS_OUT * kB * Tk) * 1e-20;
  const double kap = std::sqrt(kap2);
  const double c0 = -(n - 1) * h / 2.0;
  for (int i = 0; i < n; i++) xf[i] = yf[i] = zf[i] = c0 + i * h;

  const size_t nn = (size_t)n * n, N = nn * n;
  std::vector<unsigned char> inside(N, 0), excl(N, 0);
  // mark nodes inside atom spheres / inside the ion-exclusion shell
  parallel_for(n, [&](int k) {
    double z = c0 + k * h;
    for (int a = 0; a < natoms; a++) {
      double re = ar[a] + 2.0, dz = z - az[a];
      if (std::fabs(dz) > re) continue;
      int j0 = (int)std::ceil((ay[a] - re - c0) / h),
          j1 = (int)std::floor((ay[a] + re - c0) / h);
      int i0 = (int)std::ceil((ax[a] - re - c0) / h),
          i1 = (int)std::floor((ax[a] + re - c0) / h);
      if (j0 < 0) j0 = 0;
      if (i0 < 0) i0 = 0;
      if (j1 > n - 1) j1 = n - 1;
      if (i1 > n - 1) i1 = n - 1;
      for (int j = j0; j <= j1; j++) {
        double dy = c0 + j * h - ay[a];
        for (int i = i0; i <= i1; i++) {
          double dx = c0 + i * h - ax[a];
          double d2 = dx * dx + dy * dy + dz * dz;
          size_t p = (size_t)k * nn + (size_t)j * n + i;
          if (d2 < re * re) excl[p] = 1;
          if (d2 < ar[a] * ar[a]) inside[p] = 1;
        }
      }
    }
  });
  auto eps = [&](size_t p) { return inside[p] ? EPS_IN : EPS_OUT; };
  auto harm = [](double a, double b) { return 2.0 * a * b / (a + b); };
  parallel_for(n, [&](int k) {
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        size_t p = (size_t)k * nn + (size_t)j * n + i;
        double e0 = eps(p);
        a1cf[p] = i + 1 < n ? harm(e0, eps(p + 1)) : e0;
        a2cf[p] = j + 1 < n ? harm(e0, eps(p + n)) : e0;
        a3cf[p] = k + 1 < n ? harm(e0, eps(p + nn)) : e0;
        ccf[p] = excl[p] ? 0.0 : EPS_OUT * kap2;
        fcf[p] = 0.0;
        tcf[p] = 0.0;
      }
  });
  // trilinear charge spreading
  const double scale = 4.0 * M_PI * lB / (h * h * h);
  for (int a = 0; a < natoms; a++) {
    double gx = (ax[a] - c0) / h, gy = (ay[a] - c0) / h, gz = (az[a] - c0) / h;
    int i = (int)std::floor(gx), j = (int)std::floor(gy),
        k = (int)std::floor(gz);
    if (i < 0 || j < 0 || k < 0 || i + 1 > n - 1 || j + 1 > n - 1 ||
        k + 1 > n - 1)
      continue;  // outside the box: dropped
    double fx = gx - i, fy = gy - j, fz = gz - k;
    for (int c = 0; c < 8; c++) {
      int di = c & 1, dj = (c >> 1) & 1, dk = (c >> 2) & 1;
      double w = (di ? fx : 1.0 - fx) * (dj ? fy : 1.0 - fy) *
                 (dk ? fz : 1.0 - fz);
      fcf[(size_t)(k + dk) * nn + (size_t)(j + dj) * n + (i + di)] +=
          scale * aq[a] * w;
    }
  }
  // Debye-Hueckel boundary values on the six faces
  auto dh = [&](double x, double y, double z) {
    double s = 0.0;
    for (int a = 0; a < natoms; a++) {
      if (aq[a] == 0.0) continue;
      double dx = x - ax[a], dy = y - ay[a], dz = z - az[a];
      double d = std::sqrt(dx * dx + dy * dy + dz * dz);
      s += lB * aq[a] / EPS_OUT * std::exp(-kap * (d - ar[a])) /
           ((1.0 + kap * ar[a]) * d);
    }
    return s;
  };
  std::memset(gxcf, 0, sizeof(double) * 4 * nn);
  std::memset(gycf, 0, sizeof(double) * 4 * nn);
  std::memset(gzcf, 0, sizeof(double) * 4 * nn);
  const double lo = c0, hi = c0 + (n - 1) * h;
  parallel_for(n, [&](int b) {
    double vb = c0 + b * h;
    for (int a = 0; a < n; a++) {
      double va = c0 + a * h;
      // gxcf(j,k,1|2): j = a, k = b
      gxcf[(size_t)b * n + a] = dh(lo, va, vb);
      gxcf[nn + (size_t)b * n + a] = dh(hi, va, vb);
      // gycf(i,k,1|2): i = a, k = b
      gycf[(size_t)b * n + a] = dh(va, lo, vb);
      gycf[nn + (size_t)b * n + a] = dh(va, hi, vb);
      // gzcf(i,j,1|2): i = a, j = b
      gzcf[(size_t)b * n + a] = dh(va, vb, lo);
      gzcf[nn + (size_t)b * n + a] = dh(va, vb, hi);
    }
  });
  return PMG_OK;
}
