// pbinput.cpp -- seeded synthetic molecules for the Poisson-Boltzmann
// workloads.  The reference reads atoms from PQR files
// (import_molecular_structure.cc); here a molecule is drawn from a seeded RNG:
//   atoms   uniform in a ball of radius R (direction = normalised Gaussian
//           triplet, radius R*U^(1/3)), radii U[1.2, 2.0] A;
//   charges mode 0: floor(N/2) of +1 and of -1 (0 for an odd remainder),
//           shuffled; mode 1: uniform over {-1,-0.5,0,0.5,1}.
// RNG: SplitMix64 seeding xoshiro256**, doubles from the top 53 bits.
// The grid inputs are then assembled on the GPU (pmg_pbe.hip).
#include <cmath>
#include <cstdint>
#include <utility>
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

}  // namespace

extern "C" int pmg_pbe_molecule(int natoms, double sphere_r, int charge_mode,
                                uint64_t seed, double *atoms) {
  if (natoms < 0 || !(sphere_r >= 0.0) ||
      (charge_mode != 0 && charge_mode != 1) || (natoms > 0 && !atoms))
    return PMG_EBADPARM;
  Xoshiro rng(seed);
  for (int a = 0; a < natoms; a++) {
    double gx = rng.gauss(), gy = rng.gauss(), gz = rng.gauss();
    double nrm = std::sqrt(gx * gx + gy * gy + gz * gz);
    if (nrm == 0.0) nrm = 1.0, gx = 1.0;
    double rad = sphere_r * std::cbrt(rng.uniform());
    atoms[5 * a + 0] = gx / nrm * rad;
    atoms[5 * a + 1] = gy / nrm * rad;
    atoms[5 * a + 2] = gz / nrm * rad;
    atoms[5 * a + 3] = 1.2 + 0.8 * rng.uniform();
  }
  std::vector<double> aq(natoms);
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
  for (int a = 0; a < natoms; a++) atoms[5 * a + 4] = aq[a];
  return PMG_OK;
}

extern "C" int pmg_pbe_synth(int n, double h, int natoms, double sphere_r,
                             double ionic_strength, int charge_mode,
                             uint64_t seed, double *xf, double *yf,
                             double *zf, double *gxcf, double *gycf,
                             double *gzcf, double *a1cf, double *a2cf,
                             double *a3cf, double *ccf, double *fcf,
                             double *tcf, double *atoms) {
  if (natoms < 0) return PMG_EBADPARM;
  std::vector<double> at(5 * (size_t)natoms);
  int rc = pmg_pbe_molecule(natoms, sphere_r, charge_mode, seed, at.data());
  if (rc) return rc;
  if (atoms)
    for (size_t q = 0; q < at.size(); q++) atoms[q] = at[q];
  return pmg_pbe_assemble(n, h, ionic_strength, natoms, at.data(), xf, yf, zf,
                          gxcf, gycf, gzcf, a1cf, a2cf, a3cf, ccf, fcf, tcf);
}
