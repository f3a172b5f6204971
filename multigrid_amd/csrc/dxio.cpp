// dxio.cpp -- OpenDX scalar fields on regular grids (the format the
// reference writes with Domain.h:136-185 and APBS writes for potentials).
//
// Files hold "object 1 class gridpositions counts nx ny nz", an origin, three
// delta vectors, "object 3 class array type double rank 0 items N data
// follows" and N values with the last axis fastest (x slowest, z fastest, the
// OpenDX regular-grid order APBS uses).  In memory the field is in the PMG
// layout, x fastest.  Host-only; needs no GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/pmg_mi355x.h"

extern "C" int pmg_dx_write(const char *path, int nx, int ny, int nz,
                            double x0, double y0, double z0, double h,
                            const double *u, const char *title) {
  if (!path || !u || nx < 1 || ny < 1 || nz < 1 || !(h > 0.0))
    return PMG_EBADPARM;
  FILE *f = std::fopen(path, "w");
  if (!f) return PMG_EBADPARM;
  std::fprintf(f, "# Data from multigrid_amd (PMG on MI355X)\n#\n# %s\n#\n",
               title ? title : "POTENTIAL (kT/e)");
  std::fprintf(f, "object 1 class gridpositions counts %d %d %d\n", nx, ny,
               nz);
  std::fprintf(f, "origin %.6e %.6e %.6e\n", x0, y0, z0);
  std::fprintf(f, "delta %.6e %.6e %.6e\n", h, 0.0, 0.0);
  std::fprintf(f, "delta %.6e %.6e %.6e\n", 0.0, h, 0.0);
  std::fprintf(f, "delta %.6e %.6e %.6e\n", 0.0, 0.0, h);
  std::fprintf(f, "object 2 class gridconnections counts %d %d %d\n", nx, ny,
               nz);
  const long long N = (long long)nx * ny * nz;
  std::fprintf(f,
               "object 3 class array type double rank 0 items %lld data "
               "follows\n",
               N);
  long long c = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const double v = u[i + (size_t)nx * (j + (size_t)ny * k)];
        std::fprintf(f, "%.6e%c", v, (++c % 3 == 0 || c == N) ? '\n' : ' ');
      }
  std::fprintf(f,
               "attribute \"dep\" string \"positions\"\n"
               "object \"regular positions regular connections\" class field\n"
               "component \"positions\" value 1\n"
               "component \"connections\" value 2\n"
               "component \"data\" value 3\n");
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? PMG_OK : PMG_EBADPARM;
}

extern "C" int pmg_dx_read(const char *path, int *counts, double *origin,
                           double *delta, double *u, size_t cap) {
  if (!path || !counts) return PMG_EBADPARM;
  FILE *f = std::fopen(path, "r");
  if (!f) return PMG_EBADPARM;
  int n[3] = {0, 0, 0}, nd = 0;
  double org[3] = {0, 0, 0}, dl[3] = {0, 0, 0};
  long long items = -1;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    if (std::strstr(line, "gridpositions counts")) {
      const char *q = std::strstr(line, "counts") + 6;
      if (std::sscanf(q, "%d %d %d", &n[0], &n[1], &n[2]) != 3) break;
    } else if (!std::strncmp(line, "origin", 6)) {
      std::sscanf(line + 6, "%lf %lf %lf", &org[0], &org[1], &org[2]);
    } else if (!std::strncmp(line, "delta", 5) && nd < 3) {
      double d[3];
      if (std::sscanf(line + 5, "%lf %lf %lf", &d[0], &d[1], &d[2]) == 3)
        dl[nd] = d[nd];
      nd++;
    } else if (std::strstr(line, "data follows")) {
      const char *q = std::strstr(line, "items");
      if (q) items = std::atoll(q + 5);
      break;
    }
  }
  const long long N = (long long)n[0] * n[1] * n[2];
  int rc = (N > 0 && items == N && nd == 3) ? PMG_OK : PMG_EBADPARM;
  if (!rc) {
    counts[0] = n[0], counts[1] = n[1], counts[2] = n[2];
    if (origin) origin[0] = org[0], origin[1] = org[1], origin[2] = org[2];
    if (delta) delta[0] = dl[0], delta[1] = dl[1], delta[2] = dl[2];
  }
  if (!rc && u) {
    if ((size_t)N > cap) {
      rc = PMG_EBADPARM;
    } else {
      long long c = 0;
      for (int i = 0; i < n[0] && !rc; i++)
        for (int j = 0; j < n[1] && !rc; j++)
          for (int k = 0; k < n[2]; k++, c++) {
            double v;
            if (std::fscanf(f, "%lf", &v) != 1) {
              rc = PMG_EBADPARM;
              break;
            }
            u[i + (size_t)n[0] * (j + (size_t)n[1] * k)] = v;
          }
    }
  }
  std::fclose(f);
  return rc;
}
