// pmg_solver.hip -- PMG hierarchy, V-cycle driver and the C ABI.
//
// Host-side restatement of the reference's drivers with device-resident
// state: Vmgdriv/Vmgdriv2 (pmgc/mgdrvd.cc:60-503), Vbuildstr/Vbuildops
// (pmgc/mgsubd.cc:63-347), Vmvcs (pmgc/mgcsd.cc:60-530) and Vnewton
// (pmgc/newtond.cc:140-486).  Everything runs on one HIP stream per
// context; reductions finish on the device (xdamp never leaves HBM) and the
// host reads back one scalar per V-cycle for the stopping test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/pmg_mi355x.h"
#include "pmg_internal.h"
#include "pmg_ctx.h"

using namespace pmg;

namespace pmg {
namespace drv {



int maxlev(int n1, int n2, int n3) {  // Vmaxlev (mgsubd.cc:456-484)
  int lev = 0, done = 0;
  do {
    lev++;
    int iden = 1 << (lev - 1);
    int c1 = (n1 - 1) / iden + 1, c2 = (n2 - 1) / iden + 1,
        c3 = (n3 - 1) / iden + 1;
    if ((c1 - 1) * iden != n1 - 1 || c1 <= 2) done = 1;
    if ((c2 - 1) * iden != n2 - 1 || c2 <= 2) done = 1;
    if ((c3 - 1) * iden != n3 - 1 || c3 <= 2) done = 1;
  } while (!done && lev < 30);
  return lev - 1;
}

int mgsz(int mgcoar, int mgdisc, int mgsolv, int nx, int ny, int nz, int nlev,
         int *out) {  // Vmgsz (mgdrvd.cc:505-601)
  if (nlev < 1 || nlev > MAXLEV) return PMG_EBADPARM;
  long long nf = (long long)nx * ny * nz, narr = nf;
  int nxc = nx, nyc = ny, nzc = nz;
  for (int l = 2; l <= nlev; l++) {
    nxc = corsr(nxc), nyc = corsr(nyc), nzc = corsr(nzc);
    narr += (long long)nxc * nyc * nzc;
  }
  long long nc = (long long)nxc * nyc * nzc, narrc = narr - nf;
  int nfop, ncop;
  long long nband = 0;
  if (mgdisc == 0) nfop = 4;
  else if (mgdisc == 1) nfop = 14;
  else return PMG_EBADPARM;
  if ((mgcoar == 0 || mgcoar == 1) && mgdisc == 0) ncop = 4;
  else if (mgcoar == 2) ncop = 14;
  else return PMG_EBADPARM;
  if (mgsolv == 1) {
    long long nb = ((mgcoar == 0 || mgcoar == 1) && mgdisc == 0)
                       ? 1 + (long long)(nxc - 2) * (nyc - 2)
                       : 1 + (long long)(nxc - 2) * (nyc - 2) + (nxc - 2) + 1;
    nband = (long long)(nxc - 2) * (nyc - 2) * (nzc - 2) * nb;
  } else if (mgsolv != 0) {
    return PMG_EBADPARM;
  }
  long long nrpc = 100LL * (nlev + 1);
  long long iretot = 2 * narr + nfop * nf + (27 + ncop) * narrc + nband + nrpc;
  if (iretot > 0x7fffffffLL) return PMG_EBADPARM;  // int arrays of the ABI
  out[0] = nxc, out[1] = nyc, out[2] = nzc;
  out[3] = (int)nf, out[4] = (int)nc, out[5] = (int)narr, out[6] = (int)narrc;
  out[7] = (int)nrpc, out[8] = 50 * (nlev + 1), out[9] = 100 * (nlev + 1);
  out[10] = (int)iretot, out[11] = out[8] + out[9];
  return PMG_OK;
}

// Decode iparm/rparm (mgdrvd.cc:111-132, 247-265) and reject what this
// build does not provide.
int decode(const int *iparm, const double *rparm, Params &p) {
  p.nrwk = iparm[0], p.niwk = iparm[1], p.nx = iparm[2], p.ny = iparm[3],
  p.nz = iparm[4], p.nlev = iparm[5], p.nu1 = iparm[6], p.nu2 = iparm[7],
  p.mgkey = iparm[8], p.itmax = iparm[9], p.istop = iparm[10],
  p.iinfo = iparm[11], p.ipkey = iparm[13], p.mode = iparm[15],
  p.mgprol = iparm[16], p.mgcoar = iparm[17], p.mgdisc = iparm[18],
  p.mgsmoo = iparm[19], p.mgsolv = iparm[20], p.iperf = iparm[21];
  p.errtol = rparm[0];
  if (p.nlev < 1 || p.nlev > MAXLEV || p.nx < 3 || p.ny < 3 || p.nz < 3)
    return PMG_EBADPARM;
  if (p.nlev > 1 && p.nlev > maxlev(p.nx, p.ny, p.nz)) return PMG_EBADPARM;
  if (p.mgcoar != 2 || p.mgdisc != 0 || p.mgsmoo != 1 ||
      (p.mgsolv != 0 && p.mgsolv != 1) ||
      (p.mgprol != 0 && p.mgprol != 1) || p.istop < 0 || p.istop > 5 ||
      p.iperf != 0 || p.ipkey != 0 || p.nu1 < 0 || p.nu2 < 0 || p.itmax < 0 ||
      (p.mgkey != 0 && p.mgkey != 1))
    return PMG_EBADPARM;
  return PMG_OK;
}



thread_local int t_iters = 0, t_ierror = 0;
thread_local double t_relres = 0.0;
thread_local std::vector<double> t_hist;

int device_ok() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
  return 1;
}

Red red_to(pmg_ctx *c, int slot) {
  Red r;
  r.part = c->part;
  r.out = c->scal + slot;
  return r;
}

// read a device scalar back (waits for the stream)
double host_scalar(pmg_ctx *c, int slot) {
  hipMemcpyAsync(c->hscal + slot, c->scal + slot, sizeof(double),
                 hipMemcpyDeviceToHost, c->s);
  hipStreamSynchronize(c->s);
  return c->hscal[slot];
}

template <class T>
LV<T> view(Level &l);
template <>
LV<double> view<double>(Level &l) {
  return {l.d, l.numdia, l.ac, l.cc, l.pc, l.x, l.w0, l.r, l.t, l.dg};
}
template <>
LV<float> view<float>(Level &l) {
  return {l.d,    l.numdia, l.ac32, l.cc32, l.pc32,
          l.x32,  l.w032,   l.r32,  l.t32,  l.dg32};
}

// zero_init: x starts at zero (Vazeros).  bnd_zero: x's boundary entries
// are zero already, so a 7-point level skips clearing x and the first sweep
// does not read it.  uadd: where the level's fused FP32 sweep can, its last
// sweep adds x to this FP64 iterate instead of storing x (returns true).
template <class T>
bool smooth(pmg_ctx *c, const LV<T> &l, const T *f, T *x, int nu, int iadj,
            bool fine, bool zero_init = false, bool bnd_zero = false,
            double *uadd = nullptr) {
  if (nu <= 0) {
    if (zero_init) fill<T>(x, l.d.n(), (T)0, c->s);
    return false;
  }
  bool timed = c->ktime && !c->capturing && fine &&
               c->nev + 2 <= (int)c->ev.size();
  if (timed) hipEventRecord(c->ev[c->nev], c->s);
  bool added = false;
  if (l.numdia == 7) {
    if (l.t && l.t != x) {
      if (zero_init && !bnd_zero) fill<T>(x, l.d.n(), (T)0, c->s);
      added = gs7_fused<T>(l.d, l.ac, l.cc, f, x, l.t, nu, iadj, c->s,
                           zero_init && bnd_zero, l.dg, uadd);
    } else {
      if (zero_init) fill<T>(x, l.d.n(), (T)0, c->s);
      gs7<T>(l.d, l.ac, l.cc, f, x, nu, iadj, c->s);
    }
  } else {
    // coarse x and t both have zero boundary entries (nothing writes them)
    gs27<T>(l.d, l.ac, l.cc, f, x, nu, iadj, c->s, zero_init,
            l.t != x ? l.t : nullptr);
  }
  if (timed) {
    hipEventRecord(c->ev[c->nev + 1], c->s);
    c->nev += 2;
    c->ktime_count += nu;
  }
  return added;
}

// Coarsest-level solve: Vcghs with 100 iterations and errtol = epsiln = 0
// (mgsolv 0; epsiln 0 gives exactly 100 iterations), or the factored band (mgsolv 1, mgcsd.cc:391-407, which
// ignores cc as the reference's band does).  FP64 inside either way.
// errtol > 0: Vcghs's relative stopping test sqrt(r.r)/sqrt(r0.r0) <= errtol
// with that tolerance instead of epsiln (at most 100 iterations).
template <class T>
void coarse_solve(pmg_ctx *c, int l, const double *ac64, const double *cc64,
                  const T *f, T *x, double errtol = 0.0) {
  const Level &L = c->L[l];
  if (c->p.mgsolv == 1)
    band_solve<T>(L.d, L.numdia, c->band_rt, c->band_ct, f, x, c->s);
  else
    cg_single<T>(L.d, L.numdia, ac64, cc64, f, x, 100, errtol, c->s);
}

// Relative tolerance of the FP64 coarsest CG inside the FP32 V-cycle of the
// mixed-precision solves (the linear solve and Newton's inner linear
// solves).  The cycle is a correction step of the FP64 defect correction,
// whose stopping test is the FP64 residual; its coarse
// correction is scaled by the Hackbusch-Reusken factor on the way up.  On
// the benchmark hierarchies (coarsest 7^3 to 11^3) the residual history of
// the cycle is the same to three digits whether the coarsest level is
// solved to 1e-1, 1e-2, 1e-3 or by 100 iterations; 1e-2 takes 7 to 11
// iterations.
constexpr double kMixedCoarseTol = 1e-2;

// The coarsest level fits the coarse solver of p.
bool coarse_fits(pmg_ctx *c) {
  const Level &L = c->L[c->nlev - 1];
  return c->p.mgsolv == 1 ? band_dims(L.d, L.numdia).ok : cg_fits(L.d);
}

// One V-cycle (Vmvcs body, mgcsd.cc:299-480).  L[0].x is the iterate on
// the fine level; frhs its right-hand side; cc64c the FP64 Helmholtz term of
// the coarsest level used by the FP64 coarse solve.
// fine_zero: the fine iterate starts at zero (the correction cycle of the
// mixed solve); its boundary entries are zero already.
// nlevd: the cycle runs on the last nlevd levels of the hierarchy (L[0] is
// the first of them; 0: all levels).
// uadd: the FP64 iterate the cycle's fine result corrects.  Where the last
// fine post-smoothing sweep can, it makes the update uadd += (double)x
// itself and the cycle returns true (L[0].x then does not hold the
// result); otherwise the caller adds L[0].x.
template <class T>
bool vcycle(pmg_ctx *c, LV<T> *L, const T *frhs, const double *ac64c,
            const double *cc64c, bool fine_zero = false,
            double coarse_tol = 0.0, int nlevd = 0, double *uadd = nullptr) {
  const int nlev = nlevd > 0 ? nlevd : c->nlev;
  const Params &p = c->p;
  hipStream_t s = c->s;
  // nu1 pre-smoothings on the fine grid, then the residual f - A x that
  // Vmvcs restricts (the reference's Vgsrb drops its iresid output)
  smooth<T>(c, L[0], frhs, L[0].x, p.nu1, 0, true, fine_zero, fine_zero);
  resid<T, T>(L[0].d, L[0].numdia, L[0].ac, L[0].cc, frhs, L[0].x, L[0].r,
              nullptr, s, L[0].dg);
  for (int l = 1; l < nlev; l++) {  // down (mgcsd.cc:325-360)
    restrict_<T>(L[l - 1].d, L[l].d, L[l - 1].r, L[l].w0, L[l].pc, s);
    if (l != nlev - 1) {
      // Vazeros + nu1 smoothings (zero start folded into the sweep)
      smooth<T>(c, L[l], L[l].w0, L[l].x, p.nu1, 0, false, true);
      resid<T, T>(L[l].d, L[l].numdia, L[l].ac, L[l].cc, L[l].w0, L[l].x,
                  L[l].r, nullptr, s);
    }
  }
  coarse_solve<T>(c, c->nlev - 1, ac64c, cc64c, L[nlev - 1].w0,
                  L[nlev - 1].x, coarse_tol);
  for (int l = nlev - 2; l >= 0; l--) {  // up (mgcsd.cc:418-473)
    const int cl = l + 1;
    const unsigned nb =
        matvec_dots_part<T>(L[cl].d, L[cl].numdia, L[cl].ac, L[cl].cc,
                            L[cl].x, L[cl].w0, L[cl].t, c->part, s);
    interp_damped<T>(L[cl].d, L[l].d, L[cl].x, L[cl].pc, L[l].x, c->part, nb,
                     s);
    if (l == 0)
      return smooth<T>(c, L[0], frhs, L[0].x, p.nu2, 1, true, false, false,
                       uadd);
    smooth<T>(c, L[l], L[l].w0, L[l].x, p.nu2, 1, false);
  }
  return false;
}

// Linear solve with FP64 V-cycles.  L[0].x holds the iterate (zeroed by the
// caller), frhs the fine right-hand side.  istop 0: absolute L1 residual,
// 1: relative to ||frhs||_1 (mgcsd.cc:133-163, 483-527).
// l0: the solve runs on levels l0.. of the hierarchy (L[0] is level l0).
int linear_fp64(pmg_ctx *c, LV<double> *L, const double *frhs,
                const double *cc64c, int istop, double errtol, int itmax,
                int *iters_out, double *rel_out, int l0 = 0) {
  Level &F = c->L[l0];
  double rsden = 1.0;
  const Red rden = red_to(c, S_RSDEN), rnrm = red_to(c, S_RSNRM);
  if (istop == 1) {
    nrm1<double>(F.d, frhs, rden, c->s);
    rsden = host_scalar(c, S_RSDEN);
  }
  if (rsden == 0.0) rsden = 1.0;
  double rsnrm = rsden;
  const int nlev = c->nlev - l0;
  if (nlev == 1) {  // direct coarse solve on the only level (mgcsd.cc:170-291)
    coarse_solve<double>(c, l0, F.ac, L[0].cc, frhs, L[0].x);
    resid<double, double>(F.d, F.numdia, F.ac, L[0].cc, frhs, L[0].x,
                          (double *)nullptr, &rnrm, c->s, L[0].dg);
    rsnrm = host_scalar(c, S_RSNRM);
    *iters_out = 1;
    *rel_out = rsnrm / rsden;
    return PMG_OK;
  }
  int iters = 0;
  do {
    vcycle<double>(c, L, frhs, c->L[c->nlev - 1].ac, cc64c, false, 0.0, nlev);
    iters++;
    resid<double, double>(F.d, F.numdia, F.ac, L[0].cc, frhs, L[0].x,
                          (double *)nullptr, &rnrm, c->s, L[0].dg);
    rsnrm = host_scalar(c, S_RSNRM);
  } while (iters < itmax && rsnrm / rsden > errtol);
  *iters_out = iters;
  *rel_out = rsnrm / rsden;
  return iters < itmax ? PMG_OK : PMG_ITMAX;
}

// Mixed precision: FP64 defect correction around FP32 V-cycles.
//   r = f - A u (FP64, stored as the FP32 right-hand side of the cycle)
//   e = V-cycle_FP32(0, r);  u += e (FP64)
// The stopping test is the FP64 L1 residual, made after each cycle as in
// linear_fp64 (at least one cycle, so the starting residual's norm is not
// read back).
int linear_mixed(pmg_ctx *c, LV<double> *L64, LV<float> *L32,
                 const double *frhs, const double *cc64c, double *u,
                 int istop, double errtol, int itmax, int *iters_out,
                 double *rel_out, int l0 = 0) {
  Level &F = c->L[l0];
  double rsden = 1.0;
  const Red rden = red_to(c, S_RSDEN), rnrm = red_to(c, S_RSNRM);
  if (istop == 1) {
    nrm1<double>(F.d, frhs, rden, c->s);
    rsden = host_scalar(c, S_RSDEN);
  }
  if (rsden == 0.0) rsden = 1.0;
  float *r32 = F.w032;  // fine right-hand side of the FP32 cycle
  resid<double, float>(F.d, F.numdia, F.ac, L64[0].cc, frhs, u, r32,
                       (const Red *)nullptr, c->s, L64[0].dg);
  double rsnrm = 0.0;
  int iters = 0;
  const int nlev = c->nlev - l0;
  do {
    bool added = false;
    if (nlev == 1) {
      fill<float>(L32[0].x, F.d.n(), 0.0f, c->s);
      coarse_solve<float>(c, l0, F.ac, L64[0].cc, r32, L32[0].x);
    } else {
      added = vcycle<float>(c, L32, r32, c->L[c->nlev - 1].ac, cc64c, true,
                            kMixedCoarseTol, nlev, u);
    }
    if (!added) add_correction<float>(F.d, L32[0].x, u, c->s);
    iters++;
    resid<double, float>(F.d, F.numdia, F.ac, L64[0].cc, frhs, u, r32, &rnrm,
                         c->s, L64[0].dg);
    rsnrm = host_scalar(c, S_RSNRM);
  } while (iters < itmax && rsnrm / rsden > errtol);
  *iters_out = iters;
  *rel_out = rsnrm / rsden;
  return (iters < itmax || rsnrm / rsden <= errtol) ? PMG_OK : PMG_ITMAX;
}

int alloc_hierarchy(pmg_ctx *c) {
  const Params &p = c->p;
  c->nlev = p.nlev;
  Dim d{p.nx, p.ny, p.nz};
  for (int l = 0; l < c->nlev; l++) {
    Level &L = c->L[l];
    L.d = d;
    L.numdia = l == 0 ? 7 : 27;
    size_t n = d.n();
    L.ac = c->alloc<double>((l == 0 ? 4 : 14) * n);
    L.cc = c->alloc<double>(n);
    L.fc = c->alloc<double>(n);
    if (l == 0) L.dg = c->alloc<double>(n);
    if (l > 0) L.pc = c->alloc<double>(27 * n);
    L.x = c->alloc<double>(n);
    L.w0 = c->alloc<double>(n);
    L.r = c->alloc<double>(n);
    L.t = c->alloc<double>(n);
    d = Dim{corsr(d.nx), corsr(d.ny), corsr(d.nz)};
  }
  if (p.mgsolv == 1) {  // Vbuildband storage (mgsubd.cc:249-262)
    const Level &C = c->L[c->nlev - 1];
    const BandDims b = band_dims(C.d, C.numdia);
    if (!b.ok) return PMG_EBADPARM;
    c->band = c->alloc<double>((size_t)b.n * b.lda);
    c->band_rt = c->alloc<double>((size_t)b.nrow * b.W);
    c->band_ct = c->alloc<double>((size_t)b.nrow * b.W);
  }
  size_t nf = c->L[0].d.n();
  c->u = c->alloc<double>(nf);
  c->part = c->alloc<double>(max_partials(c->L[0].d) * 2);
  c->scal = c->alloc<double>(16);
  if (hipHostMalloc((void **)&c->hscal, 16 * sizeof(double)) != hipSuccess)
    c->ok = false;
  return c->ok ? PMG_OK : PMG_ENOMEM;
}

int alloc_fp32(pmg_ctx *c) {
  if (c->have32) return PMG_OK;
  for (int l = 0; l < c->nlev; l++) {
    Level &L = c->L[l];
    size_t n = L.d.n();
    L.ac32 = c->alloc<float>((l == 0 ? 4 : 14) * n);
    L.cc32 = c->alloc<float>(n);
    if (l == 0) L.dg32 = c->alloc<float>(n);
    if (l > 0) L.pc32 = c->alloc<float>(27 * n);
    L.x32 = c->alloc<float>(n);
    L.w032 = c->alloc<float>(n);
    L.r32 = c->alloc<float>(n);
    L.t32 = c->alloc<float>(n);
  }
  if (!c->ok) return PMG_ENOMEM;
  c->have32 = true;
  return PMG_OK;
}

void convert_ops32(pmg_ctx *c) {
  for (int l = 0; l < c->nlev; l++) {
    Level &L = c->L[l];
    size_t n = L.d.n();
    convert<double, float>(L.ac, L.ac32, (l == 0 ? 4 : 14) * n, c->s);
    convert<double, float>(L.cc, L.cc32, n, c->s);
    if (l == 0) diag_sum<float>(L.ac32, L.cc32, L.dg32, n, c->s);
    if (l > 0) convert<double, float>(L.pc, L.pc32, 27 * n, c->s);
  }
}

// Vbuildops ido = 0 and 1 (mgsubd.cc:63-264), Galerkin coarsening.
void build_ops(pmg_ctx *c) {
  hipStream_t s = c->s;
  Level &F = c->L[0];
  // With the FP32 hierarchy allocated, the build kernels write its copies
  // as they go (the same values convert_ops32 would give: every entry the
  // kernels do not write is zero in both precisions).
  const bool w32 = c->have32;
  buildA(F.d, F.ac, F.cc, F.fc, c->xf, c->yf, c->zf, c->gx, c->gy, c->gz,
         c->a1, c->a2, c->a3, c->ccf, c->fcf, s, F.dg,
         w32 ? F.ac32 : nullptr, w32 ? F.cc32 : nullptr,
         w32 ? F.dg32 : nullptr);
  for (int l = 1; l < c->nlev; l++) {
    Level &Lf = c->L[l - 1], &Lc = c->L[l];
    buildP(Lf.d, Lc.d, c->p.mgprol, Lf.numdia, Lf.ac, Lc.pc, s,
           w32 ? Lc.pc32 : nullptr);
    buildG(Lf.d, Lc.d, Lf.numdia, Lc.pc, Lf.ac, Lc.ac, s,
           w32 ? Lc.ac32 : nullptr);
    restrict2(Lf.d, Lc.d, Lf.cc, Lf.fc, Lc.cc, Lc.fc, Lc.pc, s,
              w32 ? Lc.cc32 : nullptr);
  }
  if (c->p.mgsolv == 1) {  // Vbuildband + Vdpbfa (mgsubd.cc:249-262)
    const Level &C = c->L[c->nlev - 1];
    band_build_factor(C.d, C.numdia, C.ac, c->band, c->band_rt, c->band_ct,
                      c->scal + S_BINFO, s);
  }
  c->ops32_ok = w32;
}

template <class T>
void upload(pmg_ctx *c, T *dst, const T *src, size_t n) {
  if (dst && src) hipMemcpyAsync(dst, src, n * sizeof(T),
                                 hipMemcpyHostToDevice, c->s);
}

// Allocate a context for p: stream, hierarchy and the resident fine-level
// inputs (zeroed, no data yet).
int ctx_new(pmg_ctx **out, const Params &p) {
  *out = nullptr;
  if (!device_ok()) return PMG_EDEVICE;
  pmg_ctx *c = new (std::nothrow) pmg_ctx;
  if (!c) return PMG_ENOMEM;
  c->p = p;
  if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PMG_EDEVICE;
  }
  int rc = alloc_hierarchy(c);
  if (rc) {
    delete c;
    return rc;
  }
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const size_t nf = (size_t)nx * ny * nz;
  c->xf = c->alloc<double>(nx);
  c->yf = c->alloc<double>(ny);
  c->zf = c->alloc<double>(nz);
  c->gx = c->alloc<double>(2 * (size_t)ny * nz);
  c->gy = c->alloc<double>(2 * (size_t)nx * nz);
  c->gz = c->alloc<double>(2 * (size_t)nx * ny);
  c->a1 = c->alloc<double>(nf);
  c->a2 = c->alloc<double>(nf);
  c->a3 = c->alloc<double>(nf);
  c->ccf = c->alloc<double>(nf);
  c->fcf = c->alloc<double>(nf);
  if (!c->ok) {
    delete c;
    return PMG_ENOMEM;
  }
  *out = c;
  return PMG_OK;
}

// Enqueue the hierarchy build (and, for a mixed-precision context, the FP32
// operator copies) bracketed by the setup events.
void setup_start(pmg_ctx *c) {
  if (!c->su0_ev) {
    hipEventCreate(&c->su0_ev);
    hipEventCreate(&c->su1_ev);
  }
  hipEventRecord(c->su0_ev, c->s);
}

void setup_enqueue(pmg_ctx *c, bool started) {
  if (!started) setup_start(c);
  build_ops(c);
  hipEventRecord(c->su1_ev, c->s);
}

// Wait for the setup and record its device time.
int setup_wait(pmg_ctx *c) {
  if (hipStreamSynchronize(c->s) != hipSuccess) return PMG_EDEVICE;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->su0_ev, c->su1_ev);
  c->setup_ms = ms;
  // the reference stops when Vdpbfa finds a non-positive pivot
  if (c->p.mgsolv == 1 && host_scalar(c, S_BINFO) != 0.0) return PMG_ENOTSPD;
  return PMG_OK;
}

int ctx_create(pmg_ctx **out, const Params &p, const double *xf,
               const double *yf, const double *zf, const double *gxcf,
               const double *gycf, const double *gzcf, const double *a1cf,
               const double *a2cf, const double *a3cf, const double *ccf,
               const double *fcf) {
  pmg_ctx *c = nullptr;
  int rc = ctx_new(&c, p);
  if (rc) return rc;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const size_t nf = (size_t)nx * ny * nz;
  upload(c, c->xf, xf, nx);
  upload(c, c->yf, yf, ny);
  upload(c, c->zf, zf, nz);
  upload(c, c->gx, gxcf, 2 * (size_t)ny * nz);
  upload(c, c->gy, gycf, 2 * (size_t)nx * nz);
  upload(c, c->gz, gzcf, 2 * (size_t)nx * ny);
  upload(c, c->a1, a1cf, nf);
  upload(c, c->a2, a2cf, nf);
  upload(c, c->a3, a3cf, nf);
  upload(c, c->ccf, ccf, nf);
  upload(c, c->fcf, fcf, nf);
  setup_enqueue(c);
  rc = setup_wait(c);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return PMG_OK;
}

// Stage an atom list and enqueue the GPU assembly of the fine-level inputs
// (pmg_pbe.hip) into the context's resident input arrays.  The assembly
// scratch is the fine residual buffer, which no solve is using yet.
int pbe_enqueue(pmg_ctx *c, double h, double ionic, int natoms,
                const double *atoms) {
  const Params &p = c->p;
  if (p.nx != p.ny || p.nx != p.nz || !(h > 0.0) || !(ionic >= 0.0) ||
      natoms < 0 || (natoms > 0 && !atoms))
    return PMG_EBADPARM;
  if (c->stage_ev) hipEventSynchronize(c->stage_ev);  // staging buffer free
  if (natoms > c->atoms_cap) {
    hipStreamSynchronize(c->s);
    if (c->atoms_d) hipFree(c->atoms_d);
    if (c->atoms_h) hipHostFree(c->atoms_h);
    c->atoms_d = nullptr, c->atoms_h = nullptr, c->atoms_cap = 0;
    int cap = std::max(natoms, 1024);
    // atoms (5 doubles) and the cell-list links (1 int) per atom
    if (hipMalloc((void **)&c->atoms_d, (size_t)cap * 6 * sizeof(double)) !=
            hipSuccess ||
        hipHostMalloc((void **)&c->atoms_h,
                      (size_t)cap * 5 * sizeof(double)) != hipSuccess)
      return PMG_ENOMEM;
    c->atoms_cap = cap;
  }
  if (!c->stage_ev) hipEventCreateWithFlags(&c->stage_ev, hipEventDisableTiming);
  if (natoms > 0) {
    memcpy(c->atoms_h, atoms, (size_t)natoms * 5 * sizeof(double));
    hipMemcpyAsync(c->atoms_d, c->atoms_h, (size_t)natoms * 5 * sizeof(double),
                   hipMemcpyHostToDevice, c->s);
  }
  hipEventRecord(c->stage_ev, c->s);
  int *next = (int *)(c->atoms_d + 5 * (size_t)c->atoms_cap);
  int rc = pbe_assemble(p.nx, h, ionic, natoms, c->atoms_d, next, c->L[0].r,
                        c->xf, c->yf, c->zf, c->gx, c->gy, c->gz, c->a1,
                        c->a2, c->a3, c->ccf, c->fcf, c->s);
  return rc ? PMG_EBADPARM : PMG_OK;
}

int ctx_create_pbe(pmg_ctx **out, const Params &p, double h, double ionic,
                   int natoms, const double *atoms) {
  pmg_ctx *c = nullptr;
  int rc = ctx_new(&c, p);
  if (rc) return rc;
  setup_start(c);  // setup time includes the input assembly
  rc = pbe_enqueue(c, h, ionic, natoms, atoms);
  if (!rc) {
    setup_enqueue(c, true);
    rc = setup_wait(c);
  }
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return PMG_OK;
}

// ---- linear solve as a session of enqueue / poll steps ----
// One step enqueues one iteration of Vmvcs's loop (mgcsd.cc:299-527): the
// V-cycle and the FP64 L1 stopping norm (for PMG_MIXED: the FP32 V-cycle on
// the FP64 residual, the FP64 update, the new FP64 residual and its norm),
// then an 8-byte copy of the norm to pinned host memory and an event.  The
// host decides whether to continue when the event has completed.  The step
// body is captured once per mode as a HIP graph and replayed.

void views(pmg_ctx *c, LV<double> *L64, LV<float> *L32, int mode) {
  for (int l = 0; l < c->nlev; l++) {
    L64[l] = view<double>(c->L[l]);
    if (mode == PMG_MIXED) L32[l] = view<float>(c->L[l]);
  }
  L64[0].x = c->u;
}

// The stopping measure of istop 2-5 into the device scalar `slot`
// (mgcsd.cc:497-526): 2 the L1 difference to tru, which then takes u; 3, 4
// the squared 2-norm of tru - u; 5 its energy (tru-u) . A (tru-u) (the host
// takes the square roots).  The fine FP64 scratch r and t hold tru - u and
// its product with A.
void stop_measure(pmg_ctx *c, int slot) {
  Level &F = c->L[0];
  const Red rd = red_to(c, slot);
  const int istop = c->p.istop;
  if (istop == 2) {
    stop_diff(F.d, 2, c->tru, c->u, nullptr, rd, c->s);
  } else if (istop == 3 || istop == 4) {
    stop_diff(F.d, 3, c->tru, c->u, nullptr, rd, c->s);
  } else {
    stop_diff(F.d, 5, c->tru, c->u, F.r, rd, c->s);
    dot_energy<double>(F.d, F.numdia, F.ac, F.cc, F.r, F.t, rd, c->s);
  }
}

void step_body(pmg_ctx *c, int mode) {
  LV<double> L64[MAXLEV];
  LV<float> L32[MAXLEV];
  views(c, L64, L32, mode);
  Level &F = c->L[0];
  const int nlev = c->nlev;
  const double *cc64c = c->L[nlev - 1].cc;
  const Red rnrm = red_to(c, S_RSNRM);
  // istop 0, 1: the L1 norm of the FP64 residual; 2-5 their own measure
  const bool resid_norm = c->p.istop <= 1;
  if (mode == PMG_MIXED) {
    float *r32 = F.w032;
    bool added = false;
    if (nlev == 1) {
      fill<float>(L32[0].x, F.d.n(), 0.0f, c->s);
      coarse_solve<float>(c, 0, F.ac, F.cc, r32, L32[0].x);
    } else {
      added = vcycle<float>(c, L32, r32, c->L[nlev - 1].ac, cc64c, true,
                            kMixedCoarseTol, 0, c->u);
    }
    if (!added) add_correction<float>(F.d, L32[0].x, c->u, c->s);
    resid<double, float>(F.d, F.numdia, F.ac, F.cc, F.fc, c->u, r32,
                         resid_norm ? &rnrm : nullptr, c->s, F.dg);
  } else {
    if (nlev == 1)
      coarse_solve<double>(c, 0, F.ac, F.cc, F.fc, c->u);
    else
      vcycle<double>(c, L64, F.fc, c->L[nlev - 1].ac, cc64c);
    if (resid_norm)
      resid<double, double>(F.d, F.numdia, F.ac, F.cc, F.fc, c->u,
                            (double *)nullptr, &rnrm, c->s, F.dg);
  }
  if (!resid_norm) stop_measure(c, S_RSNRM);
}

// Vbuildalg (mgdrvd.cc:424-433, mgsubd.cc:891-950) for istop 4 and 5: the
// right-hand side of every level becomes A tcf, tcf carried to the coarse
// levels by injection (Vextrac in Vbuildops), so that tcf is the solution.
// The coarse levels' r arrays hold the injected tcf.
void build_alg(pmg_ctx *c) {
  const double *t = c->tru;
  for (int l = 0; l < c->nlev; l++) {
    Level &L = c->L[l];
    if (l > 0) {
      extrac(c->L[l - 1].d, L.d, t, L.r, c->s);
      t = L.r;
    }
    matvec_dots<double>(L.d, L.numdia, L.ac, L.cc, t, nullptr, L.fc, nullptr,
                        c->s);
  }
}

int capture(pmg_ctx *c, int mode) {
  if (c->gexec[mode]) return PMG_OK;
  if (hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal) !=
      hipSuccess)
    return PMG_EDEVICE;
  c->capturing = true;
  step_body(c, mode);
  c->capturing = false;
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture(c->s, &g) != hipSuccess || !g) return PMG_EDEVICE;
  if (hipGraphInstantiate(&c->gexec[mode], g, nullptr, nullptr, 0) !=
      hipSuccess) {
    hipGraphDestroy(g);
    c->gexec[mode] = nullptr;
    return PMG_EDEVICE;
  }
  c->graph[mode] = g;
  return PMG_OK;
}

// Start a solve: Vazeros + VfboundPMG00 on u (mgdrvd.cc:419, 439), the
// stopping-test denominator (mgcsd.cc:133-163) and, for PMG_MIXED, the
// first residual; the two scalars come back with the start event, which
// lin_poll reads.  Never blocks.
int lin_begin(pmg_ctx *c, int mode) {
  const Params &p = c->p;
  Level &F = c->L[0];
  if (mode == PMG_MIXED) {
    int rc = alloc_fp32(c);
    if (rc) return rc;
    if (!c->ops32_ok) convert_ops32(c);
    c->ops32_ok = true;
  }
  if (!coarse_fits(c)) return PMG_EBADPARM;
  if (p.istop >= 2 && !c->tru) return PMG_EBADPARM;  // no tcf given
  if (!c->cyc_ev) {
    hipEventCreateWithFlags(&c->cyc_ev, hipEventDisableTiming);
    hipEventCreateWithFlags(&c->fin_ev, hipEventDisableTiming);
    hipEventCreate(&c->t0_ev);
    hipEventCreate(&c->t1_ev);
  }
  // the iteration graph is captured before the start is enqueued (capture
  // records, it does not run, the work issued while it lasts)
  if (c->graphs && !c->ktime) {
    int rc = capture(c, mode);
    if (rc) return rc;
  }
  auto &S = c->ses;
  S = pmg_ctx::Session{};
  S.mode = mode;
  S.done = false;
  S.phase = 0;
  c->nev = 0;
  c->hist.clear();
  hipEventRecord(c->t0_ev, c->s);
  if (p.istop == 4 || p.istop == 5) build_alg(c);
  fill<double>(c->u, F.d.n(), 0.0, c->s);
  if (p.istop == 1) {
    const Red rden = red_to(c, S_RSDEN);
    nrm1<double>(F.d, F.fc, rden, c->s);
  } else if (p.istop >= 3) {  // ||tru||_2^2, or tru . A tru (u = 0 here)
    const Red rden = red_to(c, S_RSDEN);
    if (p.istop == 5) {
      stop_diff(F.d, 5, c->tru, c->u, F.r, rden, c->s);
      dot_energy<double>(F.d, F.numdia, F.ac, F.cc, F.r, F.t, rden, c->s);
    } else {
      stop_diff(F.d, 3, c->tru, c->u, nullptr, rden, c->s);
    }
  }
  if (mode == PMG_MIXED) {
    // residual of u = 0, the first right-hand side of the FP32 cycle
    const Red rnrm = red_to(c, S_RSNRM);
    resid<double, float>(F.d, F.numdia, F.ac, F.cc, F.fc, c->u, F.w032,
                         p.istop <= 1 ? &rnrm : nullptr, c->s, F.dg);
  }
  // the two scalars and the band factor's info (slot S_BINFO)
  hipMemcpyAsync(c->hscal, c->scal, 4 * sizeof(double), hipMemcpyDeviceToHost,
                 c->s);
  hipEventRecord(c->cyc_ev, c->s);
  S.pending = true;
  return PMG_OK;
}

void lin_enqueue(pmg_ctx *c) {
  auto &S = c->ses;
  if (S.done || S.pending || S.phase != 1) return;
  if (c->gexec[S.mode] && !c->ktime)
    hipGraphLaunch(c->gexec[S.mode], c->s);
  else
    step_body(c, S.mode);
  hipMemcpyAsync(c->hscal + S_RSNRM, c->scal + S_RSNRM, sizeof(double),
                 hipMemcpyDeviceToHost, c->s);
  hipEventRecord(c->cyc_ev, c->s);
  S.pending = true;
}

// Returns 1 when the session has finished (after this poll), 0 otherwise.
int lin_poll(pmg_ctx *c, bool wait) {
  auto &S = c->ses;
  const Params &p = c->p;
  if (S.done) return 1;
  if (!S.pending) return 0;
  if (wait) {
    hipEventSynchronize(c->cyc_ev);
  } else if (hipEventQuery(c->cyc_ev) != hipSuccess) {
    return 0;
  }
  S.pending = false;
  if (S.phase == 0) {  // the start has landed
    S.phase = 1;
    if (p.mgsolv == 1 && c->hscal[S_BINFO] != 0.0) {  // band not SPD
      S.done = true;
      S.rc = PMG_ENOTSPD;
      return 1;
    }
    // the stopping-test denominator (mgcsd.cc:133-157)
    if (p.istop == 1) S.rsden = c->hscal[S_RSDEN];
    else if (p.istop == 2) S.rsden = std::sqrt((double)c->L[0].d.n());
    else if (p.istop >= 3) S.rsden = std::sqrt(c->hscal[S_RSDEN]);
    if (S.rsden == 0.0) S.rsden = 1.0;  // mgcsd.cc:154-157
    const bool mixed_res = S.mode == PMG_MIXED && p.istop <= 1;
    S.rsnrm = mixed_res ? c->hscal[S_RSNRM] : S.rsden;
    // mixed: while (iters < itmax && relres > errtol) may do nothing
    if (mixed_res && !(S.iters < p.itmax && S.rsnrm / S.rsden > p.errtol)) {
      S.done = true;
      S.rc = PMG_OK;
      return 1;
    }
    return 0;
  }
  return ses_cycle_done(c);
}

// An iteration of the session has landed, its stopping scalar in
// c->hscal[S_RSNRM]: count it, record and report it, decide whether the
// session goes on (mgcsd.cc:483-527).  Returns 1 when it has finished.
int ses_cycle_done(pmg_ctx *c) {
  auto &S = c->ses;
  const Params &p = c->p;
  S.iters++;
  const double orsnrm = S.rsnrm;
  S.rsnrm = p.istop >= 3 ? std::sqrt(c->hscal[S_RSNRM]) : c->hscal[S_RSNRM];
  c->hist.push_back(S.rsnrm / S.rsden);
  if (p.iinfo > 0)  // Vprtstp (mgsubd.cc:486-536)
    printf("iteration = %d relative residual = %e contraction number = %e\n",
           S.iters, S.rsnrm / S.rsden,
           orsnrm != 0.0 ? S.rsnrm / orsnrm : 0.0);
  const bool more = (c->nlev > 1 || S.mode == PMG_MIXED) && S.iters < p.itmax &&
                    S.rsnrm / S.rsden > p.errtol;
  if (!more) {
    S.done = true;
    const bool conv = S.rsnrm / S.rsden <= p.errtol;
    S.rc = (S.iters < p.itmax || conv) ? PMG_OK : PMG_ITMAX;
    if (c->nlev == 1 && S.mode == PMG_FP64) S.rc = PMG_OK;
  }
  return S.done ? 1 : 0;
}

// VfboundPMG(ibound=1): Dirichlet values on u (mgdrvd.cc:500-502), then the
// optional copy of the solution to pinned host memory.
void lin_finish(pmg_ctx *c, double *host_u) {
  Level &F = c->L[0];
  fbound_dirichlet(F.d, c->u, c->gx, c->gy, c->gz, c->s);
  hipEventRecord(c->t1_ev, c->s);
  if (host_u)
    hipMemcpyAsync(host_u, c->u, F.d.n() * sizeof(double),
                   hipMemcpyDeviceToHost, c->s);
  hipEventRecord(c->fin_ev, c->s);
  c->ses.phase = 2;
}

int lin_finished(pmg_ctx *c, bool wait) {
  if (wait) return hipEventSynchronize(c->fin_ev) == hipSuccess ? 1 : -1;
  const hipError_t e = hipEventQuery(c->fin_ev);
  if (e == hipSuccess) return 1;
  return e == hipErrorNotReady ? 0 : -1;
}

int lin_end(pmg_ctx *c) {
  auto &S = c->ses;
  hipError_t se = hipEventSynchronize(c->fin_ev);
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->t0_ev, c->t1_ev);
  c->solve_ms = ms;
  if (c->ktime) {
    for (int q = 0; q + 1 < c->nev; q += 2) {
      float kms = 0.f;
      hipEventElapsedTime(&kms, c->ev[q], c->ev[q + 1]);
      c->ktime_ms += kms;
    }
  }
  if (se != hipSuccess || hipGetLastError() != hipSuccess) return PMG_EDEVICE;
  c->iters = S.iters;
  c->relres = S.rsnrm / S.rsden;
  c->ierror = S.rc == PMG_ITMAX ? 1 : 0;
  return S.rc;
}

// ---- grouped iterations (the serving loop) ----
// One iteration of several sessions at once: the step of step_body for
// every context, each kernel launched once for all of them (g_* launchers),
// so that a level's work for up to GMAX requests costs one launch.  Every
// context runs exactly the kernels, workgroups and arithmetic of its own
// step, in its own order: results are bit-identical to solving it alone.

// Contexts whose iterations can run grouped: a residual stopping test, two
// or more levels, no per-kernel timing.
bool group_ok(const pmg_ctx *c) {
  return c->p.istop <= 1 && c->nlev >= 2 && !c->ktime;
}

// smooth() of level lev for the contexts qs (x = L[lev].x, f[k] its
// right-hand side).
template <class T>
void g_smooth_level(LV<T> *const *L, const std::vector<int> &qs, int lev,
                    const std::vector<const T *> &f, int nu, int iadj,
                    bool zero_init, bool bnd_zero, hipStream_t s) {
  if (qs.empty()) return;
  if (nu <= 0) {
    if (zero_init)
      for (int q : qs) fill<T>(L[q][lev].x, L[q][lev].d.n(), (T)0, s);
    return;
  }
  std::vector<GSmooth<T>> j7, j27;
  for (size_t k = 0; k < qs.size(); k++) {
    const LV<T> &l = L[qs[k]][lev];
    T *x = l.x;
    if (l.numdia == 7) {
      if (l.t && l.t != x) {
        if (zero_init && !bnd_zero) fill<T>(x, l.d.n(), (T)0, s);
        j7.push_back(GSmooth<T>{l.d, l.ac, l.cc, f[k], x, l.t, l.dg});
      } else {
        if (zero_init) fill<T>(x, l.d.n(), (T)0, s);
        gs7<T>(l.d, l.ac, l.cc, f[k], x, nu, iadj, s);
      }
    } else {
      j27.push_back(GSmooth<T>{l.d, l.ac, l.cc, f[k], x,
                               l.t != x ? l.t : nullptr});
    }
  }
  g_gs7_fused<T>(j7.data(), (int)j7.size(), nu, iadj, zero_init && bnd_zero,
                 s);
  g_gs27<T>(j27.data(), (int)j27.size(), nu, iadj, zero_init, s);
}

// vcycle() for n contexts at once (same nu1, nu2): level by level from the
// top; a context takes part in the levels it has, and its coarsest solve
// runs once every context has restricted to its own coarsest level.
template <class T>
void g_vcycle(pmg_ctx *const *c, int n, LV<T> *const *L,
              const std::vector<const T *> &frhs, bool fine_zero,
              double coarse_tol, hipStream_t s) {
  const int nu1 = c[0]->p.nu1, nu2 = c[0]->p.nu2;
  int lmax = 0;
  for (int q = 0; q < n; q++) lmax = std::max(lmax, c[q]->nlev);
  std::vector<int> all(n);
  for (int q = 0; q < n; q++) all[q] = q;
  g_smooth_level<T>(L, all, 0, frhs, nu1, 0, fine_zero, fine_zero, s);
  {
    std::vector<GResid<T, T>> j;
    for (int q = 0; q < n; q++) {
      const LV<T> &l = L[q][0];
      j.push_back(GResid<T, T>{l.d, l.numdia, l.ac, l.cc, frhs[q], l.x, l.r,
                               nullptr, nullptr, l.dg});
    }
    g_resid<T, T>(j.data(), n, s);
  }
  for (int l = 1; l < lmax; l++) {  // down
    std::vector<GRestrict<T>> jr;
    std::vector<int> mid;
    std::vector<const T *> fm;
    for (int q = 0; q < n; q++) {
      if (c[q]->nlev <= l) continue;
      jr.push_back(GRestrict<T>{L[q][l - 1].d, L[q][l].d, L[q][l - 1].r,
                                L[q][l].w0, L[q][l].pc});
      if (l != c[q]->nlev - 1) {
        mid.push_back(q);
        fm.push_back(L[q][l].w0);
      }
    }
    g_restrict<T>(jr.data(), (int)jr.size(), s);
    g_smooth_level<T>(L, mid, l, fm, nu1, 0, true, false, s);
    std::vector<GResid<T, T>> j;
    for (int q : mid) {
      const LV<T> &v = L[q][l];
      j.push_back(GResid<T, T>{v.d, v.numdia, v.ac, v.cc, v.w0, v.x, v.r,
                               nullptr, nullptr});
    }
    g_resid<T, T>(j.data(), (int)j.size(), s);
  }
  {  // coarsest levels
    std::vector<GCoarse<T>> jc;
    for (int q = 0; q < n; q++) {
      const int lc = c[q]->nlev - 1;
      const Level &C = c[q]->L[lc];
      if (c[q]->p.mgsolv == 1) {
        band_solve<T>(C.d, C.numdia, c[q]->band_rt, c[q]->band_ct,
                      L[q][lc].w0, L[q][lc].x, s);
      } else {
        jc.push_back(GCoarse<T>{C.d, C.numdia, C.ac, C.cc, L[q][lc].w0,
                                L[q][lc].x});
      }
    }
    g_cg<T>(jc.data(), (int)jc.size(), 100, coarse_tol, s);
  }
  for (int l = lmax - 2; l >= 0; l--) {  // up
    std::vector<int> qs;
    for (int q = 0; q < n; q++)
      if (c[q]->nlev - 2 >= l) qs.push_back(q);
    if (qs.empty()) continue;
    const int cl = l + 1;
    std::vector<GMatvec<T>> jm;
    for (int q : qs) {
      const LV<T> &v = L[q][cl];
      jm.push_back(GMatvec<T>{v.d, v.numdia, v.ac, v.cc, v.x, v.w0, v.t,
                              c[q]->part});
    }
    std::vector<unsigned> nb(qs.size());
    g_matvec_dots_part<T>(jm.data(), (int)jm.size(), nb.data(), s);
    std::vector<GInterp<T>> ji;
    std::vector<const T *> fu;
    for (size_t k = 0; k < qs.size(); k++) {
      const int q = qs[k];
      ji.push_back(GInterp<T>{L[q][cl].d, L[q][l].d, L[q][cl].x, L[q][cl].pc,
                              L[q][l].x, c[q]->part, nb[k]});
      fu.push_back(l == 0 ? frhs[q] : L[q][l].w0);
    }
    g_interp_damped<T>(ji.data(), (int)ji.size(), s);
    g_smooth_level<T>(L, qs, l, fu, nu2, 1, false, false, s);
  }
}

// One iteration (step_body) of the sessions of c[0..n-1], all in mode and
// with the same nu1, nu2, on stream s; each context's stopping norm goes
// to its device scalar and to norm[q].
void group_step(pmg_ctx *const *c, int n, int mode, hipStream_t s,
                double *norm) {
  std::vector<LV<double>> v64((size_t)n * MAXLEV);
  std::vector<LV<float>> v32((size_t)n * MAXLEV);
  std::vector<LV<double> *> L64(n);
  std::vector<LV<float> *> L32(n);
  std::vector<Red> rn(n);
  for (int q = 0; q < n; q++) {
    L64[q] = &v64[(size_t)q * MAXLEV];
    L32[q] = &v32[(size_t)q * MAXLEV];
    views(c[q], L64[q], L32[q], mode);
    rn[q] = red_to(c[q], S_RSNRM);
  }
  if (mode == PMG_MIXED) {
    std::vector<const float *> f(n);
    for (int q = 0; q < n; q++) f[q] = c[q]->L[0].w032;
    g_vcycle<float>(c, n, L32.data(), f, true, kMixedCoarseTol, s);
    std::vector<GAddcorr<float>> ja;
    std::vector<GResid<double, float>> jr;
    for (int q = 0; q < n; q++) {
      Level &F = c[q]->L[0];
      ja.push_back(GAddcorr<float>{F.d, L32[q][0].x, c[q]->u});
      jr.push_back(GResid<double, float>{F.d, F.numdia, F.ac, F.cc, F.fc,
                                         c[q]->u, F.w032, &rn[q], norm + q,
                                         F.dg});
    }
    g_addcorr<float>(ja.data(), n, s);
    g_resid<double, float>(jr.data(), n, s);
  } else {
    std::vector<const double *> f(n);
    for (int q = 0; q < n; q++) f[q] = c[q]->L[0].fc;
    g_vcycle<double>(c, n, L64.data(), f, false, 0.0, s);
    std::vector<GResid<double, double>> jr;
    for (int q = 0; q < n; q++) {
      Level &F = c[q]->L[0];
      jr.push_back(GResid<double, double>{F.d, F.numdia, F.ac, F.cc, F.fc,
                                          c[q]->u, nullptr, &rn[q], norm + q,
                                          F.dg});
    }
    g_resid<double, double>(jr.data(), n, s);
  }
}

// Solve the contexts' linear systems together, their iterations grouped
// (group_step) until each has met its stopping test; a context leaves the
// group when it finishes.  Blocking.
int ctx_solve_group(pmg_ctx *const *cs, int n, int mode) {
  for (int q = 0; q < n; q++)
    if (!cs[q] || cs[q]->p.mode != 0 || !group_ok(cs[q]) ||
        cs[q]->p.nu1 != cs[0]->p.nu1 || cs[q]->p.nu2 != cs[0]->p.nu2)
      return PMG_EBADPARM;
  for (int q = 0; q < n; q++)
    for (int r = 0; r < q; r++)
      if (cs[q] == cs[r]) return PMG_EBADPARM;
  int rc = PMG_OK;
  std::vector<pmg_ctx *> act;
  for (int q = 0; q < n; q++) {
    const int r = lin_begin(cs[q], mode);
    if (r) {  // close the sessions already started, then report
      for (int k = 0; k < q; k++) {
        lin_poll(cs[k], true);
        lin_finish(cs[k], nullptr);
        lin_end(cs[k]);
      }
      return r;
    }
  }
  for (int q = 0; q < n; q++) {
    if (lin_poll(cs[q], true))
      lin_finish(cs[q], nullptr);
    else
      act.push_back(cs[q]);
  }
  hipStream_t gs = nullptr;
  hipEvent_t ev = nullptr;
  double *dn = nullptr, *hn = nullptr;
  if (hipStreamCreateWithFlags(&gs, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&dn, sizeof(double) * (n + 1)) != hipSuccess ||
      hipHostMalloc((void **)&hn, sizeof(double) * (n + 1)) != hipSuccess)
    rc = PMG_EDEVICE;
  while (!rc && !act.empty()) {
    group_step(act.data(), (int)act.size(), mode, gs, dn);
    hipMemcpyAsync(hn, dn, sizeof(double) * act.size(), hipMemcpyDeviceToHost,
                   gs);
    hipEventRecord(ev, gs);
    if (hipEventSynchronize(ev) != hipSuccess) {
      rc = PMG_EDEVICE;
      break;
    }
    std::vector<pmg_ctx *> next;
    for (size_t k = 0; k < act.size(); k++) {
      pmg_ctx *c = act[k];
      c->hscal[S_RSNRM] = hn[k];
      c->ses.pending = false;
      if (ses_cycle_done(c)) {
        hipStreamWaitEvent(c->s, ev, 0);
        lin_finish(c, nullptr);
      } else {
        next.push_back(c);
      }
    }
    act.swap(next);
  }
  for (pmg_ctx *c : act) {  // only after a device error
    hipStreamWaitEvent(c->s, ev, 0);
    lin_finish(c, nullptr);
  }
  if (gs) hipStreamSynchronize(gs);
  for (int q = 0; q < n; q++) {
    const int r = lin_end(cs[q]);
    if (!rc && r != PMG_OK && r != PMG_ITMAX) rc = r;
  }
  if (ev) hipEventDestroy(ev);
  if (dn) hipFree(dn);
  if (hn) hipHostFree(hn);
  if (gs) hipStreamDestroy(gs);
  return rc;
}

int ctx_solve(pmg_ctx *c, int mode) {
  int rc = lin_begin(c, mode);
  if (rc) return rc;
  while (!lin_poll(c, true)) lin_enqueue(c);
  lin_finish(c, nullptr);
  return lin_end(c);
}

// HBM bytes of a context for p, as allocated by ctx_new (+ alloc_fp32 for
// PMG_MIXED): per level the operator planes, cc, fc, P and four work
// vectors; the resident fine inputs; padding of every allocation.
size_t ctx_bytes_estimate(const Params &p, int mode) {
  size_t b64 = 0, b32 = 0, nalloc = 0;
  Dim d{p.nx, p.ny, p.nz}, dc = d;
  for (int l = 0; l < p.nlev; l++) {
    const size_t n = d.n();
    dc = d;
    b64 += ((l == 0 ? 4 : 14) + 2 + (l ? 27 : 0) + 4) * n;
    b32 += ((l == 0 ? 4 : 14) + 1 + (l ? 27 : 0) + 4) * n;
    nalloc += 16;
    d = Dim{corsr(d.nx), corsr(d.ny), corsr(d.nz)};
  }
  const size_t nf = (size_t)p.nx * p.ny * p.nz;
  b64 += nf + 2 * max_partials(Dim{p.nx, p.ny, p.nz}) + 16;  // u, partials
  b64 += (size_t)p.nx + p.ny + p.nz + 2 * ((size_t)p.ny * p.nz +
         (size_t)p.nx * p.nz + (size_t)p.nx * p.ny) + 5 * nf;  // inputs
  nalloc += 16;
  if (p.mgsolv == 1) {  // band and its row and column copies
    const BandDims b = band_dims(dc, p.nlev == 1 ? 7 : 27);
    b64 += (size_t)b.n * b.lda + 2 * (size_t)b.nrow * b.W;
    nalloc += 3;
  }
  size_t bytes = 8 * b64 + (mode == PMG_MIXED ? 4 * b32 : 0) + 256 * nalloc;
  return bytes + 6 * sizeof(double) * 1024 + (1 << 20);  // atoms, slack
}

// ---- Newton (Vnewton, newtond.cc:140-486) ----
int alloc_newton(pmg_ctx *c) {
  if (c->haveN) return PMG_OK;
  for (int l = 0; l < c->nlev; l++) {
    Level &L = c->L[l];
    L.cpr = c->alloc<double>(L.d.n());
    L.rhs = c->alloc<double>(L.d.n());
    if (l == 0) L.dgn = c->alloc<double>(L.d.n());
    // the Newton iterate of a coarse level (nested iteration, mgkey 1)
    if (l > 0 && c->p.mgkey == 1) L.xn = c->alloc<double>(L.d.n());
  }
  for (int q = 0; q < 4; q++) c->work[q] = c->alloc<double>(c->L[0].d.n());
  if (!c->ok) return PMG_ENOMEM;
  c->haveN = true;
  return PMG_OK;
}

// L1 norm of the nonlinear residual of level l at x (Vnmresid + Vxnrm1),
// r receives the residual.
double nm_norm(pmg_ctx *c, int l, const double *x, double *r) {
  Level &F = c->L[l];
  const Red rd = red_to(c, S_TMP);
  nmresid(F.d, F.numdia, F.ac, F.cc, F.fc, x, r, &rd, c->s);
  return host_scalar(c, S_TMP);
}

// Vnewton (newtond.cc:140-486) on levels l0.. of the hierarchy (level l0 as
// the finest) from the iterate x of level l0 as it stands: damped inexact
// Newton whose inner solve (FP64, or FP32 cycles under FP64 defect
// correction) runs to min(0.9 |r|, |r|^2) in the absolute L1 residual;
// stopping test istop 0 / 1 on the nonlinear residual.
int newton_level(pmg_ctx *c, int mode, int l0, double *x, int itmax,
                 int *iters_out, double *rel_out) {
  const Params &p = c->p;
  Level &F = c->L[l0];
  const Dim d = F.d;
  const size_t nf = d.n();
  hipStream_t s = c->s;
  double *w0n = c->work[0], *w1 = c->work[1], *t1 = c->work[2],
         *t2 = c->work[3];
  for (int q = 0; q < 4; q++) fill<double>(c->work[q], nf, 0.0, s);
  double rsden = 1.0;
  if (p.istop == 1) rsden = nm_norm(c, l0, w1, t2);  // residual at zero
  if (rsden == 0.0) rsden = 1.0;
  double xold = nm_norm(c, l0, x, w0n);
  double xnew = xold, xmed = xold, rsnrm = rsden;
  int idamp = 1, iters = 0;
  LV<double> L64[MAXLEV];
  LV<float> L32[MAXLEV];
  // The Jacobian term replaces cc; on the 7-point fine level its diagonal
  // oC + cpr is written beside cpr each iteration (dc_vec), elsewhere the
  // kernels form it themselves.
  const bool fine7 = l0 == 0 && F.numdia == 7 && F.dgn &&
                     (mode != PMG_MIXED || F.dgn32);
  for (int l = l0; l < c->nlev; l++) {
    L64[l] = view<double>(c->L[l]);
    L64[l].cc = c->L[l].cpr;
    L64[l].dg = (fine7 && l == l0) ? F.dgn : nullptr;
    if (mode == PMG_MIXED) {
      L32[l] = view<float>(c->L[l]);
      L32[l].cc = c->L[l].cpr32;
      L32[l].dg = (fine7 && l == l0) ? F.dgn32 : nullptr;
    }
  }
  double *xtmp = F.x;  // iterate of the inner linear solve
  L64[l0].x = xtmp;
  for (;;) {
    iters++;
    // Vgetjac (newtond.cc:488-531)
    // the right-hand side of the Newton correction is the nonlinear
    // residual w0n itself (interior; the boundary is zero), read in place;
    // the FP32 copies of the Helmholtz term for mixed inner solves are
    // written by the kernels that produce it
    const bool m32 = mode == PMG_MIXED;
    dc_vec(F.cc, x, F.cpr, nf, s, m32 ? F.cpr32 : nullptr,
           fine7 ? F.ac : nullptr, fine7 ? F.dgn : nullptr,
           fine7 && m32 ? F.ac32 : nullptr, fine7 && m32 ? F.dgn32 : nullptr);
    for (int l = l0 + 1; l < c->nlev; l++)
      restrict2(c->L[l - 1].d, c->L[l].d, l - 1 == l0 ? w0n : c->L[l - 1].rhs,
                c->L[l - 1].cpr, c->L[l].rhs, c->L[l].cpr, c->L[l].pc, s,
                nullptr, m32 ? c->L[l].cpr32 : nullptr);
    const double errtol_s = std::min(0.9 * xold, 1.0 * std::pow(xold, 2.0));
    fill<double>(xtmp, nf, 0.0, s);
    int its = 0;
    double rel = 0.0;
    const double *cc64c = c->L[c->nlev - 1].cpr;
    if (mode == PMG_MIXED)
      linear_mixed(c, L64 + l0, L32 + l0, w0n, cc64c, xtmp, 0, errtol_s,
                   1000, &its, &rel, l0);
    else
      linear_fp64(c, L64 + l0, w0n, cc64c, 0, errtol_s, 1000, &its, &rel,
                  l0);
    if (idamp == 1) {  // damping (newtond.cc:313-376)
      xpay_interior(d, x, 1.0, xtmp, w1, s);
      xnew = nm_norm(c, l0, w1, w0n);
      double damp = 1.0;
      int iter_d = 0, dmode = 0;
      while (iter_d < 10) {
        if (dmode == 0) {
          if (xnew < xold) dmode = 1;
        } else if (xnew > xmed) {
          break;
        }
        copy_interior(d, w1, t1, s);
        copy_interior(d, w0n, t2, s);
        xmed = xnew;
        damp = damp / 2.0;
        xpay_interior(d, x, damp, xtmp, w1, s);
        xnew = nm_norm(c, l0, w1, w0n);
        iter_d++;
      }
      copy_interior(d, t1, x, s);
      copy_interior(d, t2, w0n, s);
      xnew = xmed;
      xold = xnew;
      if (iter_d - 1 == 0) idamp = 0;
    } else {
      axpy_interior<double>(d, 1.0, xtmp, x, s);
      xnew = nm_norm(c, l0, x, w0n);
      xold = xnew;
    }
    rsnrm = xnew;
    if (l0 == 0) {
      c->hist.push_back(rsnrm / rsden);
      if (p.iinfo > 0)  // Vprtstp (mgsubd.cc:486-536)
        printf("iteration = %d relative residual = %e\n", iters,
               rsnrm / rsden);
    }
    if (rsnrm / rsden <= p.errtol) break;
    if (iters >= itmax) break;
  }
  *iters_out = iters;
  *rel_out = rsnrm / rsden;
  return *rel_out <= p.errtol ? PMG_OK : PMG_ITMAX;
}

// Vnewton (mgkey 0) or Vfnewton (mgkey 1, newtond.cc:60-138, its level loop
// run from the coarsest level up, as its level-- intends):
// Newton on each level from the coarsest up, each level's solution
// prolongated (VinterpPMG) as the next finer level's starting iterate; the
// coarse iterates start from zero.  The fine iterate starts from c->u when
// u_given is set (the caller's u of pmg_newdriv, boundary zeroed,
// VfboundPMG00 newdrvd.cc:263), else from zero.
int ctx_newton(pmg_ctx *c, int mode) {
  const Params &p = c->p;
  if (p.istop > 1) return PMG_EBADPARM;  // the residual tests only
  int rc = alloc_newton(c);
  if (rc) return rc;
  if (mode == PMG_MIXED) {
    rc = alloc_fp32(c);
    if (rc) return rc;
    if (!c->ops32_ok) convert_ops32(c);
    c->ops32_ok = true;
    for (int l = 0; l < c->nlev; l++)
      if (!c->L[l].cpr32) c->L[l].cpr32 = c->alloc<float>(c->L[l].d.n());
    if (!c->L[0].dgn32) c->L[0].dgn32 = c->alloc<float>(c->L[0].d.n());
    if (!c->ok) return PMG_ENOMEM;
  }
  if (!coarse_fits(c)) return PMG_EBADPARM;
  if (p.mgsolv == 1 && host_scalar(c, S_BINFO) != 0.0) return PMG_ENOTSPD;
  Level &F = c->L[0];
  const Dim d = F.d;
  hipStream_t s = c->s;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, s);
  c->hist.clear();
  if (c->u_given)
    fbound00(d, c->u, s);
  else
    fill<double>(c->u, d.n(), 0.0, s);
  c->u_given = false;
  int iters = 0;
  double rel = 0.0;
  if (p.mgkey == 1) {
    for (int l = 1; l < c->nlev; l++)
      fill<double>(c->L[l].xn, c->L[l].d.n(), 0.0, s);
    for (int l = c->nlev - 1; l >= 1; l--) {
      newton_level(c, mode, l, c->L[l].xn, 1000, &iters, &rel);
      interp<double>(c->L[l].d, c->L[l - 1].d, c->L[l].xn, c->L[l].pc,
                     l - 1 == 0 ? c->u : c->L[l - 1].xn, nullptr, s);
    }
  }
  rc = newton_level(c, mode, 0, c->u, p.itmax, &iters, &rel);
  fbound_dirichlet(d, c->u, c->gx, c->gy, c->gz, s);
  hipEventRecord(e1, s);
  hipError_t se = hipStreamSynchronize(s);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  c->solve_ms = ms;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (se != hipSuccess) return PMG_EDEVICE;
  c->iters = iters;
  c->relres = rel;
  c->ierror = rc == PMG_OK ? 0 : 1;
  return rc;
}

// ---- class-style solver (PBE_pmgc_class.h:1070-1133, CG_Solver.h:45-242)
// CG preconditioned by the class V-cycle, on this context's hierarchy.

enum { S_PCG_RS = 6, S_PCG_PAP = 8, S_PCG_RN1 = 9 };  // RS: 6 and 7 alternate

// x_fine = M rhs: per level two red-black sweeps from zero (iadjoint 0),
// the defect restricted, the recursion, the Hackbusch-Reusken factor of the
// coarse solution folded into the prolongation (the same partial sums and
// k_interp as the PMG cycle), two sweeps again with iadjoint 0.  The
// coarsest level runs the class CG from its previous solution.
template <class T>
void class_vcycle(pmg_ctx *c, LV<T> *L, const T *frhs) {
  const int nlev = c->nlev;
  hipStream_t s = c->s;
  for (int l = 0; l < nlev - 1; l++) {
    const T *rhs = l == 0 ? frhs : L[l].w0;
    smooth<T>(c, L[l], rhs, L[l].x, 2, 0, false, true);
    resid<T, T>(L[l].d, L[l].numdia, L[l].ac, L[l].cc, rhs, L[l].x, L[l].r,
                nullptr, s);
    restrict_<T>(L[l].d, L[l + 1].d, L[l].r, L[l + 1].w0, L[l + 1].pc, s);
  }
  {
    const Level &C = c->L[nlev - 1];
    const int ni = (int)C.d.interior();
    cg_class<T>(C.d, C.numdia, C.ac, C.cc, L[nlev - 1].w0, L[nlev - 1].x,
                std::max(1000, 4 * ni), 1e-8, s);
  }
  for (int l = nlev - 2; l >= 0; l--) {
    const int cl = l + 1;
    const unsigned nb =
        matvec_dots_part<T>(L[cl].d, L[cl].numdia, L[cl].ac, L[cl].cc,
                            L[cl].x, L[cl].w0, L[cl].t, c->part, s);
    interp_damped<T>(L[cl].d, L[l].d, L[cl].x, L[cl].pc, L[l].x, c->part, nb,
                     s);
    smooth<T>(c, L[l], l == 0 ? frhs : L[l].w0, L[l].x, 2, 0, false);
  }
}

// Preconditioner application s = M r (FP64 r and s; PMG_MIXED runs the
// cycle in FP32).
void pcg_precond(pmg_ctx *c, int mode, LV<double> *L64, LV<float> *L32,
                 const double *r, double *sv) {
  Level &F = c->L[0];
  const size_t nf = F.d.n();
  if (mode == PMG_MIXED) {
    convert<double, float>(r, F.w032, nf, c->s);
    class_vcycle<float>(c, L32, F.w032);
    convert<float, double>(F.x32, sv, nf, c->s);
  } else {
    class_vcycle<double>(c, L64, r);  // result in F.x == sv
  }
}

// num_iters > 0: exactly that many CG iterations (counting_criterion,
// CG_Solver.h:38-43); 0: until ||r||_1 / ||f||_1 <= errtol, or itmax.
int ctx_pcg(pmg_ctx *c, int mode, int num_iters) {
  const Params &p = c->p;
  if (c->nlev < 2 || num_iters < 0) return PMG_EBADPARM;
  const Level &C = c->L[c->nlev - 1];
  if (!cg_class_fits(C.d, C.numdia)) return PMG_EBADPARM;
  if (mode == PMG_MIXED) {
    int rc = alloc_fp32(c);
    if (rc) return rc;
    if (!c->ops32_ok) convert_ops32(c);
    c->ops32_ok = true;
  }
  Level &F = c->L[0];
  const Dim d = F.d;
  const size_t nf = d.n();
  if (!c->pcg[0]) {
    for (int q = 0; q < 4; q++) c->pcg[q] = c->alloc<double>(nf);
    if (!c->ok) return PMG_ENOMEM;
  }
  hipStream_t s = c->s;
  double *r = c->pcg[0], *pv = c->pcg[1], *q = c->pcg[2];
  double *sv = mode == PMG_MIXED ? c->pcg[3] : F.x;
  LV<double> L64[MAXLEV];
  LV<float> L32[MAXLEV];
  for (int l = 0; l < c->nlev; l++) {
    L64[l] = view<double>(c->L[l]);
    if (mode == PMG_MIXED) L32[l] = view<float>(c->L[l]);
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, s);
  // x0 = 0; the coarsest solution, which the class CG starts from, too
  fill<double>(c->u, nf, 0.0, s);
  if (mode == PMG_MIXED)
    fill<float>(c->L[c->nlev - 1].x32, C.d.n(), 0.0f, s);
  else
    fill<double>(c->L[c->nlev - 1].x, C.d.n(), 0.0, s);
  const Red rden = red_to(c, S_RSDEN), rn1 = red_to(c, S_PCG_RN1),
            rpap = red_to(c, S_PCG_PAP);
  nrm1<double>(d, F.fc, rden, s);
  double rsden = host_scalar(c, S_RSDEN);
  if (rsden == 0.0) rsden = 1.0;
  // r = f - A x0, s = M r, p = s, rs = r.s, pap = p.A p
  resid<double, double>(d, F.numdia, F.ac, F.cc, F.fc, c->u, r, nullptr, s);
  pcg_precond(c, mode, L64, L32, r, sv);
  copy_interior(d, sv, pv, s);
  dot_interior(d, r, sv, red_to(c, S_PCG_RS), s);
  matvec_dots<double>(d, F.numdia, F.ac, F.cc, pv, pv, q, nullptr, s);
  dot_interior(d, pv, q, rpap, s);
  nrm1<double>(d, r, rn1, s);
  hipMemcpyAsync(c->hscal + S_PCG_PAP, c->scal + S_PCG_PAP,
                 2 * sizeof(double), hipMemcpyDeviceToHost, s);
  hipStreamSynchronize(s);
  const double pap0 = c->hscal[S_PCG_PAP];
  double rel = c->hscal[S_PCG_RN1] / rsden;
  int iters = 0, ierr = 0;
  if (pap0 != 0.0) {  // CG_Solver.h:124-125
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
      const double *rs_old = c->scal + S_PCG_RS + (iters & 1);
      double *rs_new = c->scal + S_PCG_RS + ((iters + 1) & 1);
      iters++;
      pcg_xr(d, c->u, r, pv, q, rs_old, c->scal + S_PCG_PAP, s);
      pcg_precond(c, mode, L64, L32, r, sv);
      Red rrs = red_to(c, S_PCG_RS);
      rrs.out = rs_new;
      dot_interior(d, r, sv, rrs, s);
      pcg_p(d, pv, sv, rs_new, rs_old, s);
      matvec_dots<double>(d, F.numdia, F.ac, F.cc, pv, pv, q, nullptr, s);
      dot_interior(d, pv, q, rpap, s);
      if (num_iters == 0) {
        nrm1<double>(d, r, rn1, s);
        rel = host_scalar(c, S_PCG_RN1) / rsden;
      }
    }
  }
  // the reported residual: that of the final u
  const Red rnrm = red_to(c, S_RSNRM);
  resid<double, double>(d, F.numdia, F.ac, F.cc, F.fc, c->u, (double *)nullptr,
                        &rnrm, s);
  const double rsnrm = host_scalar(c, S_RSNRM);
  fbound_dirichlet(d, c->u, c->gx, c->gy, c->gz, s);
  hipEventRecord(e1, s);
  hipError_t se = hipStreamSynchronize(s);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  c->solve_ms = ms;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (se != hipSuccess || hipGetLastError() != hipSuccess) return PMG_EDEVICE;
  c->iters = iters;
  c->relres = rsnrm / rsden;
  c->ierror = ierr;
  return ierr ? PMG_ITMAX : PMG_OK;
}

// Upload tcf (fine level) as the comparison solution of istop 2-5.
int ctx_set_tru(pmg_ctx *c, const double *tcf) {
  const size_t n = c->L[0].d.n();
  if (!c->tru) {
    c->tru = c->alloc<double>(n);
    if (!c->ok) return PMG_ENOMEM;
  }
  if (hipMemcpyAsync(c->tru, tcf, n * sizeof(double), hipMemcpyHostToDevice,
                     c->s) != hipSuccess)
    return PMG_EDEVICE;
  return hipStreamSynchronize(c->s) == hipSuccess ? PMG_OK : PMG_EDEVICE;
}

// Fill iwork/rwork as the reference leaves them (mgdrvd.cc:145-166,
// mgsubd.cc:266-373): iz, ipc per level; rpc | cc | fc | pc | ac.
int export_work(pmg_ctx *c, int *iwork, double *rwork, bool newton) {
  int sz[12];
  const Params &p = c->p;
  mgsz(p.mgcoar, p.mgdisc, p.mgsolv, p.nx, p.ny, p.nz, p.nlev, sz);
  const int nf = sz[3], narr = sz[5], narrc = sz[6], nrpc = sz[7], niz = sz[8];
  std::memset(iwork, 0, sizeof(int) * (size_t)sz[11]);
  int *iz = iwork, *ipc = iwork + niz;
  auto IZ = [&](int a, int lev) -> int & { return iz[(lev - 1) * 50 + a - 1]; };
  // Vbuildstr
  Dim d{p.nx, p.ny, p.nz};
  int n = (int)d.n();
  for (int a = 1; a <= 11; a++) IZ(a, 1) = 1;
  IZ(1, 2) = IZ(1, 1) + n;
  IZ(2, 2) = IZ(2, 1) + 4 * d.ny * d.nz;
  IZ(3, 2) = IZ(3, 1) + 4 * d.nx * d.nz;
  IZ(4, 2) = IZ(4, 1) + 4 * d.nx * d.ny;
  IZ(5, 2) = IZ(5, 1) + 100;
  IZ(6, 2) = IZ(6, 1) + 100;
  IZ(8, 2) = IZ(8, 1) + d.nx;
  IZ(9, 2) = IZ(9, 1) + d.ny;
  IZ(10, 2) = IZ(10, 1) + d.nz;
  for (int lev = 2; lev <= p.nlev; lev++) {
    d = Dim{corsr(d.nx), corsr(d.ny), corsr(d.nz)};
    n = (int)d.n();
    IZ(1, lev + 1) = IZ(1, lev) + n;
    IZ(2, lev + 1) = IZ(2, lev) + 4 * d.ny * d.nz;
    IZ(3, lev + 1) = IZ(3, lev) + 4 * d.nx * d.nz;
    IZ(4, lev + 1) = IZ(4, lev) + 4 * d.nx * d.ny;
    IZ(5, lev + 1) = IZ(5, lev) + 100;
    IZ(6, lev + 1) = IZ(6, lev) + 100;
    IZ(7, lev + 1) = IZ(7, lev) + 4 * n;
    IZ(8, lev + 1) = IZ(8, lev) + d.nx;
    IZ(9, lev + 1) = IZ(9, lev) + d.ny;
    IZ(10, lev + 1) = IZ(10, lev) + d.nz;
    IZ(11, lev) = IZ(11, lev - 1) + 27 * n;
  }
  // operator offsets as Vbuildops marks them
  IZ(7, 2) = IZ(7, 1) + 4 * (int)c->L[0].d.n();
  for (int lev = 2; lev <= p.nlev; lev++)
    IZ(7, lev + 1) = IZ(7, lev) + 14 * (int)c->L[lev - 1].d.n();
  for (int lev = 1; lev <= p.nlev; lev++) {
    int *ipcl = ipc + IZ(5, lev) - 1;
    ipcl[9] = p.ipkey;
    ipcl[10] = lev == 1 ? 7 : 27;
    if (lev == 1) ipcl[11] = 1;
  }
  size_t rtot = (size_t)sz[10] + (newton ? 2 * (size_t)nf : 0);
  std::memset(rwork, 0, sizeof(double) * rtot);
  double *ccb = rwork + nrpc, *fcb = ccb + narr;
  double *after = fcb + narr + (newton ? 2 * (size_t)nf : 0);
  double *pcb = after, *acb = pcb + 27 * (size_t)narrc;
  for (int lev = 1; lev <= p.nlev; lev++) {
    Level &L = c->L[lev - 1];
    size_t nl = L.d.n();
    hipMemcpyAsync(ccb + IZ(1, lev) - 1, L.cc, nl * sizeof(double),
                   hipMemcpyDeviceToHost, c->s);
    hipMemcpyAsync(fcb + IZ(1, lev) - 1, L.fc, nl * sizeof(double),
                   hipMemcpyDeviceToHost, c->s);
    hipMemcpyAsync(acb + IZ(7, lev) - 1, L.ac,
                   (lev == 1 ? 4 : 14) * nl * sizeof(double),
                   hipMemcpyDeviceToHost, c->s);
    if (lev > 1)
      hipMemcpyAsync(pcb + IZ(11, lev - 1) - 1, L.pc,
                     27 * nl * sizeof(double), hipMemcpyDeviceToHost, c->s);
  }
  if (p.mgsolv == 1) {  // factored band at ac(iz(7, nlev+1)) (mgsubd.cc:255)
    const Level &C = c->L[p.nlev - 1];
    const BandDims b = band_dims(C.d, C.numdia);
    int *ipcb = ipc + IZ(5, p.nlev + 1) - 1;  // buildBd.cc:82-85, 295
    ipcb[0] = b.n, ipcb[1] = b.m, ipcb[2] = b.lda, ipcb[3] = 1;
    hipMemcpyAsync(acb + IZ(7, p.nlev + 1) - 1, c->band,
                   (size_t)b.n * b.lda * sizeof(double),
                   hipMemcpyDeviceToHost, c->s);
  }
  return hipStreamSynchronize(c->s) == hipSuccess ? PMG_OK : PMG_EDEVICE;
}

int check_sizes(const Params &p, bool newton) {
  int sz[12];
  int rc = mgsz(p.mgcoar, p.mgdisc, p.mgsolv, p.nx, p.ny, p.nz, p.nlev, sz);
  if (rc) return rc;
  long long need = (long long)sz[10] + (newton ? 2LL * sz[3] : 0);
  if (p.nrwk < need || p.niwk < sz[11]) return PMG_EBADPARM;
  return PMG_OK;
}

// kind 0: Vmgdriv, 1: Vnewdriv, 2: the class-style PCG (num_iters as
// ctx_pcg)
int driver(int *iparm, double *rparm, int *iwork, double *rwork, double *u,
           const double *xf, const double *yf, const double *zf,
           const double *gxcf, const double *gycf, const double *gzcf,
           const double *a1cf, const double *a2cf, const double *a3cf,
           const double *ccf, const double *fcf, const double *tcf, int kind,
           int num_iters = 0) {
  const bool newton = kind == 1;
  t_iters = 0, t_ierror = 0, t_relres = 0.0;
  t_hist.clear();
  if (!iparm || !rparm || !iwork || !rwork || !u || !xf || !yf || !zf ||
      !gxcf || !gycf || !gzcf || !a1cf || !a2cf || !a3cf || !ccf || !fcf)
    return PMG_EBADPARM;
  Params p;
  int rc = decode(iparm, rparm, p);
  if (rc) return rc;
  if (!newton && p.mode != 0) return PMG_EBADPARM;  // FAS not provided
  if (kind == 2 && (p.nlev < 2 || num_iters < 0)) return PMG_EBADPARM;

  if (p.istop >= 2 && (kind != 0 || !tcf)) return PMG_EBADPARM;
  rc = check_sizes(p, newton);
  if (rc) return rc;
  pmg_ctx *c = nullptr;
  rc = ctx_create(&c, p, xf, yf, zf, gxcf, gycf, gzcf, a1cf, a2cf, a3cf, ccf,
                  fcf);
  if (rc) return rc;
  if (newton) {  // the starting iterate (interior of the caller's u)
    if (hipMemcpyAsync(c->u, u, c->L[0].d.n() * sizeof(double),
                       hipMemcpyHostToDevice, c->s) != hipSuccess) {
      delete c;
      return PMG_EDEVICE;
    }
    c->u_given = true;
  }
  if (p.istop >= 2) {  // the comparison solution of the stopping test
    rc = ctx_set_tru(c, tcf);
    if (rc) {
      delete c;
      return rc;
    }
  }
  int src = kind == 1   ? ctx_newton(c, PMG_FP64)
            : kind == 2 ? ctx_pcg(c, PMG_FP64, num_iters)
                        : ctx_solve(c, PMG_FP64);
  if (src == PMG_OK || src == PMG_ITMAX) {
    hipMemcpyAsync(u, c->u, c->L[0].d.n() * sizeof(double),
                   hipMemcpyDeviceToHost, c->s);
    int erc = export_work(c, iwork, rwork, newton);
    if (erc) src = erc;
  }
  t_iters = c->iters;
  t_ierror = c->ierror;
  t_relres = c->relres;
  t_hist = c->hist;
  delete c;
  return src;
}

// ---- single-kernel helpers on host arrays ----
struct Scratch {
  std::vector<void *> v;
  bool ok = true;
  template <class T>
  T *put(const T *h, size_t n) {
    void *d = nullptr;
    if (hipMalloc(&d, n * sizeof(T) + 64) != hipSuccess) {
      ok = false;
      return nullptr;
    }
    v.push_back(d);
    if (h)
      hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice);
    else
      hipMemset(d, 0, n * sizeof(T) + 64);
    return (T *)d;
  }
  ~Scratch() {
    for (void *d : v) hipFree(d);
  }
};

template <class T>
int finish(Scratch &S, T *host, const T *dev, size_t n) {
  if (!S.ok) return PMG_ENOMEM;
  if (hipDeviceSynchronize() != hipSuccess) return PMG_EDEVICE;
  if (hipMemcpy(host, dev, n * sizeof(T), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return PMG_EDEVICE;
  return PMG_OK;
}

// ---- fixed device scratch of the allocation-free entry point ----
// The block partials of pmg_mresid_mixed_noalloc's L1 norm live in this
// array (one copy per GPU, held until the process ends).  Host code uses
// only the device address hipGetSymbolAddress returns for it, looked up
// once per device with that device current; the array's name itself is a
// host-side stand-in, not a device address.  Calls that overlap on one GPU
// share the array and race.
constexpr size_t kScratchDoubles = (size_t)1 << 17;  // 1 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of the current device's scratch (nullptr on failure).
double *scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "pmg: no device address for the scratch array: %s\n",
              hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = (double *)p;
  }
  return g_scratch_dev[dev];
}

bool dims_ok(int nx, int ny, int nz) { return nx >= 3 && ny >= 3 && nz >= 3; }
bool coarse_of(int f, int c) { return f == 2 * c - 1; }
bool coarse_dims(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc) {
  return dims_ok(nxc, nyc, nzc) && coarse_of(nxf, nxc) &&
         coarse_of(nyf, nyc) && coarse_of(nzf, nzc);
}

// The single-kernel entry points in either precision: each runs exactly the
// launcher the V-cycle of that precision runs.
template <class T>
int mresid_t(int nx, int ny, int nz, int numdia, const T *ac, const T *cc,
             const T *fc, const T *x, T *r) {
  if (!dims_ok(nx, ny, nz) || (numdia != 7 && numdia != 27) || !ac || !cc ||
      !fc || !x || !r)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim d{nx, ny, nz};
  size_t n = d.n();
  Scratch S;
  T *dac = S.put(ac, (numdia == 7 ? 4 : 14) * n), *dcc = S.put(cc, n),
    *dfc = S.put(fc, n), *dx = S.put(x, n), *dr = S.put(r, n);
  // 7-point levels: the precomputed diagonal, as the cycle's residual
  T *ddg = numdia == 7 ? S.put((const T *)nullptr, n) : nullptr;
  if (!S.ok) return PMG_ENOMEM;
  if (ddg) diag_sum<T>(dac, dcc, ddg, n, 0);
  resid<T, T>(d, numdia, dac, dcc, dfc, dx, dr, nullptr, 0, ddg);
  return finish(S, r, dr, n);
}

// zero_start: the sweeps start from x = 0 (Vazeros) and x's interior is
// not read, as in the first pre-smoothing of a correction cycle; x's
// boundary entries must be zero then (as on every level of that cycle).
template <class T>
int gsrb_t(int nx, int ny, int nz, int numdia, const T *ac, const T *cc,
           const T *fc, T *x, int itmax, int iadjoint, int zero_start) {
  if (!dims_ok(nx, ny, nz) || (numdia != 7 && numdia != 27) || !ac || !cc ||
      !fc || !x || itmax < 0 || (iadjoint != 0 && iadjoint != 1) ||
      (zero_start != 0 && zero_start != 1))
    return PMG_EBADPARM;
  Dim d{nx, ny, nz};
  size_t n = d.n();
  if (zero_start)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++)
          if ((i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) &&
              x[((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1)] != (T)0)
            return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Scratch S;
  T *dac = S.put(ac, (numdia == 7 ? 4 : 14) * n), *dcc = S.put(cc, n),
    *dfc = S.put(fc, n), *dx = S.put(x, n);
  // second buffer of the sweeps: the same boundary values as x
  T *tmp = S.put(x, n);
  if (!S.ok) return PMG_ENOMEM;
  if (numdia == 7)
    gs7_fused<T>(d, dac, dcc, dfc, dx, tmp, itmax, iadjoint, 0,
                 zero_start != 0);
  else
    gs27<T>(d, dac, dcc, dfc, dx, itmax, iadjoint, 0, zero_start != 0, tmp);
  return finish(S, x, dx, n);
}

template <class T>
int restrc_t(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
             const T *xin, T *xout, const T *pc) {
  if (!coarse_dims(nxf, nyf, nzf, nxc, nyc, nzc) || !xin || !xout || !pc)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim f{nxf, nyf, nzf}, c{nxc, nyc, nzc};
  Scratch S;
  T *dxi = S.put(xin, f.n()), *dxo = S.put(xout, c.n()),
    *dpc = S.put(pc, 27 * c.n());
  if (!S.ok) return PMG_ENOMEM;
  restrict_<T>(f, c, dxi, dxo, dpc, 0);
  return finish(S, xout, dxo, c.n());
}

template <class T>
int interp_t(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
             const T *xin, T *xout, const T *pc) {
  if (!coarse_dims(nxf, nyf, nzf, nxc, nyc, nzc) || !xin || !xout || !pc)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim f{nxf, nyf, nzf}, c{nxc, nyc, nzc};
  Scratch S;
  T *dxi = S.put(xin, c.n()), *dxo = S.put(xout, f.n()),
    *dpc = S.put(pc, 27 * c.n());
  if (!S.ok) return PMG_ENOMEM;
  interp<T>(c, f, dxi, dpc, dxo, nullptr, 0);
  return finish(S, xout, dxo, f.n());
}

// xf(interior) += xdamp * (P xc), xdamp = (xc . w0c) / (xc . A_c xc): the
// coarse-grid correction of the V-cycle's way up (mgcsd.cc:429-458), as
// the cycle runs it: k_matvec_dots partials, then k_interp<T, true>.
template <class T>
int correct_t(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
              int numdia, const T *acc, const T *ccc, const T *xc,
              const T *w0c, const T *pc, T *xf) {
  if (!coarse_dims(nxf, nyf, nzf, nxc, nyc, nzc) ||
      (numdia != 7 && numdia != 27) || !acc || !ccc || !xc || !w0c || !pc ||
      !xf)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim f{nxf, nyf, nzf}, c{nxc, nyc, nzc};
  Scratch S;
  T *dac = S.put(acc, (numdia == 7 ? 4 : 14) * c.n()), *dcc = S.put(ccc, c.n()),
    *dxc = S.put(xc, c.n()), *dw0 = S.put(w0c, c.n()),
    *dpc = S.put(pc, 27 * c.n()), *dxf = S.put(xf, f.n()),
    *dy = S.put((const T *)nullptr, c.n());
  // two partials (x.w0, x.Ax) per block of the matvec
  double *part = S.put((const double *)nullptr, 2 * max_partials(c));
  if (!S.ok) return PMG_ENOMEM;
  const unsigned nb =
      matvec_dots_part<T>(c, numdia, dac, dcc, dxc, dw0, dy, part, 0);
  interp_damped<T>(c, f, dxc, dpc, dxf, part, nb, 0);
  return finish(S, xf, dxf, f.n());
}

}  // namespace drv
}  // namespace pmg

using namespace pmg::drv;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int pmg_mgsz(int mgcoar, int mgdisc, int mgsolv, int nx, int ny, int nz,
             int nlev, int *out) {
  if (!out) return PMG_EBADPARM;
  return mgsz(mgcoar, mgdisc, mgsolv, nx, ny, nz, nlev, out);
}

int pmg_maxlev(int nx, int ny, int nz) { return maxlev(nx, ny, nz); }

int pmg_Vmgsz(int *mgcoar, int *mgdisc, int *mgsolv, int *nx, int *ny,
              int *nz, int *nlev, int *nxc, int *nyc, int *nzc, int *nf,
              int *nc, int *narr, int *narrc, int *n_rpc, int *n_iz,
              int *n_ipc, int *iretot, int *iintot) {
  if (!mgcoar || !mgdisc || !mgsolv || !nx || !ny || !nz || !nlev || !nxc ||
      !nyc || !nzc || !nf || !nc || !narr || !narrc || !n_rpc || !n_iz ||
      !n_ipc || !iretot || !iintot)
    return PMG_EBADPARM;
  int out[12];
  const int rc =
      mgsz(*mgcoar, *mgdisc, *mgsolv, *nx, *ny, *nz, *nlev, out);
  if (rc) return rc;
  int *dst[12] = {nxc, nyc, nzc, nf, nc, narr, narrc, n_rpc, n_iz, n_ipc,
                  iretot, iintot};
  for (int q = 0; q < 12; q++) *dst[q] = out[q];
  return PMG_OK;
}

int pmg_Vpackmg(int *iparm, double *rparm, size_t *nrwk, int *niwk, int *nx,
                int *ny, int *nz, int *nlev, int *nu1, int *nu2, int *mgkey,
                int *itmax, int *istop, int *ipcon, int *nonlin, int *mgsmoo,
                int *mgprol, int *mgcoar, int *mgsolv, int *mgdisc, int *iinfo,
                double *errtol, int *ipkey, double *omegal, double *omegan,
                int *irite, int *iperf) {
  if (!iparm || !rparm || !nrwk || !niwk || !nx || !ny || !nz || !nlev ||
      !nu1 || !nu2 || !mgkey || !itmax || !istop || !ipcon || !nonlin ||
      !mgsmoo || !mgprol || !mgcoar || !mgsolv || !mgdisc || !iinfo ||
      !errtol || !ipkey || !omegal || !omegan || !irite || !iperf)
    return PMG_EBADPARM;
  if (*nrwk > 0x7fffffffULL) return PMG_EBADPARM;  // iparm(1) is an int
  const int v[22] = {(int)*nrwk, *niwk,  *nx,     *ny,     *nz,     *nlev,
                     *nu1,       *nu2,   *mgkey,  *itmax,  *istop,  *iinfo,
                     *irite,     *ipkey, *ipcon,  *nonlin, *mgprol, *mgcoar,
                     *mgdisc,    *mgsmoo, *mgsolv, *iperf};
  for (int q = 0; q < 22; q++) iparm[q] = v[q];
  rparm[0] = *errtol;
  rparm[8] = *omegal;
  rparm[9] = *omegan;
  return PMG_OK;
}

int pmg_mgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
               double *u, const double *xf, const double *yf, const double *zf,
               const double *gxcf, const double *gycf, const double *gzcf,
               double *a1cf, double *a2cf, double *a3cf, double *ccf,
               double *fcf, double *tcf) {
  return driver(iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf, gzcf,
                a1cf, a2cf, a3cf, ccf, fcf, tcf, 0);
}

int pmg_pcgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                double *u, const double *xf, const double *yf,
                const double *zf, const double *gxcf, const double *gycf,
                const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                double *ccf, double *fcf, double *tcf, int num_iters) {
  return driver(iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf, gzcf,
                a1cf, a2cf, a3cf, ccf, fcf, tcf, 2, num_iters);
}

int pmg_ctx_pcg(pmg_ctx *ctx, int mode, int num_iters) {
  if (!ctx || (mode != PMG_FP64 && mode != PMG_MIXED) || ctx->p.mode != 0)
    return PMG_EBADPARM;
  return ctx_pcg(ctx, mode, num_iters);
}

int pmg_newdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                double *u, const double *xf, const double *yf,
                const double *zf, const double *gxcf, const double *gycf,
                const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                double *ccf, double *fcf, double *tcf) {
  return driver(iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf, gzcf,
                a1cf, a2cf, a3cf, ccf, fcf, tcf, 1);
}

int pmg_mypdefinit(int nion, const double *charge, const double *sconc) {
  if (nion < 0 || (nion > 0 && (!charge || !sconc))) return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  set_ions(nion, charge, sconc);
  return hipDeviceSynchronize() == hipSuccess ? PMG_OK : PMG_EDEVICE;
}

int pmg_Vmypdefinitlpbe(int *nion, double *charge, double *sconc) {
  if (!nion) return PMG_EBADPARM;
  return pmg_mypdefinit(*nion, charge, sconc);
}

int pmg_Vmypdefinitnpbe(int *nion, double *charge, double *sconc) {
  if (!nion) return PMG_EBADPARM;
  return pmg_mypdefinit(*nion, charge, sconc);
}

int pmg_last_stats(int *iters, int *ierror, double *relres) {
  if (iters) *iters = t_iters;
  if (ierror) *ierror = t_ierror;
  if (relres) *relres = t_relres;
  return PMG_OK;
}

int pmg_last_history(double *relres, int cap) {
  if (cap < 0 || (cap > 0 && !relres)) return -PMG_EBADPARM;
  const int m = std::min(cap, (int)t_hist.size());
  for (int q = 0; q < m; q++) relres[q] = t_hist[q];
  return (int)t_hist.size();
}

int pmg_ctx_history(pmg_ctx *ctx, double *relres, int cap) {
  if (!ctx || cap < 0 || (cap > 0 && !relres)) return -PMG_EBADPARM;
  const int m = std::min(cap, (int)ctx->hist.size());
  for (int q = 0; q < m; q++) relres[q] = ctx->hist[q];
  return (int)ctx->hist.size();
}

int pmg_ctx_set_tru(pmg_ctx *ctx, const double *tcf) {
  if (!ctx || !tcf) return PMG_EBADPARM;
  return ctx_set_tru(ctx, tcf);
}

int pmg_ctx_create(pmg_ctx **out, const int *iparm, const double *rparm,
                   const double *xf, const double *yf, const double *zf,
                   const double *gxcf, const double *gycf, const double *gzcf,
                   const double *a1cf, const double *a2cf, const double *a3cf,
                   const double *ccf, const double *fcf) {
  if (!out || !iparm || !rparm) return PMG_EBADPARM;
  Params p;
  int rc = decode(iparm, rparm, p);
  if (rc) return rc;
  if (!xf || !yf || !zf || !gxcf || !gycf || !gzcf || !a1cf || !a2cf ||
      !a3cf || !ccf || !fcf)
    return PMG_EBADPARM;
  return ctx_create(out, p, xf, yf, zf, gxcf, gycf, gzcf, a1cf, a2cf, a3cf,
                    ccf, fcf);
}

int pmg_ctx_solve(pmg_ctx *ctx, int mode) {
  if (!ctx || (mode != PMG_FP64 && mode != PMG_MIXED)) return PMG_EBADPARM;
  if (ctx->p.mode == 1) return ctx_newton(ctx, mode);
  return ctx_solve(ctx, mode);
}

int pmg_ctx_solve_group(pmg_ctx *const *ctx, int n, int mode) {
  if (!ctx || n < 1 || n > 64 || (mode != PMG_FP64 && mode != PMG_MIXED))
    return PMG_EBADPARM;
  return ctx_solve_group(ctx, n, mode);
}

int pmg_ctx_get_u(pmg_ctx *ctx, double *u) {
  if (!ctx || !u) return PMG_EBADPARM;
  if (hipMemcpyAsync(u, ctx->u, ctx->L[0].d.n() * sizeof(double),
                     hipMemcpyDeviceToHost, ctx->s) != hipSuccess)
    return PMG_EDEVICE;
  return hipStreamSynchronize(ctx->s) == hipSuccess ? PMG_OK : PMG_EDEVICE;
}

int pmg_ctx_stats(pmg_ctx *ctx, int *iters, int *ierror, double *relres,
                  double *setup_ms, double *solve_ms) {
  if (!ctx) return PMG_EBADPARM;
  if (iters) *iters = ctx->iters;
  if (ierror) *ierror = ctx->ierror;
  if (relres) *relres = ctx->relres;
  if (setup_ms) *setup_ms = ctx->setup_ms;
  if (solve_ms) *solve_ms = ctx->solve_ms;
  return PMG_OK;
}

int pmg_ctx_rebuild(pmg_ctx *ctx) {
  if (!ctx) return PMG_EBADPARM;
  setup_enqueue(ctx);
  return setup_wait(ctx);
}

int pmg_ctx_create_pbe(pmg_ctx **out, const int *iparm, const double *rparm,
                       double h, double ionic_strength, int natoms,
                       const double *atoms) {
  if (!out || !iparm || !rparm) return PMG_EBADPARM;
  *out = nullptr;
  Params p;
  int rc = decode(iparm, rparm, p);
  if (rc) return rc;
  if (p.nx != p.ny || p.nx != p.nz || !(h > 0.0) || !(ionic_strength >= 0.0) ||
      natoms < 0 || (natoms > 0 && !atoms))
    return PMG_EBADPARM;
  return ctx_create_pbe(out, p, h, ionic_strength, natoms, atoms);
}

int pmg_ctx_load_pbe(pmg_ctx *ctx, double h, double ionic_strength,
                     int natoms, const double *atoms) {
  if (!ctx) return PMG_EBADPARM;
  setup_start(ctx);
  int rc = pbe_enqueue(ctx, h, ionic_strength, natoms, atoms);
  if (rc) {
    hipStreamSynchronize(ctx->s);
    return rc;
  }
  setup_enqueue(ctx, true);
  return setup_wait(ctx);
}

size_t pmg_ctx_bytes(const pmg_ctx *ctx) { return ctx ? ctx->bytes : 0; }

void pmg_ctx_destroy(pmg_ctx *ctx) { delete ctx; }

int pmg_ctx_get_level(pmg_ctx *ctx, int lev, double *ac, double *cc,
                      double *fc, double *pc) {
  if (!ctx || lev < 1 || lev > ctx->nlev) return PMG_EBADPARM;
  Level &L = ctx->L[lev - 1];
  size_t n = L.d.n();
  if (ac)
    hipMemcpyAsync(ac, L.ac, (lev == 1 ? 4 : 14) * n * sizeof(double),
                   hipMemcpyDeviceToHost, ctx->s);
  if (cc)
    hipMemcpyAsync(cc, L.cc, n * sizeof(double), hipMemcpyDeviceToHost,
                   ctx->s);
  if (fc)
    hipMemcpyAsync(fc, L.fc, n * sizeof(double), hipMemcpyDeviceToHost,
                   ctx->s);
  if (pc && lev < ctx->nlev) {
    Level &C = ctx->L[lev];
    hipMemcpyAsync(pc, C.pc, 27 * C.d.n() * sizeof(double),
                   hipMemcpyDeviceToHost, ctx->s);
  }
  return hipStreamSynchronize(ctx->s) == hipSuccess ? PMG_OK : PMG_EDEVICE;
}

int pmg_matvec(int nx, int ny, int nz, int numdia, const double *ac,
               const double *cc, const double *x, double *y) {
  if (!dims_ok(nx, ny, nz) || (numdia != 7 && numdia != 27) || !ac || !cc ||
      !x || !y)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim d{nx, ny, nz};
  size_t n = d.n();
  Scratch S;
  double *dac = S.put(ac, (numdia == 7 ? 4 : 14) * n), *dcc = S.put(cc, n),
         *dx = S.put(x, n), *dy = S.put(y, n);
  if (!S.ok) return PMG_ENOMEM;
  matvec_dots<double>(d, numdia, dac, dcc, dx, nullptr, dy, nullptr, 0);
  return finish(S, y, dy, n);
}

int pmg_mresid(int nx, int ny, int nz, int numdia, const double *ac,
               const double *cc, const double *fc, const double *x,
               double *r) {
  if (!dims_ok(nx, ny, nz) || (numdia != 7 && numdia != 27) || !ac || !cc ||
      !fc || !x || !r)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim d{nx, ny, nz};
  size_t n = d.n();
  Scratch S;
  double *dac = S.put(ac, (numdia == 7 ? 4 : 14) * n), *dcc = S.put(cc, n),
         *dfc = S.put(fc, n), *dx = S.put(x, n), *dr = S.put(r, n);
  if (!S.ok) return PMG_ENOMEM;
  resid<double, double>(d, numdia, dac, dcc, dfc, dx, dr, nullptr, 0);
  return finish(S, r, dr, n);
}

int pmg_gsrb(int nx, int ny, int nz, int numdia, const double *ac,
             const double *cc, const double *fc, double *x, int itmax,
             int iadjoint) {
  if (!dims_ok(nx, ny, nz) || (numdia != 7 && numdia != 27) || !ac || !cc ||
      !fc || !x || itmax < 0 || (iadjoint != 0 && iadjoint != 1))
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim d{nx, ny, nz};
  size_t n = d.n();
  Scratch S;
  double *dac = S.put(ac, (numdia == 7 ? 4 : 14) * n), *dcc = S.put(cc, n),
         *dfc = S.put(fc, n), *dx = S.put(x, n);
  if (!S.ok) return PMG_ENOMEM;
  if (numdia == 7) {
    double *tmp = S.put(x, n);  // same boundary values as x
    if (!S.ok) return PMG_ENOMEM;
    gs7_fused<double>(d, dac, dcc, dfc, dx, tmp, itmax, iadjoint, 0);
  } else {
    // x's boundary values: the second buffer of the two-launch sweeps
    double *tmp = S.put(x, n);
    if (!S.ok) return PMG_ENOMEM;
    gs27<double>(d, dac, dcc, dfc, dx, itmax, iadjoint, 0, false, tmp);
  }
  return finish(S, x, dx, n);
}

int pmg_restrc(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
               const double *xin, double *xout, const double *pc) {
  if (!dims_ok(nxc, nyc, nzc) || !coarse_of(nxf, nxc) ||
      !coarse_of(nyf, nyc) || !coarse_of(nzf, nzc) || !xin || !xout || !pc)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim f{nxf, nyf, nzf}, c{nxc, nyc, nzc};
  Scratch S;
  double *dxi = S.put(xin, f.n()), *dxo = S.put(xout, c.n()),
         *dpc = S.put(pc, 27 * c.n());
  if (!S.ok) return PMG_ENOMEM;
  restrict_<double>(f, c, dxi, dxo, dpc, 0);
  return finish(S, xout, dxo, c.n());
}

int pmg_interp(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
               const double *xin, double *xout, const double *pc) {
  if (!dims_ok(nxc, nyc, nzc) || !coarse_of(nxf, nxc) ||
      !coarse_of(nyf, nyc) || !coarse_of(nzf, nzc) || !xin || !xout || !pc)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim f{nxf, nyf, nzf}, c{nxc, nyc, nzc};
  Scratch S;
  double *dxi = S.put(xin, c.n()), *dxo = S.put(xout, f.n()),
         *dpc = S.put(pc, 27 * c.n());
  if (!S.ok) return PMG_ENOMEM;
  interp<double>(c, f, dxi, dpc, dxo, nullptr, 0);
  return finish(S, xout, dxo, f.n());
}

int pmg_correct(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                int numdia, const double *acc, const double *ccc,
                const double *xc, const double *w0c, const double *pc,
                double *xf) {
  return correct_t<double>(nxc, nyc, nzc, nxf, nyf, nzf, numdia, acc, ccc, xc,
                           w0c, pc, xf);
}

int pmg_mresid_f32(int nx, int ny, int nz, int numdia, const float *ac,
                   const float *cc, const float *fc, const float *x,
                   float *r) {
  return mresid_t<float>(nx, ny, nz, numdia, ac, cc, fc, x, r);
}

int pmg_gsrb_f32(int nx, int ny, int nz, int numdia, const float *ac,
                 const float *cc, const float *fc, float *x, int itmax,
                 int iadjoint, int zero_start) {
  return gsrb_t<float>(nx, ny, nz, numdia, ac, cc, fc, x, itmax, iadjoint,
                       zero_start);
}

int pmg_gsrb_zero(int nx, int ny, int nz, int numdia, const double *ac,
                  const double *cc, const double *fc, double *x, int itmax,
                  int iadjoint) {
  return gsrb_t<double>(nx, ny, nz, numdia, ac, cc, fc, x, itmax, iadjoint,
                        1);
}

int pmg_restrc_f32(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                   const float *xin, float *xout, const float *pc) {
  return restrc_t<float>(nxf, nyf, nzf, nxc, nyc, nzc, xin, xout, pc);
}

int pmg_interp_f32(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                   const float *xin, float *xout, const float *pc) {
  return interp_t<float>(nxc, nyc, nzc, nxf, nyf, nzf, xin, xout, pc);
}

int pmg_correct_f32(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                    int numdia, const float *acc, const float *ccc,
                    const float *xc, const float *w0c, const float *pc,
                    float *xf) {
  return correct_t<float>(nxc, nyc, nzc, nxf, nyf, nzf, numdia, acc, ccc, xc,
                          w0c, pc, xf);
}

int pmg_mresid_mixed(int nx, int ny, int nz, const double *ac,
                     const double *cc, const double *fc, const double *x,
                     float *r, double *l1) {
  if (!dims_ok(nx, ny, nz) || !ac || !cc || !fc || !x || !r)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim d{nx, ny, nz};
  size_t n = d.n();
  Scratch S;
  double *dac = S.put(ac, 4 * n), *dcc = S.put(cc, n), *dfc = S.put(fc, n),
         *dx = S.put(x, n);
  float *dr = S.put(r, n);
  double *part = S.put((const double *)nullptr, max_partials(d) + 8);
  if (!S.ok) return PMG_ENOMEM;
  Red red;
  red.part = part;
  red.out = part + max_partials(d);
  resid<double, float>(d, 7, dac, dcc, dfc, dx, dr, &red, 0);
  int rc = finish(S, r, dr, n);
  if (!rc && l1 &&
      hipMemcpy(l1, red.out, sizeof(double), hipMemcpyDeviceToHost) !=
          hipSuccess)
    rc = PMG_EDEVICE;
  return rc;
}

int pmg_mresid_mixed_noalloc(int nx, int ny, int nz, const double *ac,
                             const double *cc, const double *fc,
                             const double *x, float *r, double *l1,
                             void *stream) {
  if (!dims_ok(nx, ny, nz) || !ac || !cc || !fc || !x || !r || !l1)
    return PMG_EBADPARM;
  const Dim d{nx, ny, nz};
  if (resid_partials(d) > kScratchDoubles) return PMG_ESCRATCH;
  if (!device_ok()) return PMG_EDEVICE;
  double *part = scratch_ptr();
  if (!part) return PMG_EDEVICE;
  // every partial the final sum reads is written by this call's residual
  // kernel, so nothing left by an earlier call needs clearing
  Red red;
  red.part = part;
  red.out = l1;
  resid<double, float>(d, 7, ac, cc, fc, x, r, &red, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? PMG_OK : PMG_EDEVICE;
}

int pmg_addcorr_f32(int nx, int ny, int nz, const float *e, double *u) {
  if (!dims_ok(nx, ny, nz) || !e || !u) return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim d{nx, ny, nz};
  size_t n = d.n();
  Scratch S;
  float *de = S.put(e, n);
  double *du = S.put(u, n);
  if (!S.ok) return PMG_ENOMEM;
  add_correction<float>(d, de, du, 0);
  return finish(S, u, du, n);
}

int pmg_buildP(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
               int mgprol, int numdia, const double *acf, double *pc) {
  if (!dims_ok(nxc, nyc, nzc) || !coarse_of(nxf, nxc) ||
      !coarse_of(nyf, nyc) || !coarse_of(nzf, nzc) ||
      (numdia != 7 && numdia != 27) || (mgprol != 0 && mgprol != 1) || !acf ||
      !pc)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim f{nxf, nyf, nzf}, c{nxc, nyc, nzc};
  Scratch S;
  double *dac = S.put(acf, (numdia == 7 ? 4 : 14) * f.n()),
         *dpc = S.put(pc, 27 * c.n());
  if (!S.ok) return PMG_ENOMEM;
  buildP(f, c, mgprol, numdia, dac, dpc, 0);
  return finish(S, pc, dpc, 27 * c.n());
}

int pmg_buildG(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
               int numdia, const double *pc, const double *acf, double *acc) {
  if (!dims_ok(nxc, nyc, nzc) || !coarse_of(nxf, nxc) ||
      !coarse_of(nyf, nyc) || !coarse_of(nzf, nzc) ||
      (numdia != 7 && numdia != 27) || !acf || !pc || !acc)
    return PMG_EBADPARM;
  if (!device_ok()) return PMG_EDEVICE;
  Dim f{nxf, nyf, nzf}, c{nxc, nyc, nzc};
  Scratch S;
  double *dpc = S.put(pc, 27 * c.n()),
         *dac = S.put(acf, (numdia == 7 ? 4 : 14) * f.n()),
         *dacc = S.put(acc, 14 * c.n());
  if (!S.ok) return PMG_ENOMEM;
  buildG(f, c, numdia, dpc, dac, dacc, 0);
  return finish(S, acc, dacc, 14 * c.n());
}

const char *pmg_version(void) { return "pmg_mi355x 0.1 (gfx950)"; }

// Kernel timing hooks used by the benchmark: when enabled, the fine-level
// smoothing launches of each solve are bracketed by HIP events on the
// context stream; pmg_ctx_ktime returns their accumulated device time and
// the number of full sweeps timed.
int pmg_ctx_ktime_enable(pmg_ctx *ctx, int max_events) {
  if (!ctx || max_events < 2) return PMG_EBADPARM;
  for (auto e : ctx->ev) hipEventDestroy(e);
  ctx->ev.assign((size_t)max_events, nullptr);
  for (auto &e : ctx->ev) hipEventCreate(&e);
  ctx->ktime = true;
  ctx->ktime_ms = 0.0;
  ctx->ktime_count = 0;
  return PMG_OK;
}

int pmg_ctx_set_graphs(pmg_ctx *ctx, int on) {
  if (!ctx) return PMG_EBADPARM;
  ctx->graphs = on != 0;
  return PMG_OK;
}

int pmg_ctx_ktime(pmg_ctx *ctx, double *ms, long long *sweeps) {
  if (!ctx) return PMG_EBADPARM;
  if (ms) *ms = ctx->ktime_ms;
  if (sweeps) *sweeps = ctx->ktime_count;
  return PMG_OK;
}

}  // extern "C"
