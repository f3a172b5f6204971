// pmg_internal.h -- kernel launchers shared by the PMG driver and the C ABI.
//
// Layout on the device is the reference's: 1-based Fortran order, i fastest,
// operators as planes of n = nx*ny*nz (pmgc/vmatrix.h:79-80,
// pmgc/matvecd.cc:185-189), P as 27 coarse-sized planes
// (pmgc/matvecd.cc:809-816).  Every launcher checks the shapes it is given
// against the grid it launches before launching.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pmg {

struct Dim {
  int nx = 0, ny = 0, nz = 0;
  size_t n() const { return (size_t)nx * ny * nz; }
  size_t interior() const {
    return (size_t)(nx - 2) * (ny - 2) * (nz - 2);
  }
};

inline int corsr(int n) { return (n - 1) / 2 + 1; }

// Largest number of per-block partial sums any reduction launcher writes for
// a level of dims d (the caller sizes the partials buffer with it).
size_t max_partials(const Dim &d);

// Destination of a reduction computed alongside the kernel that produces
// its terms: per-block partials (2 per block at most) and the device scalar
// a final single-workgroup pass writes after summing them in block order.
struct Red {
  double *part = nullptr;
  double *out = nullptr;
};

// ---- smoothers (pmgc/gsd.cc) ----
template <class T>
void gs7(const Dim &d, const T *ac, const T *cc, const T *fc, T *x,
         int nsweeps, int iadjoint, hipStream_t s);
// nsweeps full red-black sweeps, each one pass over memory (x and tmp
// ping-pong; tmp must share x's boundary values).  Same result as gs7.
// zero_init: the sweeps start from x = 0 (Vazeros); x's interior is not
// read, its boundary entries must already be zero.
// dg (may be null): the diagonal oC + cc of the operator, precomputed
// (diag_sum); the sweep then reads it instead of oC and cc.
// uadd (FP32 sweeps with dg, not zero_init): the last sweep adds its result
// to this FP64 iterate (u += (double)x at the interior points) instead of
// storing it in x; returns true when it did.
template <class T>
bool gs7_fused(const Dim &d, const T *ac, const T *cc, const T *fc, T *x,
               T *tmp, int nsweeps, int iadjoint, hipStream_t s,
               bool zero_init = false, const T *dg = nullptr,
               double *uadd = nullptr);
// dg[p] = ac[p] + cc[p] for the n points of a 7-point level (ac: its oC
// plane), the diagonal the 7-point sweeps and residuals use.
template <class T>
void diag_sum(const T *ac, const T *cc, T *dg, size_t n, hipStream_t s);
// zero_init: x starts from 0 (the caller need not clear it).  tmp (may be
// null): an array of the level's size with the same boundary values as x,
// the second buffer of the two-launch sweeps; its interior is overwritten.
template <class T>
void gs27(const Dim &d, const T *ac, const T *cc, const T *fc, T *x,
          int nsweeps, int iadjoint, hipStream_t s, bool zero_init = false,
          T *tmp = nullptr);

// ---- residual r = f - A x (pmgc/matvecd.cc:381-528) ----
// TO is the output type (FP64 operator may write an FP32 residual for the
// mixed-precision inner solve); r may be null.  If red != nullptr the L1
// norm of the residual (in double) is written to red->out[0].
// dg (7-point, may be null): the precomputed diagonal oC + cc.
template <class T, class TO>
void resid(const Dim &d, int numdia, const T *ac, const T *cc, const T *fc,
           const T *x, TO *r, const Red *red, hipStream_t s,
           const T *dg = nullptr);
// Number of block partials resid() writes (all of them, every call) when
// given a Red.
size_t resid_partials(const Dim &d);
// Nonlinear residual r = f - A x - c_vec(cc, x) (pmgc/matvecd.cc:530-693),
// 7- or 27-point operator, FP64.
void nmresid(const Dim &d, int numdia, const double *ac, const double *cc,
             const double *fc, const double *x, double *r, const Red *red,
             hipStream_t s);

// ---- y = A x; with red, red->out[0] = (x.w0) / (x.y), the Hackbusch-
// Reusken damping of the coarse correction (pmgc/mgcsd.cc:429-452) ----
template <class T>
void matvec_dots(const Dim &d, int numdia, const T *ac, const T *cc,
                 const T *x, const T *w0, T *y, const Red *red, hipStream_t s);

// y = A x and the per-block partials of (x.w0, x.y) into part (2 per
// block), for interp_damped; returns the number of blocks.
template <class T>
unsigned matvec_dots_part(const Dim &d, int numdia, const T *ac, const T *cc,
                          const T *x, const T *w0, T *y, double *part,
                          hipStream_t s);

// ---- grid transfer (pmgc/matvecd.cc:695-983) ----
template <class T>
// out32, when non-null, also receives the coarse interior rounded to FP32.
void restrict_(const Dim &f, const Dim &c, const T *xin, T *xout, const T *pc,
               hipStream_t s, float *out32 = nullptr);
// xf(interior) += (*xdamp) * (P xc); xdamp == nullptr means plain
// interpolation into xf with the fine boundary set to zero.
template <class T>
void interp(const Dim &c, const Dim &f, const T *xc, const T *pc, T *xf,
            const double *xdamp, hipStream_t s);

// xf(interior) += xdamp * (P xc) with xdamp = sum(part[2b]) / sum(part[2b+1])
// over the nblk partials of matvec_dots_part (0 when the denominator is 0),
// summed by every workgroup in one fixed order.
template <class T>
void interp_damped(const Dim &c, const Dim &f, const T *xc, const T *pc,
                   T *xf, const double *part, unsigned nblk, hipStream_t s);

// ---- stopping tests istop 2-5 (pmgc/mgcsd.cc:126-152, 497-526) ----
// d = tru - u on the interior.  kind 2: red.out[0] = sum |d|, then tru = u;
// kind 3: red.out[0] = sum d^2; kind 5: w = d (no reduction).
void stop_diff(const Dim &d, int kind, double *tru, const double *u,
               double *w, const Red &red, hipStream_t s);
// red.out[0] = w . (A w) over the interior (y receives A w).
template <class T>
void dot_energy(const Dim &d, int numdia, const T *ac, const T *cc,
                const T *w, T *y, const Red &red, hipStream_t s);
// Vextrac: coarse interior values by injection from the fine level.
void extrac(const Dim &f, const Dim &c, const double *xf, double *xc,
            hipStream_t s);

// ---- reductions ----
// Deterministic interior L1 norm into red.out[0].
template <class T>
void nrm1(const Dim &d, const T *x, const Red &red, hipStream_t s);

// ---- coarsest-level CG (pmgc/cgd.cc:60-124), one workgroup, FP64 ----
// Returns false if the level is too large for the single-workgroup kernel.
bool cg_fits(const Dim &d);
template <class T>
void cg_single(const Dim &d, int numdia, const double *ac, const double *cc,
               const T *f, T *x, int itmax, double errtol, hipStream_t s);

// ---- the class-style solver (PBE_pmgc_class.h:1070-1133,
// CG_Solver.h:45-242) ----
// Coarsest CG of the class V-cycle: starts from x as it is and stops when
// sqrt(r.r) <= thresh (first test on sqrt(sqrt(r0.r0))) or after itmax
// iterations; 27-point levels that fit one workgroup's row-segment kernel.
bool cg_class_fits(const Dim &d, int numdia);
template <class T>
void cg_class(const Dim &d, int numdia, const double *ac, const double *cc,
              const T *f, T *x, int itmax, double thresh, hipStream_t s);
// red.out[0] = x.y over the interior (deterministic order)
void dot_interior(const Dim &d, const double *x, const double *y,
                  const Red &red, hipStream_t s);
// x += (rs/pap) p, r -= (rs/pap) q; p = s + (rsn/rs) p (interior; device
// scalars, a zero denominator gives a zero coefficient)
void pcg_xr(const Dim &d, double *x, double *r, const double *p,
            const double *q, const double *rs, const double *pap,
            hipStream_t s);
void pcg_p(const Dim &d, double *p, const double *sv, const double *rsn,
           const double *rs, hipStream_t s);

// ---- banded coarsest solve, mgsolv = 1 (pmgc/buildBd.cc:62-304,
// pmgc/mlinpckd.cc:167-235), FP64 ----
// n unknowns (interior points), half-bandwidth m, lda = m + 1, W = row
// stride of the row-wise factor (a multiple of 64 above m); ok is false
// for levels the band path does not take (one interior point in x or y
// with 27 points, or a factor window m + 1 beyond the LDS of one CU).
struct BandDims {
  int n = 0, m = 0, lda = 0, W = 0;
  int npad = 0, nrow = 0;  // n rounded up to a multiple of W; rows of rt/ct
  bool ok = false;
};
BandDims band_dims(const Dim &d, int numdia);
// Band of the operator ac of level d (abd: n * lda), its factor in place,
// and the factor by rows and by columns for the substitution (rt, ct:
// nrow * W each); info[0] = 0, or the first non-positive pivot.
void band_build_factor(const Dim &d, int numdia, const double *ac,
                       double *abd, double *rt, double *ct, double *info,
                       hipStream_t s);
// x = A_band^{-1} f on the interior, 0 on the boundary.
template <class T>
void band_solve(const Dim &d, int numdia, const double *rt, const double *ct,
                const T *f, T *x, hipStream_t s);

// ---- BLAS-1 helpers ----
template <class T>
void fill(T *x, size_t n, T v, hipStream_t s);
template <class T>
void axpy_interior(const Dim &d, double alpha, const T *x, T *y,
                   hipStream_t s);
template <class TA, class TB>
void convert(const TA *a, TB *b, size_t n, hipStream_t s);
// u(interior) += e(interior) (u FP64, e any type)
template <class T>
void add_correction(const Dim &d, const T *e, double *u, hipStream_t s);
void copy_interior(const Dim &d, const double *x, double *y, hipStream_t s);
void fbound00(const Dim &d, double *x, hipStream_t s);
void fbound_dirichlet(const Dim &d, double *x, const double *gx,
                      const double *gy, const double *gz, hipStream_t s);

// ---- ion terms (pmgc/mypdec.cc) ----
void set_ions(int nion, const double *charge, const double *sconc);
// out32: FP32 copy of out; oc, dg: dg = oc + out (and dg32 = oc32 + out32
// in FP32), the diagonal of the Jacobian's 7-point operator.
void dc_vec(const double *coef, const double *u, double *out, size_t n,
            hipStream_t s, float *out32 = nullptr, const double *oc = nullptr,
            double *dg = nullptr, const float *oc32 = nullptr,
            float *dg32 = nullptr);
// x_out = x + damp * dx on the interior (boundary of x_out = boundary of x)
void xpay_interior(const Dim &d, const double *x, double damp,
                   const double *dx, double *out, hipStream_t s);

// ---- setup (pmgc/buildAd.cc, buildPd.cc, buildGd.cc), FP64 ----
// A non-null *32 output also receives the interior values rounded to FP32
// (the boundary entries are left as they are: zero from allocation).
void buildA(const Dim &d, double *ac, double *cc, double *fc,
            const double *xf, const double *yf, const double *zf,
            const double *gx, const double *gy, const double *gz,
            const double *a1, const double *a2, const double *a3,
            const double *ccf, const double *fcf, hipStream_t s,
            double *dg, float *ac32 = nullptr, float *cc32 = nullptr,
            float *dg32 = nullptr);
// The two restrictions of a level's setup (cc and fc, Vbuildgaler0
// mgsubd.cc:370-372) in one pass over P, each sum in restrict_'s term order.
void restrict2(const Dim &f, const Dim &c, const double *a, const double *b,
               double *ao, double *bo, const double *pc, hipStream_t s,
               float *ao32, float *bo32 = nullptr);
void buildP(const Dim &f, const Dim &c, int mgprol, int numdia,
            const double *acf, double *pc, hipStream_t s,
            float *pc32 = nullptr);
void buildG(const Dim &f, const Dim &c, int numdia, const double *pc,
            const double *acf, double *acc, hipStream_t s,
            float *acc32 = nullptr);

// ---- Poisson-Boltzmann input assembly from atoms (PBE_pmgc_class.h:191-344)
// atoms: device array natoms x {x, y, z, radius, charge}; next: natoms ints;
// scratch: pbe_scratch_bytes(n).  Writes the fine-level inputs of an n^3
// grid of spacing h centred on the origin; gx/gy/gz get their two planes
// (low, high).  Returns 0, or 2 for invalid arguments.
// eps_in / eps_out: dielectric inside / outside the atom spheres; shell:
// ion-exclusion shell width (A) beyond each radius.
void pbe_constants(double ionic, double eps_out, double *lambda_b,
                   double *kappa2);
size_t pbe_scratch_bytes(int n);
int pbe_assemble(int n, double h, double ionic, int natoms,
                 const double *atoms, int *next, void *scratch, double *xf,
                 double *yf, double *zf, double *gx, double *gy, double *gz,
                 double *a1, double *a2, double *a3, double *ccf, double *fcf,
                 hipStream_t s, double eps_in = 2.0, double eps_out = 78.54,
                 double shell = 2.0);

// ---- grouped launches (the serving loop): the same step for up to GMAX
// independent problems per launch, each problem running exactly the
// workgroups and arithmetic of its own launch (bit-identical results).
// Problems a grouped kernel does not take run through the single launcher.
constexpr int GMAX = 8;
template <class T>
struct GSmooth {  // one level's smoothing (gs7_fused / gs27 arguments)
  Dim d;
  const T *ac, *cc, *f;
  T *x, *tmp;
  const T *dg = nullptr;
};
template <class T>
void g_gs7_fused(const GSmooth<T> *j, int n, int nsweeps, int iadjoint,
                 bool zero_init, hipStream_t s);
template <class T>
void g_gs27(const GSmooth<T> *j, int n, int nsweeps, int iadjoint,
            bool zero_init, hipStream_t s);
template <class T, class TO>
struct GResid {  // resid(); out2, when non-null, also receives the norm
  Dim d;
  int numdia;
  const T *ac, *cc, *fc, *x;
  TO *r;
  const Red *red;
  double *out2;
  const T *dg = nullptr;
};
template <class T, class TO>
void g_resid(const GResid<T, TO> *j, int n, hipStream_t s);
template <class T>
struct GRestrict {
  Dim f, c;
  const T *xin;
  T *xout;
  const T *pc;
};
template <class T>
void g_restrict(const GRestrict<T> *j, int n, hipStream_t s);
template <class T>
struct GMatvec {  // matvec_dots_part()
  Dim d;
  int numdia;
  const T *ac, *cc, *x, *w0;
  T *y;
  double *part;
};
// nb[q] receives problem q's number of partials.
template <class T>
void g_matvec_dots_part(const GMatvec<T> *j, int n, unsigned *nb,
                        hipStream_t s);
template <class T>
struct GInterp {  // interp_damped()
  Dim c, f;
  const T *xc, *pc;
  T *xf;
  const double *part;
  unsigned nblk;
};
template <class T>
void g_interp_damped(const GInterp<T> *j, int n, hipStream_t s);
template <class T>
struct GCoarse {  // cg_single()
  Dim d;
  int numdia;
  const double *ac, *cc;
  const T *f;
  T *x;
};
template <class T>
void g_cg(const GCoarse<T> *j, int n, int itmax, double errtol,
          hipStream_t s);
template <class T>
struct GAddcorr {  // add_correction()
  Dim d;
  const T *e;
  double *u;
};
template <class T>
void g_addcorr(const GAddcorr<T> *j, int n, hipStream_t s);

}  // namespace pmg
