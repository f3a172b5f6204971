// pmg_ctx.h -- resident solve context and its session interface (internal).
//
// A pmg_ctx holds one PMG hierarchy in HBM (the reference's Vbuildstr /
// Vbuildops layout per level, pmgc/mgsubd.cc:63-347) with its own HIP
// stream.  A linear solve runs as a session: lin_begin enqueues the start
// (Vazeros, VfboundPMG00 and the stopping-test denominator,
// pmgc/mgdrvd.cc:419-439, mgcsd.cc:133-163), lin_enqueue one iteration of
// Vmvcs's loop (mgcsd.cc:299-527, replayed from a captured HIP graph),
// lin_poll reads the stopping norm back without blocking, and lin_finish
// restores the Dirichlet values (mgdrvd.cc:500-502).  The blocking
// pmg_ctx_solve and the serving loop both drive contexts through it.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pmg_mi355x.h"
#include "pmg_internal.h"

namespace pmg {
namespace drv {

constexpr int MAXLEV = 16;

// device scalar slots of a context (pmg_ctx::scal, mirrored in hscal):
// stopping-test denominator and norm, xdamp, the band factor's Vdpbfa info
// (mgsolv 1), a temporary
enum { S_RSDEN = 0, S_RSNRM = 1, S_XDAMP = 2, S_BINFO = 3, S_TMP = 4 };

struct Params {
  int nrwk, niwk, nx, ny, nz, nlev, nu1, nu2, mgkey, itmax, istop, iinfo,
      ipkey, mode, mgprol, mgcoar, mgdisc, mgsmoo, mgsolv, iperf;
  double errtol;
};

struct Level {
  Dim d;
  int numdia = 7;
  double *ac = nullptr, *cc = nullptr, *fc = nullptr, *pc = nullptr;
  double *x = nullptr, *w0 = nullptr, *r = nullptr, *t = nullptr;
  float *ac32 = nullptr, *cc32 = nullptr, *pc32 = nullptr, *x32 = nullptr,
        *w032 = nullptr, *r32 = nullptr, *t32 = nullptr;
  // Newton: Jacobian diagonal term and right-hand side per level, and the
  // Newton iterate of a coarse level (nested iteration)
  double *cpr = nullptr, *rhs = nullptr, *xn = nullptr;
  float *cpr32 = nullptr;
  // Newton, fine level: the diagonal oC + cpr of the Jacobian (FP64, and the
  // FP32 sum of the FP32 copies for mixed inner solves)
  double *dgn = nullptr;
  float *dgn32 = nullptr;
  // 7-point (fine) level: the diagonal oC + cc, precomputed for the sweeps
  // and residuals (FP64; FP32 as the sum of the FP32 copies)
  double *dg = nullptr;
  float *dg32 = nullptr;
};

template <class T>
struct LV {
  Dim d;
  int numdia;
  const T *ac, *cc, *pc;
  T *x, *w0, *r, *t;
  const T *dg;  // oC + cc when it holds for ac and cc above, else null
};


// decode iparm/rparm (mgdrvd.cc:111-132, 247-265); PMG_EBADPARM when this
// build does not provide what they ask for
int decode(const int *iparm, const double *rparm, Params &p);

}  // namespace drv
}  // namespace pmg

struct pmg_ctx {
  pmg::drv::Params p{};
  int nlev = 0;
  pmg::drv::Level L[pmg::drv::MAXLEV];
  hipStream_t s = nullptr;
  std::vector<void *> allocs;
  size_t bytes = 0;
  bool ok = true;
  // resident fine-level inputs
  double *xf = nullptr, *yf = nullptr, *zf = nullptr, *gx = nullptr,
         *gy = nullptr, *gz = nullptr, *a1 = nullptr, *a2 = nullptr,
         *a3 = nullptr, *ccf = nullptr, *fcf = nullptr;
  double *u = nullptr;       // fine solution (FP64)
  bool u_given = false;      // u holds the next Newton solve's start
  // istop 2-5: the comparison solution tru (tcf, fine level; istop 2 then
  // holds the previous iterate) and the relative stopping measure after
  // each iteration of the last solve (Vprtstp's relres, mgsubd.cc:486-536)
  double *tru = nullptr;
  std::vector<double> hist;
  double *work[4] = {};      // fine FP64 scratch (Newton)
  double *pcg[4] = {};       // class-style PCG: r, p, A p, s (FP64)
  double *part = nullptr;    // reduction partials
  // mgsolv = 1: factored band of the coarsest operator (reference layout)
  // and the factor by rows and by columns for the substitution
  double *band = nullptr, *band_rt = nullptr, *band_ct = nullptr;
  double *scal = nullptr;    // device scalars
  double *hscal = nullptr;   // pinned host mirror
  bool have32 = false, haveN = false;
  bool ops32_ok = false;  // FP32 operator copies match the FP64 hierarchy
  int iters = 0, ierror = 0;
  double relres = 0.0, setup_ms = 0.0, solve_ms = 0.0;
  // linear-solve session (pmg_ctx_solve and the serving loop)
  struct Session {
    int mode = PMG_FP64, iters = 0, rc = PMG_OK;
    int phase = 0;  // 0: start in flight, 1: iterating, 2: finishing
    bool done = true, pending = false;
    double rsden = 1.0, rsnrm = 0.0;
  } ses;
  hipEvent_t cyc_ev = nullptr, t0_ev = nullptr, t1_ev = nullptr;
  hipEvent_t fin_ev = nullptr;
  double *u_pin = nullptr;  // pinned host copy of the solution (serving)
  hipEvent_t su0_ev = nullptr, su1_ev = nullptr, stage_ev = nullptr;
  // atom list of the GPU input assembly: device copy (+ links) and pinned
  // staging, capacity in atoms
  double *atoms_d = nullptr, *atoms_h = nullptr;
  int atoms_cap = 0;
  // V-cycle bodies captured as HIP graphs, one per precision mode
  bool graphs = true, capturing = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  // optional per-kernel event timing of the fine smoothing sweeps
  bool ktime = false;
  std::vector<hipEvent_t> ev;
  int nev = 0;
  double ktime_ms = 0.0;
  long long ktime_count = 0;

  template <class T>
  T *alloc(size_t n) {
    void *ptr = nullptr;
    if (!ok) return nullptr;
    if (hipMalloc(&ptr, n * sizeof(T) + 256) != hipSuccess) {
      ok = false;
      return nullptr;
    }
    hipMemsetAsync(ptr, 0, n * sizeof(T) + 256, s);
    allocs.push_back(ptr);
    bytes += n * sizeof(T) + 256;
    return (T *)ptr;
  }
  ~pmg_ctx() {
    if (s) hipStreamSynchronize(s);
    for (int m = 0; m < 2; m++) {
      if (gexec[m]) hipGraphExecDestroy(gexec[m]);
      if (graph[m]) hipGraphDestroy(graph[m]);
    }
    if (cyc_ev) hipEventDestroy(cyc_ev);
    if (t0_ev) hipEventDestroy(t0_ev);
    if (t1_ev) hipEventDestroy(t1_ev);
    if (fin_ev) hipEventDestroy(fin_ev);
    if (u_pin) hipHostFree(u_pin);
    for (hipEvent_t e : {su0_ev, su1_ev, stage_ev})
      if (e) hipEventDestroy(e);
    if (atoms_d) hipFree(atoms_d);
    if (atoms_h) hipHostFree(atoms_h);
    for (void *ptr : allocs) hipFree(ptr);
    if (hscal) hipHostFree(hscal);
    for (auto e : ev) hipEventDestroy(e);
    if (s) hipStreamDestroy(s);
  }
};


namespace pmg {
namespace drv {

// Allocate a context for p (stream, hierarchy, zeroed resident inputs).
int ctx_new(pmg_ctx **out, const Params &p);
// Bytes a context for p needs in HBM (mode PMG_MIXED adds the FP32 copies).
size_t ctx_bytes_estimate(const Params &p, int mode);
// Stage an atom list and enqueue the GPU assembly of the fine-level inputs.
int pbe_enqueue(pmg_ctx *c, double h, double ionic, int natoms,
                const double *atoms);
// Setup timing events and the hierarchy build (VbuildA/P/G, FP32 copies).
void setup_start(pmg_ctx *c);
void setup_enqueue(pmg_ctx *c, bool started = false);
int setup_wait(pmg_ctx *c);
// Session of a linear solve (see the file comment).  lin_begin never
// blocks; lin_poll returns 1 once the session has finished.
int lin_begin(pmg_ctx *c, int mode);
void lin_enqueue(pmg_ctx *c);
int lin_poll(pmg_ctx *c, bool wait);
// Enqueue the end of the session (Dirichlet values; the solution copy to
// host_u when non-null, which must stay valid until lin_end) and record
// the completion event.
void lin_finish(pmg_ctx *c, double *host_u);
// 1 when the work of lin_finish has completed (wait: block until then).
int lin_finished(pmg_ctx *c, bool wait);
// Statistics of the finished session; returns its status code.
int lin_end(pmg_ctx *c);
// An iteration of the session has landed with its stopping scalar in
// c->hscal[S_RSNRM]: count, record, decide.  1: finished.
int ses_cycle_done(pmg_ctx *c);

// Grouped iterations (the serving loop): contexts whose sessions may run
// their iterations through group_step, and one iteration of n such
// sessions (same mode, same nu1 and nu2) on stream s, each kernel launched
// once for all of them; context q's stopping scalar also goes to norm[q]
// (device memory).  Bit-identical to each context's own iteration.
bool group_ok(const pmg_ctx *c);
void group_step(pmg_ctx *const *c, int n, int mode, hipStream_t s,
                double *norm);

}  // namespace drv
}  // namespace pmg
