/*
 * pmg_mi355x.h -- C ABI of the MI355X-native PMG multigrid solver.
 *
 * Drop-in boundary for the reference's PMG driver entry points.  All arrays
 * use the reference's layouts: 1-based, Fortran order (i fastest),
 * VAT3(a,i,j,k) = a[(k-1)*nx*ny + (j-1)*nx + (i-1)] (pmgc/vmatrix.h:79-80);
 * iparm/rparm are the APBS control arrays (pmgc/mgsubd.cc:538-576);
 * iwork/rwork are caller-allocated with the sizes pmg_mgsz returns.
 *
 * Return codes (the reference prints and exits instead):
 *   PMG_OK 0, PMG_ITMAX 1 (itmax reached), PMG_EBADPARM 2 (unsupported or
 *   invalid iparm/rparm), PMG_ENOTSPD 3 (coarse operator breakdown),
 *   PMG_ENOMEM 4 (device allocation failed), PMG_EDEVICE 5 (no usable GPU
 *   or a HIP runtime error), PMG_EBUSY 6 (serving admission refused),
 *   PMG_ESCRATCH 7 (beyond the fixed scratch of an allocation-free call).
 *
 * Host pointers are accepted; the library copies inputs to HBM, builds the
 * hierarchy there, solves, and copies results back.  The `*_ctx` calls keep
 * a solve's hierarchy resident between calls.
 */
#ifndef PMG_MI355X_H
#define PMG_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMG_OK 0
#define PMG_ITMAX 1
#define PMG_EBADPARM 2
#define PMG_ENOTSPD 3
#define PMG_ENOMEM 4
#define PMG_EDEVICE 5
#define PMG_EBUSY 6
#define PMG_ESCRATCH 7

/* Replaces pmgc::Vmgsz (pmgc/mgdrvd.h:236-255, mgdrvd.cc:505-601).
 * out[12] = {nxc,nyc,nzc,nf,nc,narr,narrc,n_rpc,n_iz,n_ipc,iretot,iintot}.
 * Host-only arithmetic; needs no GPU. */
int pmg_mgsz(int mgcoar, int mgdisc, int mgsolv, int nx, int ny, int nz,
             int nlev, int *out);

/* pmgc::Vmgsz with its own pointer list (pmgc/mgdrvd.h:236-255): reads
 * mgcoar..nlev, writes nxc..iintot.  PMG_EBADPARM where the reference
 * prints "invalid ... parameter" (mgdrvd.cc:561, 570, 585).  Host-only. */
int pmg_Vmgsz(int *mgcoar, int *mgdisc, int *mgsolv, int *nx, int *ny,
              int *nz, int *nlev, int *nxc, int *nyc, int *nzc, int *nf,
              int *nc, int *narr, int *narrc, int *n_rpc, int *n_iz,
              int *n_ipc, int *iretot, int *iintot);

/* pmgc::Vpackmg (pmgc/mgsubd.h:273, mgsubd.cc:538-576): encodes iparm(1..22)
 * and rparm(1), (9), (10) from the reference's pointer list.  Host-only. */
int pmg_Vpackmg(int *iparm, double *rparm, size_t *nrwk, int *niwk, int *nx,
                int *ny, int *nz, int *nlev, int *nu1, int *nu2, int *mgkey,
                int *itmax, int *istop, int *ipcon, int *nonlin, int *mgsmoo,
                int *mgprol, int *mgcoar, int *mgsolv, int *mgdisc, int *iinfo,
                double *errtol, int *ipkey, double *omegal, double *omegan,
                int *irite, int *iperf);

/* Replaces pmgc::Vmaxlev (pmgc/mgsubd.cc:456-484).  Host-only. */
int pmg_maxlev(int nx, int ny, int nz);

/* Replaces pmgc::Vmgdriv (pmgc/mgdrvd.h:79-98, mgdrvd.cc:60-65): the linear
 * PMG solve, same pointer list minus the sycl::queue.  FP64 throughout.
 * On return u holds the solution with Dirichlet boundary values restored
 * (mgdrvd.cc:500-502); iwork holds iz/ipc and rwork holds rpc/cc/fc/pc/ac
 * as the reference leaves them (mgdrvd.cc:145-166).  Unlike the reference,
 * a1cf..ccf, fcf and tcf are not overwritten.  iparm(21) = mgsolv 0 solves
 * the coarsest level by CG (cgd.cc:60-124), 1 by the factored band
 * (buildBd.cc:251-304, mlinpckd.cc:167-235), left in rwork at
 * ac(iz(7,nlev+1)); an indefinite band returns PMG_ENOTSPD.  iparm(11) =
 * istop: 0 absolute and 1 relative L1 residual; 2 L1 difference of
 * successive iterates (the first against tcf) over sqrt(nx ny nz); 3 and 4
 * the relative 2-norm error against tcf; 5 the relative energy-norm error
 * (mgcsd.cc:126-152, 497-526).  4 and 5 first replace fc by A tcf on every
 * level, so that tcf is the solution (Vbuildalg, mgdrvd.cc:424-433). */
int pmg_mgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
               double *u, const double *xf, const double *yf, const double *zf,
               const double *gxcf, const double *gycf, const double *gzcf,
               double *a1cf, double *a2cf, double *a3cf, double *ccf,
               double *fcf, double *tcf);

/* Replaces pmgc::Vnewdriv (pmgc/newdrvd.h, newdrvd.cc:60-155): the
 * nonlinear (sinh) Poisson-Boltzmann Newton solve with PMG inner solves.
 * rwork needs iretot + 2*nf entries (newdrvd.cc:123).  The interior of u
 * is the starting iterate (its boundary is zeroed, newdrvd.cc:263).
 * iparm(9) = mgkey 0 runs Vnewton, 1 the nested iteration Vfnewton
 * (newtond.cc:60-138): Newton on each level from the coarsest up, each
 * level's solution prolongated as the next finer level's start (the
 * coarse iterates start from zero).  istop 0 and 1 only. */
int pmg_newdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                double *u, const double *xf, const double *yf,
                const double *zf, const double *gxcf, const double *gycf,
                const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                double *ccf, double *fcf, double *tcf);

/* The reference's class-style solver (PBE_pmgc_class.h:1070-1133 with
 * CG_Solver.h:45-242, run by PBE_pmgc_class_with_non_zero_IS.cc:129-136):
 * CG preconditioned by the class V-cycle (two zero-start red-black sweeps,
 * restriction, recursion, coarsest CG to an absolute 1e-8 from its previous
 * solution, Hackbusch-Reusken-damped prolongation, two sweeps in the same
 * colour order), here on the hierarchy pmg_mgdriv builds; same arguments,
 * iwork and rwork as pmg_mgdriv.  num_iters > 0: that many CG iterations
 * (the reference's counting criterion); 0: until the relative L1 residual
 * of the CG residual is <= rparm(1) (errtol) or iparm(10) iterations.
 * Needs nlev >= 2 and a 27-point coarsest level of at most 13^3. */
int pmg_pcgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                double *u, const double *xf, const double *yf,
                const double *zf, const double *gxcf, const double *gycf,
                const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                double *ccf, double *fcf, double *tcf, int num_iters);

/* Replaces pmgc::Vmypdefinitlpbe / Vmypdefinitnpbe
 * (pmgc/mypdec.cc:72-106): the ion table used by the nonlinear term (at
 * most 50 species, MAXIONS of pmgc/mypdec.h:63; further ones are ignored
 * as the reference ignores them). */
int pmg_mypdefinit(int nion, const double *charge, const double *sconc);
/* The same with the reference's pointer lists: pmgc::Vmypdefinitlpbe
 * (pmgc/mypdec.h:89, mypdec.cc:72-88) and pmgc::Vmypdefinitnpbe
 * (mypdec.h:106, mypdec.cc:90-106). */
int pmg_Vmypdefinitlpbe(int *nion, double *charge, double *sconc);
int pmg_Vmypdefinitnpbe(int *nion, double *charge, double *sconc);

/* Iterations, error flag and final relative residual of the last
 * pmg_mgdriv / pmg_newdriv call on this thread.  The reference computes
 * these and drops them (pmgc/mgdrvd.cc:194-195, 452-455). */
int pmg_last_stats(int *iters, int *ierror, double *relres);
/* The relative stopping measure after each V-cycle of the last pmg_mgdriv
 * call on this thread, or after each finest-level Newton iteration of the
 * last pmg_newdriv call (what the reference's Vprtstp reports per
 * iteration, pmgc/mgsubd.cc:486-536; with iparm(12) iinfo > 0 the library
 * also prints it, with the contraction number for V-cycles, to stdout).
 * Copies at most cap values; returns the number of iterations. */
int pmg_last_history(double *relres, int cap);

/* --- resident-hierarchy API (no reference counterpart: the reference
 * rebuilds everything per Vmgdriv call; these keep a solve's operators in
 * HBM for repeated solves, mixed precision and serving) --- */

typedef struct pmg_ctx pmg_ctx;

/* Precision modes of pmg_ctx_solve. */
#define PMG_FP64 0  /* plain FP64 V-cycles: the drop-in arithmetic      */
#define PMG_MIXED 1 /* FP32 V-cycle inside FP64 defect correction       */

/* Create a context: decodes iparm/rparm exactly as pmg_mgdriv, uploads the
 * fine-level inputs and builds the hierarchy (VbuildA, VbuildP, VbuildG)
 * in FP64 on the GPU. */
int pmg_ctx_create(pmg_ctx **out, const int *iparm, const double *rparm,
                   const double *xf, const double *yf, const double *zf,
                   const double *gxcf, const double *gycf, const double *gzcf,
                   const double *a1cf, const double *a2cf, const double *a3cf,
                   const double *ccf, const double *fcf);
/* Solve on the resident hierarchy (mode PMG_FP64 or PMG_MIXED).  The
 * solution stays in HBM until pmg_ctx_get_u. */
int pmg_ctx_solve(pmg_ctx *ctx, int mode);
/* Solve n (<= 64) distinct contexts together: their V-cycle iterations run
 * in lockstep with every kernel launched once for all of them (the
 * grouped kernels of the serving loop), each context leaving the group when
 * it meets its stopping test.  Each result, iteration count and history is
 * bit-identical to pmg_ctx_solve of that context alone.  Linear contexts
 * with istop 0 or 1, nlev >= 2, all with the same nu1 and nu2
 * (PMG_EBADPARM otherwise).  No reference counterpart. */
int pmg_ctx_solve_group(pmg_ctx *const *ctx, int n, int mode);
/* The class-style PCG of pmg_pcgdriv on a resident context; PMG_MIXED runs
 * the V-cycle preconditioner in FP32 under the FP64 CG.  Statistics as
 * pmg_ctx_solve (iterations are CG iterations). */
int pmg_ctx_pcg(pmg_ctx *ctx, int mode, int num_iters);
/* Copy the solution (with Dirichlet boundary values) to host u (nx*ny*nz). */
int pmg_ctx_get_u(pmg_ctx *ctx, double *u);
/* Stats of the last solve on ctx: iterations, ierror, relative residual,
 * device milliseconds of setup and of the solve. */
int pmg_ctx_stats(pmg_ctx *ctx, int *iters, int *ierror, double *relres,
                  double *setup_ms, double *solve_ms);
/* Rebuild the hierarchy on the GPU from the resident inputs. */
int pmg_ctx_rebuild(pmg_ctx *ctx);
/* The comparison solution (tcf, fine level, nx*ny*nz) of the stopping
 * tests iparm(11) = istop 2-5 (pmgc/mgcsd.cc:126-152, 497-526); needed
 * before pmg_ctx_solve with those.  istop 4 and 5 also replace the
 * right-hand side of every level by A tcf at the start of each solve (the
 * reference's Vbuildalg, mgdrvd.cc:424-433), until the next rebuild. */
int pmg_ctx_set_tru(pmg_ctx *ctx, const double *tcf);
/* Per-cycle relative stopping measure of the last solve on ctx, as
 * pmg_last_history. */
int pmg_ctx_history(pmg_ctx *ctx, double *relres, int cap);
/* Device bytes held by the context. */
size_t pmg_ctx_bytes(const pmg_ctx *ctx);
void pmg_ctx_destroy(pmg_ctx *ctx);

/* Copy out the FP64 operators built on the GPU for level lev (1-based):
 * ac (4 or 14 planes), cc, fc, and pc (27 planes of P from level lev+1 to
 * lev, coarse-sized; ignored on the coarsest level).  Any pointer may be
 * NULL.  Used to check VbuildA/VbuildP/VbuildG output. */
int pmg_ctx_get_level(pmg_ctx *ctx, int lev, double *ac, double *cc,
                      double *fc, double *pc);

/* Kernel timing (no reference counterpart): when enabled, the fine-level
 * smoothing launches of every later solve on ctx are bracketed by HIP
 * events on the context's stream (at most max_events events per solve);
 * pmg_ctx_ktime returns the accumulated device milliseconds and the number
 * of full red-black sweeps they cover. */
int pmg_ctx_ktime_enable(pmg_ctx *ctx, int max_events);
/* Replay each solve iteration from a captured HIP graph (default on; event
 * timing of the smoothing kernels uses the direct launches instead). */
int pmg_ctx_set_graphs(pmg_ctx *ctx, int on);
int pmg_ctx_ktime(pmg_ctx *ctx, double *ms, long long *sweeps);

/* --- single-kernel entry points on host arrays (parity checking of the
 * individual HIP kernels against the reference routines) --- */

/* Vmatvec / Vmresid (pmgc/matvecd.cc:244-261, 512-528), numdia 7 or 27. */
int pmg_matvec(int nx, int ny, int nz, int numdia, const double *ac,
               const double *cc, const double *x, double *y);
int pmg_mresid(int nx, int ny, int nz, int numdia, const double *ac,
               const double *cc, const double *fc, const double *x,
               double *r);
/* Vgsrb7x / Vgsrb27x (pmgc/gsd.cc:60-212): itmax sweeps, in place on x. */
int pmg_gsrb(int nx, int ny, int nz, int numdia, const double *ac,
             const double *cc, const double *fc, double *x, int itmax,
             int iadjoint);
/* Vrestrc / VinterpPMG (pmgc/matvecd.cc:801-817, 966-983). */
int pmg_restrc(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
               const double *xin, double *xout, const double *pc);
int pmg_interp(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
               const double *xin, double *xout, const double *pc);
/* The coarse-grid correction of Vmvcs's way up (pmgc/mgcsd.cc:429-458):
 * xf(interior) += xdamp * VinterpPMG(xc) with the Hackbusch-Reusken factor
 * xdamp = (xc . w0c) / (xc . A_c xc) of Vmatvec + 2 x Vxdot on the coarse
 * level (acc/ccc its operator, w0c its right-hand side; xdamp = 0 when the
 * denominator is 0). */
int pmg_correct(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                int numdia, const double *acc, const double *ccc,
                const double *xc, const double *w0c, const double *pc,
                double *xf);
/* pmg_gsrb starting from x = 0 (Vazeros before Vsmooth, mgcsd.cc:341-346):
 * x's interior is not read; its boundary entries must be 0. */
int pmg_gsrb_zero(int nx, int ny, int nz, int numdia, const double *ac,
                  const double *cc, const double *fc, double *x, int itmax,
                  int iadjoint);
/* FP32 instances of the same kernels, as the inner V-cycle of the
 * mixed-precision solve runs them (operators rounded to FP32 by the
 * caller).  zero_start as pmg_gsrb_zero. */
int pmg_mresid_f32(int nx, int ny, int nz, int numdia, const float *ac,
                   const float *cc, const float *fc, const float *x,
                   float *r);
int pmg_gsrb_f32(int nx, int ny, int nz, int numdia, const float *ac,
                 const float *cc, const float *fc, float *x, int itmax,
                 int iadjoint, int zero_start);
int pmg_restrc_f32(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                   const float *xin, float *xout, const float *pc);
int pmg_interp_f32(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                   const float *xin, float *xout, const float *pc);
int pmg_correct_f32(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                    int numdia, const float *acc, const float *ccc,
                    const float *xc, const float *w0c, const float *pc,
                    float *xf);
/* The outer residual of the mixed-precision solve: r = f - A x of the
 * 7-point fine operator in FP64 (Vmresid7_1s, pmgc/matvecd.cc:381-420),
 * stored rounded to FP32 (the next correction cycle's right-hand side);
 * *l1 (may be NULL) = its FP64 L1 norm (Vxnrm1, mikpckd.cc:133-159). */
int pmg_mresid_mixed(int nx, int ny, int nz, const double *ac,
                     const double *cc, const double *fc, const double *x,
                     float *r, double *l1);
/* pmg_mresid_mixed without any allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops).  All pointers are DEVICE pointers
 * the caller owns; l1 (required) receives the FP64 L1 norm; the work is
 * enqueued on stream (NULL: the null stream) and the call does not wait.
 * The norm's block partials use a fixed 1 MiB device array per GPU held
 * until the process ends, so grids needing more than 131072 partials
 * (beyond about 513^3) return PMG_ESCRATCH.  Calls that overlap on one GPU
 * share that array and race. */
int pmg_mresid_mixed_noalloc(int nx, int ny, int nz, const double *ac,
                             const double *cc, const double *fc,
                             const double *x, float *r, double *l1,
                             void *stream);
/* The update of the mixed-precision solve: u(interior) += e (FP64 += FP32,
 * Vxaxpy with alpha 1, mikpckd.cc:115-131). */
int pmg_addcorr_f32(int nx, int ny, int nz, const float *e, double *u);

/* VbuildP / VbuildG (pmgc/buildPd.cc:1118-1147, buildGd.cc:3276-3368). */
int pmg_buildP(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
               int mgprol, int numdia, const double *acf, double *pc);
int pmg_buildG(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
               int numdia, const double *pc, const double *acf, double *acc);

/* --- Poisson-Boltzmann inputs from an atom list (no reference
 * counterpart as a C call: the reference assembles them inside its solver
 * class, PBE_pmgc_class.h:191-344, from atoms read out of PQR files) --- */

/* Seeded synthetic molecule: natoms x {x, y, z, radius, charge} into atoms.
 * Atoms uniform in a ball of radius sphere_r, radii U[1.2, 2.0] A;
 * charge_mode 0: +-1 with net zero, 1: uniform over {-1,-0.5,0,0.5,1}.
 * Host-only; needs no GPU. */
int pmg_pbe_molecule(int natoms, double sphere_r, int charge_mode,
                     uint64_t seed, double *atoms);

/* Assemble the fine-level Vmgdriv inputs of an n^3 grid of spacing h
 * (centred on the origin) on the GPU: dielectric faces (2 inside the atom
 * spheres, 78.54 outside, harmonic averages), kappa^2 map for a 1:1 salt of
 * ionic_strength mol/L with a 2 A ion-exclusion shell, trilinear charge
 * source 4*pi*lambda_B*q/h^3, Debye-Hueckel boundary faces.  Host arrays:
 * xf/yf/zf n; gxcf/gycf/gzcf the two planes (low, high) of each face,
 * 2*n*n entries (the (n,n,4) arrays of the driver hold them first);
 * a1cf..tcf n^3 (tcf zeroed).  Any output may be NULL. */
int pmg_pbe_assemble(int n, double h, double ionic_strength, int natoms,
                     const double *atoms, double *xf, double *yf, double *zf,
                     double *gxcf, double *gycf, double *gzcf, double *a1cf,
                     double *a2cf, double *a3cf, double *ccf, double *fcf,
                     double *tcf);

/* pmg_pbe_assemble with the dielectric constants inside / outside the atom
 * spheres and the ion-exclusion shell width (A) given (pmg_pbe_assemble:
 * 2, 78.54, 2). */
int pmg_pbe_assemble_ex(int n, double h, double ionic_strength, double eps_in,
                        double eps_out, double ion_shell, int natoms,
                        const double *atoms, double *xf, double *yf,
                        double *zf, double *gxcf, double *gycf, double *gzcf,
                        double *a1cf, double *a2cf, double *a3cf, double *ccf,
                        double *fcf, double *tcf);

/* pmg_pbe_molecule + pmg_pbe_assemble; atoms (5*natoms) may be NULL. */
int pmg_pbe_synth(int n, double h, int natoms, double sphere_r,
                  double ionic_strength, int charge_mode, uint64_t seed,
                  double *xf, double *yf, double *zf, double *gxcf,
                  double *gycf, double *gzcf, double *a1cf, double *a2cf,
                  double *a3cf, double *ccf, double *fcf, double *tcf,
                  double *atoms);

/* Create a context whose fine-level inputs are assembled on the GPU from
 * atoms (as pmg_pbe_assemble; nx = ny = nz required), then build the
 * hierarchy.  Only the atom list crosses PCIe. */
int pmg_ctx_create_pbe(pmg_ctx **out, const int *iparm, const double *rparm,
                       double h, double ionic_strength, int natoms,
                       const double *atoms);
/* Re-assemble a context's inputs for another molecule on the same grid
 * size and rebuild its hierarchy (graphs and buffers are kept). */
int pmg_ctx_load_pbe(pmg_ctx *ctx, double h, double ionic_strength,
                     int natoms, const double *atoms);

/* --- serving loop (no reference counterpart: the reference solves one
 * problem per process).  Continuous batching of linear PB solves of
 * uneven grid sizes: requests are admitted as solve slots and device
 * memory free up, their contexts are pooled by solve shape, and the
 * iterations of all admitted solves overlap on the device (one HIP stream
 * each, graph-captured V-cycle iterations).  Results are those of the same
 * solve run alone. --- */
typedef struct pmg_server pmg_server;

/* Start a server on the current GPU: at most max_active solves in flight,
 * contexts within mem_budget bytes of HBM (0: 90% of the free memory). */
int pmg_server_create(pmg_server **out, size_t mem_budget, int max_active);
/* The same with grouped iterations: the V-cycle iterations of the solves in
 * flight run in `groups` groups (0: each solve on its own stream, as
 * pmg_server_create), a group launching each kernel of an iteration once
 * for up to 8 solves; solves join and leave a group between iterations.
 * Results are bit-identical either way (pmg_ctx_solve_group). */
int pmg_server_create_grouped(pmg_server **out, size_t mem_budget,
                              int max_active, int groups);
/* Queue the solve of an atom list (natoms x {x, y, z, radius, charge},
 * copied) on the cubic grid iparm/rparm describe (linear: iparm(16) = 0),
 * spacing h, 1:1 salt of ionic_strength mol/L, mode PMG_FP64 or PMG_MIXED.
 * u_out (n^3, may be NULL) receives the potential and must stay valid until
 * pmg_server_wait returns for the ticket.  PMG_EBUSY: the solve can never
 * fit the memory budget. */
int pmg_server_submit(pmg_server *srv, const int *iparm, const double *rparm,
                      int mode, double h, double ionic_strength, int natoms,
                      const double *atoms, double *u_out, int64_t *ticket);
/* Wait for a ticket (timeout_ms < 0: no limit; PMG_EBUSY on timeout, the
 * ticket stays valid); returns the solve's status code, releases the ticket.
 * latency_ms: submission to completion, queueing included. */
int pmg_server_wait(pmg_server *srv, int64_t ticket, int timeout_ms,
                    int *iters, double *relres, double *latency_ms);
/* Completed solves, latency percentiles (ms), completed solves per second
 * from the first submission to the last completion, budgeted bytes held,
 * most solves seen in flight at once. */
int pmg_server_stats(pmg_server *srv, long long *completed, double *p50_ms,
                     double *p99_ms, double *solves_per_s, size_t *mem_used,
                     int *max_concurrent);
/* Start the statistics of pmg_server_stats afresh (completed requests,
 * latencies, the throughput window, the concurrency high-water mark); the
 * context pool is kept.  PMG_EBUSY while requests are in flight. */
int pmg_server_reset_stats(pmg_server *srv);
/* Finish every submitted request, then free the server. */
void pmg_server_destroy(pmg_server *srv);

/* --- OpenDX scalar fields (the reference writes them with
 * Domain.h:136-185; APBS writes its potentials in the same format).  u is
 * in the PMG layout (x fastest); files list values x slowest, z fastest.
 * Host-only. --- */
int pmg_dx_write(const char *path, int nx, int ny, int nz, double x0,
                 double y0, double z0, double h, const double *u,
                 const char *title);
/* Reads counts[3], origin[3] and the grid spacings delta[3] (any may be
 * NULL except counts); u (capacity cap values) may be NULL to read only the
 * header. */
int pmg_dx_read(const char *path, int *counts, double *origin, double *delta,
                double *u, size_t cap);

/* Version string of the library build. */
const char *pmg_version(void);

#ifdef __cplusplus
}
#endif
#endif
