/*
 * pmg_oracle.h -- CPU oracle for the PMG (pmgc) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's PMG solver (the pmgc/ sources, APBS semantics) used as the checker
 * for the HIP implementation in multigrid_amd/.  Only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * path never does.
 *
 * Conventions follow the reference: arrays are 1-based, Fortran order
 * (i fastest), VAT3(a,i,j,k) = a[(k-1)*nx*ny + (j-1)*nx + (i-1)]
 * (pmgc/vmatrix.h:79-80).  Operators are stored as planes of nx*ny*nz:
 * 7-point [oC,oE,oN,uC]; 27-point symmetric
 * [oC,oE,oN,uC,oNE,oNW,uE,uW,uN,uS,uNE,uNW,uSE,uSW] (pmgc/matvecd.cc:185-189).
 * Prolongation: 27 planes per coarse point
 * [oPC,oPN,oPS,oPE,oPW,oPNE,oPNW,oPSE,oPSW, uP*(9), dP*(9)]
 * (pmgc/matvecd.cc:809-816).
 *
 * Parity status of each function is recorded beside its declaration:
 * "pinned" names the independent check in tests/ that fixes it.
 */
#ifndef PMG_ORACLE_H
#define PMG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sizing: Vmgsz (pmgc/mgdrvd.cc:505-601).  out[12] =
 * {nxc,nyc,nzc,nf,nc,narr,narrc,n_rpc,n_iz,n_ipc,iretot,iintot}.
 * Pinned: closed-form integer arithmetic (tests/test_oracle_sizes.py). */
int pmgo_mgsz(int mgcoar, int mgdisc, int mgsolv, int nx, int ny, int nz,
              int nlev, int *out);
/* Vmaxlev (pmgc/mgsubd.cc:456-484).  Pinned: closed form. */
int pmgo_maxlev(int n1, int n2, int n3);
/* Vbuildstr (pmgc/mgsubd.cc:266-347): iz(50, nlev+1), 1-based offsets. */
void pmgo_buildstr(int nx, int ny, int nz, int nlev, int *iz);

/* VbuildA_fv (pmgc/buildAd.cc:62-212), box discretisation with Dirichlet
 * values folded into fc.  ac = 4 planes.  Pinned: uniform-grid closed form. */
void pmgo_buildA_fv(int nx, int ny, int nz, int ipkey, int *ipc, double *ac,
                    double *cc, double *fc, const double *xf, const double *yf,
                    const double *zf, const double *gxcf, const double *gycf,
                    const double *gzcf, const double *a1cf, const double *a2cf,
                    const double *a3cf, const double *ccf, const double *fcf);

/* VbuildP (pmgc/buildPd.cc:1118-1147): mgprol 0 = trilinear
 * (VbuildPb_trilin :62-152), 1 = operator-based (VbuildPb_op7 :172-554 for
 * numdia 7, VbuildPb_op27 :578-1090 for numdia 27).  Writes interior coarse
 * points of pc (27 planes); boundary entries are left as found (callers
 * zero them).  Pinned: constant-coefficient -> trilinear weights, layered
 * media -> flux-weighted 1-D weights, and explicit reference formulas for a
 * set of weights (tests/test_oracle_transfer.py). */
void pmgo_buildP(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 int mgprol, int numdia, const double *acf, double *pc);

/* VbuildG (pmgc/buildGd.cc:3276-3368 -> VbuildG_7 :437-1361 /
 * VbuildG_27 :1363-3274): Galerkin P^T A P on interior coarse points,
 * 14 output planes.  Pinned: dense brute-force P^T A P
 * (tests/test_oracle_galerkin.py). */
void pmgo_buildG(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 int numdia, const double *pc, const double *acf, double *acc);

/* Vmatvec / Vmresid (pmgc/matvecd.cc:61-261, 381-528).  numdia 7 or 27.
 * Pinned: dense brute force + symmetry (tests/test_oracle_stencil.py). */
void pmgo_matvec(int nx, int ny, int nz, int numdia, const double *ac,
                 const double *cc, const double *x, double *y);
void pmgo_mresid(int nx, int ny, int nz, int numdia, const double *ac,
                 const double *cc, const double *fc, const double *x,
                 double *r);
/* Vnmresid (pmgc/matvecd.cc:530-693) with the ion term c_vec. */
void pmgo_nmresid(int nx, int ny, int nz, int numdia, const double *ac,
                  const double *cc, const double *fc, const double *x,
                  double *r, double *w1);

/* Vgsrb7x / Vgsrb27x (pmgc/gsd.cc:60-111, 113-212). Pinned: fixed point
 * of the exact discrete solution, colour independence, serial sweep. */
void pmgo_gsrb(int nx, int ny, int nz, int numdia, const double *ac,
               const double *cc, const double *fc, double *x, int itmax,
               int iadjoint);

/* Vrestrc2 / VinterpPMG2 / Vextrac (pmgc/matvecd.cc:695-1012).
 * Pinned: adjointness <R r,e> = <r,P e>, linear reproduction. */
void pmgo_restrc(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 const double *xin, double *xout, const double *pc);
void pmgo_interp(int nxc, int nyc, int nzc, int nxf, int nyf, int nzf,
                 const double *xin, double *xout, const double *pc);
void pmgo_extrac(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc,
                 const double *xin, double *xout);

/* BLAS-1 and boundary helpers (pmgc/mikpckd.cc:61-350), interior ranges,
 * deterministic plane-ordered reductions. */
double pmgo_xnrm1(int nx, int ny, int nz, const double *x);
double pmgo_xnrm2(int nx, int ny, int nz, const double *x);
double pmgo_xdot(int nx, int ny, int nz, const double *x, const double *y);
void pmgo_xaxpy(int nx, int ny, int nz, double alpha, const double *x,
                double *y);
void pmgo_xcopy(int nx, int ny, int nz, const double *x, double *y);
void pmgo_fbound00(int nx, int ny, int nz, double *x);
void pmgo_fbound(int ibound, int nx, int ny, int nz, double *x,
                 const double *gxc, const double *gyc, const double *gzc);

/* Vcghs (pmgc/cgd.cc:60-124): CG on one level; returns iterations. */
int pmgo_cghs(int nx, int ny, int nz, int numdia, const double *ac,
              const double *cc, const double *fc, double *x, double *p,
              double *ap, double *r, int itmax, double errtol);

/* Banded coarsest solve, mgsolv = 1.  pmgo_buildband: Vbuildband
 * (pmgc/buildBd.cc:251-304) with Vbuildband1_7 / Vbuildband1_27 (:62-249):
 * LINPACK band storage abd(lda, n) of the interior operator, diagonal oC
 * only, then factored in place; ipcb = {n, m, lda, 1}; returns the Vdpbfa
 * info (0, or the first non-positive pivot).  pmgo_dpbfa / pmgo_dpbsl:
 * Vdpbfa / Vdpbsl (pmgc/mlinpckd.cc:167-235).  pmgo_bandsolve: the
 * coarsest-level use (pmgc/mgcsd.cc:391-407), wk of n doubles.
 * Pinned: band entries against the dense interior matrix with cc removed,
 * R^T R == A, and solves against numpy's dense solver
 * (tests/test_oracle_band.py). */
int pmgo_buildband(int nx, int ny, int nz, int numdia, const double *ac,
                   double *abd, int *ipcb);
void pmgo_dpbfa(double *abd, int lda, int n, int m, int *info);
void pmgo_dpbsl(const double *abd, int lda, int n, int m, double *b);
void pmgo_bandsolve(int nx, int ny, int nz, const double *abd, const int *ipcb,
                    const double *rhs, double *x, double *wk);

/* Vmypdefinitlpbe / npbe (pmgc/mypdec.cc:72-106): ion table. */
void pmgo_mypdefinit(int nion, const double *charge, const double *sconc);
/* Vc_vec / Vdc_vec (pmgc/mypdec.cc:136-212, 348-414), ipkey 0. */
void pmgo_c_vec(const double *coef, const double *uin, double *uout, int nx,
                int ny, int nz);
void pmgo_dc_vec(const double *coef, const double *uin, double *uout, int nx,
                 int ny, int nz);

/* Vmgdriv (pmgc/mgdrvd.cc:60-171 + Vmgdriv2 :173-503): full linear drop-in
 * driver.  Returns 0 ok, 1 itmax reached, 2 bad iparm. */
int pmgo_mgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                double *u, const double *xf, const double *yf,
                const double *zf, const double *gxcf, const double *gycf,
                const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                double *ccf, double *fcf, double *tcf);
/* Vnewdriv (pmgc/newdrvd.cc:60-303 -> Vnewton, newtond.cc:140-486). */
int pmgo_newdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                 double *u, const double *xf, const double *yf,
                 const double *zf, const double *gxcf, const double *gycf,
                 const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                 double *ccf, double *fcf, double *tcf);
/* Class-style solver (PBE_pmgc_class.h:1070-1133, CG_Solver.h:45-242): CG
 * preconditioned by the class V-cycle, on the Vmgdriv hierarchy; see
 * pmg_oracle.c.  num_iters > 0 runs that many CG iterations, 0 runs to
 * errtol (L1 residual, relative) or itmax.  pmgo_cg_class: the coarsest
 * solve (absolute 1e-8 criterion, warm start).  Pinned: identity-free
 * checks in tests/test_oracle_pcg.py (CG on a dense SPD system, the
 * V-cycle preconditioner is linear, PCG converges to the dense solution). */
int pmgo_pcgdriv(int *iparm, double *rparm, int *iwork, double *rwork,
                 double *u, const double *xf, const double *yf,
                 const double *zf, const double *gxcf, const double *gycf,
                 const double *gzcf, double *a1cf, double *a2cf, double *a3cf,
                 double *ccf, double *fcf, double *tcf, int num_iters);
int pmgo_cg_class(int nx, int ny, int nz, int numdia, const double *ac,
                  const double *cc, const double *f, double *x, double *p,
                  double *ap, double *r, double thresh, int itmax);

/* iters / ierror / relative residual of the last driver call, plus the
 * per-cycle relative residual history (up to hist_cap entries). */
int pmgo_last_stats(int *iters, int *ierror, double *relres, double *hist,
                    int hist_cap);

/* Poisson-Boltzmann input assembly from an atom list (pbe_oracle.c;
 * PBE_pmgc_class.h:191-344).  atoms: natoms x {x, y, z, radius, charge}.
 * Writes the Vmgdriv fine-level inputs of an n^3 grid of spacing h centred
 * on the origin; gxcf/gycf/gzcf planes 1-2 of the (n,n,4) face arrays.
 * Pinned: charge conservation, the closed-form dielectric/kappa values,
 * the exclusion-shell geometry and the Debye-Hueckel sum evaluated
 * independently in tests/test_pbinput.py. */
void pbo_constants(double ionic, double *lambda_b, double *kappa2);
void pbo_constants_eps(double ionic, double eps_out, double *lambda_b,
                       double *kappa2);
/* As pbo_assemble with the dielectric constants inside / outside the atoms
 * and the ion-exclusion shell width as parameters (pbo_assemble: 2, 78.54,
 * 2 A). */
void pbo_assemble_ex(int n, double h, double ionic, double eps_in,
                     double eps_out, double shell, int natoms,
                     const double *atoms, double *xf, double *yf, double *zf,
                     double *gxcf, double *gycf, double *gzcf, double *a1cf,
                     double *a2cf, double *a3cf, double *ccf, double *fcf,
                     double *tcf);
void pbo_assemble(int n, double h, double ionic, int natoms,
                  const double *atoms, double *xf, double *yf, double *zf,
                  double *gxcf, double *gycf, double *gzcf, double *a1cf,
                  double *a2cf, double *a3cf, double *ccf, double *fcf,
                  double *tcf);

#ifdef __cplusplus
}
#endif
#endif
