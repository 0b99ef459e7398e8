/*
 * thp.h -- C ABI of libthp, the MI355X-native AMOEBA induced-dipole solver.
 *
 * Every entry point takes plain pointers and sizes.  Conventions:
 *   - return value: THP_OK (0) or a negative THP_ERR_* code; never exit().
 *     thp_last_error() returns a message for the last failure on a context.
 *   - atom indices are 0-based at this boundary.
 *   - host arrays are caller-owned and copied; device state is library-owned.
 *     The *_dev variants take device pointers (stream-ordered, zero copy).
 *   - array layouts follow the reference's Fortran arrays in C row-major
 *     order: rpole[n][13] (q, dx,dy,dz, Qxx,Qxy,Qxz,Qyx,Qyy,Qyz,Qzx,Qzy,Qzz,
 *     global frame, quadrupole in the internal /3 convention), vectors over
 *     both right-hand sides as [n][2][3] (rhs 0 = d, rhs 1 = p), cphi[n][10].
 *   - boxes are orthorhombic (three edge lengths).
 *   - a thp_ctx is used from one host thread at a time.  While its serving
 *     loop runs (thp_serve_start .. thp_serve_stop) only thp_serve_* calls
 *     may be made on it, and those from any thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   thp_init/thp_finalize ...... C_init_env, C_get_cell
 *                                (GPU/source/cu_tmatxb_pme.cu:1035-1078,
 *                                 GPU/source/MOD_interfaces.f:684-690)
 *   thp_nblist ................. nblist / mlistcell
 *                                (v1.2/source/nblist.f:17-247;
 *                                 GPU/source/nblist.f:24-186)
 *   thp_efld0_direct ........... efld0_direct / cu_efld0_direct
 *                                (v1.2/source/newinduce_pme.f:948-1389;
 *                                 GPU/source/MOD_interfaces.f:610-636)
 *   thp_efld0_recip ............ efld0_recip (v1.2/source/newinduce_pme.f:1671-1917)
 *   thp_tmatxb ................. tmatxb_pme / cu_tmatxb_pme
 *                                (v1.2/source/newinduce_pme.f:1392-1668;
 *                                 GPU/source/MOD_interfaces.f:584-606)
 *   thp_tmatxb_recip ........... tmatxbrecip (v1.2/source/newinduce_pme.f:1919-2128)
 *   thp_grid_uind .............. grid_uind_site (v1.2/source/pmestuff.f:280-400)
 *   thp_fphi_uind .............. fphi_uind_site (v1.2/source/pmestuff.f:746-1037)
 *   thp_induce ................. newinduce_pme2 + inducepcg_pme2
 *                                (v1.2/source/newinduce_pme2.f:16-636;
 *                                 GPU/source/newinduce_pme2gpu.f:17-1137)
 *   thp_induce_mu .............. the same solve with uind and uinp in one
 *                                host array, as the reference keeps them in
 *                                its mu(3,nrhs,npolebloc) work array
 *                                (v1.2/source/newinduce_pme2.f:71,230-235)
 *   thp_set_solver ............. POLAR-ALG keyword (v1.2/source/prmkey.f:396-398)
 *                                selecting inducepcg_pme2 or inducejac_pme2
 *                                (v1.2/source/newinduce_pme2.f:210-217,
 *                                 638-890; diis.f) or the truncated CG
 *                                (v1.2/source/epolar3tcg.f:18-56)
 *   thp_system_set_frames ...... rotpole + rotmat + rotsite
 *                                (v1.2/source/rotpole.f:24-339) and chkpole
 *                                (v1.2/source/chkpole.f:20-98), each step
 *   thp_set_shortreal .......... use_polarshortreal with ewaldshortcut
 *                                (v1.2/source/cutoffs.f:59,228; switch.f:60-62):
 *                                newinduce_shortreal, inducepcg_shortreal,
 *                                inducejac_shortreal
 *                                (v1.2/source/newinduce_shortreal.f:16-610)
 *   thp_set_tcg ................ tcgorder/tcgprec/tcgguess/tcgpeek/tcgomega
 *                                (v1.2/source/initprm.f:447-455) for
 *                                epolar3tcg1/epolar3tcg2
 *                                (v1.2/source/epolar3tcg.f:66-621)
 *   thp_epolar1 ................ the gradient of epolar1c after the solve:
 *                                epreal1c, eprecip1, the self torque and
 *                                torque (v1.2/source/epolar1.f:57-2543;
 *                                 v1.2/source/torque.f)
 *   thp_set_predictor .......... polpred keyword + ulspred and the dipole
 *                                history (v1.2/source/mdinit.f:93,185-190;
 *                                 v1.2/source/newinduce_pme2.f:165-178,
 *                                 253-264, 910-1010; kpolar.f:405-427)
 */
#ifndef THP_H
#define THP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int thp_status;
#define THP_OK 0
#define THP_ERR_ARG (-1)      /* invalid argument */
#define THP_ERR_HIP (-2)      /* HIP runtime failure */
#define THP_ERR_FFT (-3)      /* grid FFT failure (reserved) */
#define THP_ERR_NOMEM (-4)    /* device allocation failed / over budget */
#define THP_ERR_LIST (-5)     /* cutoff+buffer exceeds half the box, or list capacity */
#define THP_ERR_NOCONV (-6)   /* PCG did not reach poleps within politer */
#define THP_ERR_NAN (-7)      /* non-finite residual in the PCG loop */
#define THP_ERR_STATE (-8)    /* call out of order (e.g. serve not started) */
#define THP_ERR_BUSY (-9)     /* serve queue full */

#define THP_PREC_FP64 0       /* FP64 pair, grid, FFT and vectors ("plain") */
#define THP_PREC_MIXED 1      /* FP32 pair/grid/FFT, FP64 accumulation and CG */

/* induced-dipole solver (polalg) */
#define THP_POLALG_PCG 1          /* preconditioned conjugate gradient (default) */
#define THP_POLALG_JACOBI_DIIS 2  /* Jacobi iterations with DIIS extrapolation */
#define THP_POLALG_TCG 3          /* truncated conjugate gradient (thp_set_tcg) */

/* induced-dipole predictor (polpred) */
#define THP_PRED_NONE 0
#define THP_PRED_GEAR 1       /* Gear binomial coefficients */
#define THP_PRED_ASPC 2       /* always stable predictor-corrector (Kolafa 2004) */
#define THP_PRED_LSQR 3       /* least-squares fit of the stored solves */

typedef struct thp_ctx thp_ctx;
typedef struct thp_sys thp_sys;

typedef struct {
  int n;                      /* number of polarizable sites */
  double box[3];              /* orthorhombic edge lengths (Angstrom) */
  const double* x;            /* [n] coordinates (Angstrom) */
  const double* y;
  const double* z;
  const double* rpole;        /* [n][13] global-frame multipoles */
  const double* polarity;     /* [n] polarizability (Angstrom^3) */
  const double* thole;        /* [n] Thole damping parameter */
  int nscaled;                /* number of scaled pairs */
  const int* scaled_ik;       /* [nscaled][2] atom pairs (any order) */
  const double* scaled_dpu;   /* [nscaled][3] d-, p-, u-scale of each pair */
  double aewald;              /* Ewald coefficient (1/Angstrom) */
  double cut;                 /* real-space cutoff (Angstrom) */
  double buffer;              /* neighbor-list buffer (Angstrom) */
  int nfft[3];                /* PME grid */
  int order;                  /* B-spline order (5) */
  int precision;              /* THP_PREC_FP64 or THP_PREC_MIXED */
} thp_system_desc;

typedef struct {
  double poleps;              /* PCG tolerance, reference criterion sqrt(r.z/(3n)) */
  int politer;                /* iteration cap */
  int guess;                  /* 0: mu = alpha*E (polgsf=0); 1: warm start from uind/uinp inputs */
} thp_solve_opts;

typedef struct {
  int niter;                  /* PCG iterations performed */
  int status;                 /* THP_OK, THP_ERR_NOCONV or THP_ERR_NAN */
  double resnrm;              /* final residual norm (reference units) */
  double ep;                  /* polarization energy, kcal/mol */
  int n_indefinite;           /* iterations where some rhs had p.(A p) <= 0 */
  double t_field_ms;          /* device time in efld0 (direct + recip + self) */
  double t_solve_ms;          /* device time in the PCG loop */
} thp_solve_stats;

/* ---------------------------------------------------------------- context */
thp_status thp_init(thp_ctx** ctx, int device);
thp_status thp_finalize(thp_ctx* ctx);
const char* thp_last_error(thp_ctx* ctx);
const char* thp_version(void);

/* ---------------------------------------------------------------- systems */
thp_status thp_system_create(thp_ctx* ctx, const thp_system_desc* desc, thp_sys** sys);
thp_status thp_system_update_coords(thp_sys* sys, const double* x, const double* y, const double* z);
/* new global-frame multipoles rpole [n][13] (as rotpole leaves them each MD
 * step, v1.2/source/rotpole.f); the next field and solve use them */
thp_status thp_system_update_multipoles(thp_sys* sys, const double* rpole);
/* Local-frame multipoles (rotpole, v1.2/source/rotpole.f:24-339, with
 * chkpole, v1.2/source/chkpole.f:20-98): per atom the axis type
 * (THP_AXIS_*), its frame atoms (0-based; -1 where the type takes none)
 * and the multipoles in that frame, [n][13] in the rpole convention (as
 * kmpole leaves pole: dipoles in e A, quadrupoles in e A^2 divided by 3).
 * The device rotates them to the global frame now and after every
 * coordinate update; thp_system_update_multipoles drops the frames. */
#define THP_AXIS_NONE 0
#define THP_AXIS_Z_ONLY 1
#define THP_AXIS_Z_THEN_X 2
#define THP_AXIS_BISECTOR 3
#define THP_AXIS_Z_BISECT 4
#define THP_AXIS_3_FOLD 5
thp_status thp_system_set_frames(thp_sys* sys, const int* axistype, const int* zaxis, const int* xaxis,
                                 const int* yaxis, const double* pole);
thp_status thp_system_destroy(thp_sys* sys);
/* device memory held by the system (bytes) */
thp_status thp_system_footprint(const thp_sys* sys, int64_t* bytes);

/* ---------------------------------------------------- path entry points */
/* (re)build the Verlet list; force_rebuild=0 rebuilds only when atoms moved
 * more than buffer/2 since the last build */
thp_status thp_nblist(thp_sys* sys, int force_rebuild);
/* pair set of the list as (i,k) rows with i<k, sorted; pass ik=NULL to
 * query the count.  cap = rows available in ik. */
thp_status thp_nblist_get(thp_sys* sys, int64_t* npairs, int* ik, int64_t cap);

/* The pair set the solver uses for the current coordinates, as (i,k) rows
 * with i<k in caller indices, sorted: the in-cutoff entries of the
 * compacted lists that the permanent-field pass writes and the matvec
 * walks, plus the scaled pairs within the cutoff.  It replaces the list the
 * reference's field and matvec loops walk with their r^2 <= off2 test
 * (v1.2/source/newinduce_pme.f:1160, 1548-1608 over nblist.f:17-247; the
 * GPU's nelstc prefix after reorder_nblist, GPU/source/nblistgpu.f90:101).
 * THP_ERR_LIST if an entry is held by only one of its two atoms.  ik=NULL
 * queries the count; cap = rows available in ik. */
thp_status thp_pairs_in_cut(thp_sys* sys, int64_t* npairs, int* ik, int64_t cap);
thp_status thp_efld0_direct(thp_sys* sys, double* ef /* [n][2][3] */);
thp_status thp_efld0_recip(thp_sys* sys, double* cphi /* [n][10] */);
/* Polarization energy gradient (kcal/mol/A, the reference's dep) for the
 * dipoles of the system's last solve, [n][3] in caller order.  Systems with
 * local frames get the multipole torques carried to the frame atoms; with
 * global multipoles dep holds the forces at fixed orientations only and
 * torque (optional, [n][3], kcal/mol) returns the torques on the sites.
 * After a short-range solve (thp_set_shortreal) it is the short-range
 * gradient: the real-space terms within the short cutoff, no reciprocal or
 * self terms (epolar1c with use_polarshortreal, v1.2/source/epolar1.f:
 * 1347-1356).  After a truncated-CG solve it is that energy's own
 * gradient (epolar1tcg, v1.2/source/epolar1tcg.f): the truncated dipoles
 * are not stationary, and the adjoint of the truncated CG's steps adds the
 * terms the stationary form leaves out.  THP_ERR_STATE unless the
 * dipoles are a solve of the current coordinates and multipoles (after
 * thp_system_update_coords, _update_multipoles or _set_frames, solve
 * again first). */
thp_status thp_epolar1(thp_sys* sys, double* dep, double* torque);
/* thp_epolar1 with the polarization virial vir[3][3] (kcal/mol; the
 * reference's vir from epolar1c: v1.2/source/epolar1.f:326-334 self,
 * 1152-1160 reciprocal, 2332-2340 real-space pairs, 2478-2526 torques):
 * the derivative of the polarization energy along a homogeneous strain of
 * the positions and the box, symmetrised, with multipoles that follow
 * their local frames (or stay fixed in the global frame without frames).
 * Real-space pair shares, the reciprocal sum's box and moment terms and the
 * frame forces' shares accumulate in FP64 and are summed in a fixed order.
 * vir may be NULL (then the call is thp_epolar1).  No virial for
 * truncated-CG solves (THP_ERR_ARG). */
thp_status thp_epolar1_vir(thp_sys* sys, double* dep, double* torque, double* vir);
thp_status thp_tmatxb(thp_sys* sys, const double* mu, double* efi, int dodiag);
thp_status thp_tmatxb_recip(thp_sys* sys, const double* mu, double* dipfield);
/* spread fractional induced dipoles fuind/fuinp [n][3] onto the device grid;
 * qgrid (optional) receives the grid as [n3][n2][n1][2] doubles */
thp_status thp_grid_uind(thp_sys* sys, const double* fuind, const double* fuinp, double* qgrid);
/* gather first derivatives of the potential of a grid [n3][n2][n1][2]
 * (NULL: the device grid) into fdip_phi1/fdip_phi2 [n][3] */
thp_status thp_fphi_uind(thp_sys* sys, const double* qgrid, double* fdip_phi1, double* fdip_phi2);
/* multiply the device grid by the scalar-sum table between the forward and
 * backward 3-D FFTs (efld0_recip / tmatxbrecip, convolution step) */
thp_status thp_pme_convolve(thp_sys* sys);

thp_status thp_induce(thp_sys* sys, const thp_solve_opts* opts, double* uind, double* uinp,
                      double* ep, thp_solve_stats* stats);
/* thp_induce with the dipoles as one host array mu[n][2][3] (uind, uinp per
 * atom; with opts->guess 1 it holds the starting dipoles): the layout the
 * solve keeps, so the result comes back in one copy without reordering */
thp_status thp_induce_mu(thp_sys* sys, const thp_solve_opts* opts, double* mu /* [n][2][3] */, double* ep,
                         thp_solve_stats* stats);
/* Select the solver of thp_induce (THP_POLALG_*).  Jacobi/DIIS stops when
 * the Jacobi increment max_rhs sqrt(|munew - mu|^2 / 3n) < poleps and keeps
 * up to 25 subspace vectors on the device (2.4 kB per atom). */
thp_status thp_set_solver(thp_sys* sys, int polalg);
/* Options of the truncated CG (THP_POLALG_TCG): order 1 or 2 conjugate-
 * gradient steps in closed form, the diagonal (polarity) preconditioner,
 * the start alpha ef instead of zero, and the peek step omega alpha r
 * (defaults 2, 1, 0, 1, 1.0).  A truncated solve does a fixed amount of work
 * (order + 1 or order + 2 applications of the polarization matrix) and has
 * no convergence test; its stats report niter = order. */
thp_status thp_set_tcg(thp_sys* sys, int order, int precond, int guess, int peek, double omega);
/* The peek step's omega fitted to the converged energy (tcgomegafit,
 * v1.2/source/kpolar.f:1059-1326, keywords in MOD_polpot.f:62-96): a PCG
 * solve at poleps 1e-10 gives E_CG; the truncated CG energy, linear in
 * omega, is bisected over [-5, 5] toward E_CG (20 steps, within max(1e-4
 * |E_CG|, 1e-6) kcal/mol), as the reference does every omegafitfreq steps.
 * On convergence the system's omega is set (*fitted = 1), else it is kept
 * (*fitted = 0); *omega returns the omega in use.  Needs the truncated CG
 * with the peek step selected (THP_ERR_STATE otherwise); solve again after
 * it. */
thp_status thp_tcg_omega_fit(thp_sys* sys, double* omega, int* fitted);
/* Short-range real-space mode (the reference's use_polarshortreal, for the
 * inner steps of RESPA1): with shortcut > 0 (at most the system's cutoff;
 * the reference's default is 5 A) the permanent field, the matvec and the
 * solves of thp_induce (PCG, Jacobi/DIIS or the truncated CG: the
 * reference's polalgshort) use only the real-space Ewald terms of pairs
 * within shortcut: no reciprocal field and no self term.  The predictor
 * and the omega fit are not used in this mode.  0 restores the full
 * solve. */
thp_status thp_set_shortreal(thp_sys* sys, double shortcut);
/* Select the induced-dipole predictor (THP_PRED_*) and clear its history.
 * While a predictor is set, every solve of the system stores its converged
 * dipoles (the last 7 are kept); once 7 are stored, the next solve starts
 * from their extrapolation instead of opts->guess, as use_pred does in the
 * reference. */
thp_status thp_set_predictor(thp_sys* sys, int method);

/* device-pointer variants: mu, efi, ef, uind... are device addresses */
thp_status thp_tmatxb_dev(thp_sys* sys, const double* mu, double* efi, int dodiag);
thp_status thp_induce_dev(thp_sys* sys, const thp_solve_opts* opts, double* mu_dev /* [n][2][3] */,
                          double* ep, thp_solve_stats* stats);

/* Device-pointer variants ordered on the caller's stream, the analogue of
 * the reference's native layer, which takes device arrays (from
 * `!$acc host_data use_device`) and a cudaStream_t by value
 * (GPU/source/MOD_interfaces.f:577-627: cu_tmatxb_pme, cu_efld0_direct;
 * GPU/source/newinduce_pmegpu.f:1091,1352 for the device-resident
 * efld0_recipgpu and tmatxbrecipgpu).  Arrays are device addresses in the
 * layouts of the host entry points; `stream` is a hipStream_t.  With a
 * stream, the call's work runs after what the caller queued on it, the
 * stream waits for the call's results, and the host is not blocked (only a
 * list or B-spline rebuild after atoms moved waits for the device).  With
 * stream = NULL the call returns when the results are in place.
 * Replace (same order): efld0_direct (v1.2/source/newinduce_pme.f:948-1389;
 * cu_efld0_direct, GPU/source/MOD_interfaces.f:610-636), efld0_recip
 * (newinduce_pme.f:1671-1917; efld0_recipgpu, GPU/source/newinduce_pmegpu.f:1091),
 * tmatxb_pme (newinduce_pme.f:1392-1668; cu_tmatxb_pme,
 * GPU/source/MOD_interfaces.f:584-606), tmatxbrecip (newinduce_pme.f:1919-2128;
 * tmatxbrecipgpu, GPU/source/newinduce_pmegpu.f:1352), grid_uind_site
 * (v1.2/source/pmestuff.f:280-400; grid_uind_site_p,
 * GPU/source/MOD_interfaces.f:224-227) and fphi_uind_site
 * (pmestuff.f:746-1037; fphi_uind_site2_p, MOD_interfaces.f:416-419). */
thp_status thp_efld0_direct_dev(thp_sys* sys, double* ef /* [n][2][3] */, void* stream);
thp_status thp_efld0_recip_dev(thp_sys* sys, double* cphi /* [n][10] */, void* stream);
thp_status thp_tmatxb_dev_async(thp_sys* sys, const double* mu, double* efi, int dodiag, void* stream);
thp_status thp_tmatxb_recip_dev(thp_sys* sys, const double* mu, double* dipfield, void* stream);
/* qgrid (optional): the spread grid as [n3][n2][n1][2] doubles */
thp_status thp_grid_uind_dev(thp_sys* sys, const double* fuind /* [n][3] */, const double* fuinp /* [n][3] */,
                             double* qgrid, void* stream);
/* qgrid NULL: the device grid (of the last spreading) */
thp_status thp_fphi_uind_dev(thp_sys* sys, const double* qgrid, double* fdip_phi1 /* [n][3] */,
                             double* fdip_phi2 /* [n][3] */, void* stream);

/* No counterpart in the reference: thp_epolar1 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, hosts that forbid hipMalloc).  dep and torque (optional) are device
 * addresses ([n][3], caller order, the values thp_epolar1 returns), ordered
 * on `stream` like the entry points above.  The forces and torques before
 * their gather live in a file-scope device array of 6 * THP_NOALLOC_MAX_ATOMS
 * doubles (6 MiB) that each device holds until the process ends, so the call
 * allocates nothing: THP_ERR_ARG for systems of more than
 * THP_NOALLOC_MAX_ATOMS atoms (rounded up to 64) and for truncated-CG
 * solves (their gradient's adjoint allocates), THP_ERR_STATE unless the
 * last solve of the current coordinates left its lists and B-spline tables
 * in place (as thp_induce does).  Every call on a device shares that
 * device's array: calls that overlap on one device (two streams, two
 * threads, two contexts) race. */
#define THP_NOALLOC_MAX_ATOMS 131072
thp_status thp_epolar1_noalloc(thp_sys* sys, double* dep, double* torque, void* stream);

/* ------------------------------------------------------------ measurement */
/* Measurement hook, no reference counterpart: average device time (ms) of
 * one launch of a path kernel over `reps` back-to-back launches, timed with
 * HIP events on the library's stream.  kernel: 0 = real-space matvec
 * (tmatxb), 1 = induced-dipole spreading, 2 = induced-dipole gather,
 * 3 = direct permanent-field pass (efld0 with the matvec's lists),
 * 4 = rebuild of the solver's lists (block candidates, block-union lists,
 * host round trips for their counts, and the permanent-field pass that
 * compacts them), 5 = rebuild of the Verlet pair list (thp_nblist),
 * 6 = the grid convolution (forward 3-D FFT, scalar sum, backward FFT).
 * Kernels 4 and 5 take one system.  Also
 * returns the number of unscaled pairs within the cutoff (i<k), the unit
 * of the matvec's algorithmic FLOP count. */
thp_status thp_time_kernel(thp_sys* sys, int kernel, int reps, double* ms_per_launch, int64_t* pairs_in_cut);

/* The same timing over one launch that covers several systems (distinct,
 * one context, one precision), with the work tables laid out as the
 * serving loop lays out its batches.  pairs_in_cut is the sum over the
 * systems. */
thp_status thp_time_kernel_batch(thp_sys* const* systems, int nsys, int kernel, int reps, double* ms_per_launch,
                                 int64_t* pairs_in_cut);

/* --------------------------------------------------------------- serving */
/* Serving loop, no single reference counterpart: it replaces an MD driver
 * that calls newinduce_pme2 once per system and step
 * (v1.2/source/newinduce_pme2.f:16-283) with a persistent worker thread
 * that solves many independent systems at once.  Requests are admitted as
 * slots (max_systems) and device memory (hbm_budget) free up; the solves
 * in flight share every kernel launch (each kernel walks a work table over
 * all of them, systems at different CG iterations included) and their
 * reciprocal grids of equal size are transformed by one batched FFT.
 * Each system keeps its coordinates, lists, tables and last dipoles
 * resident; the dipoles seed the next solve of the same system (the
 * polgsf = 1 branch, newinduce_pme2.f:189-196). */
typedef struct {
  int max_systems;            /* solves in flight at once (slots), >= 1 */
  int64_t hbm_budget;         /* device bytes for resident systems; 0 = 80% of free memory */
  int iters_per_graph;        /* CG iterations per launch chunk between status reads, >= 1 */
  int autotune;               /* 1: pick the pair kernel's wavefronts per block by measurement */
  double poleps;              /* PCG tolerance of every solve */
  int politer;                /* iteration cap of every solve */
  int warm_start;             /* 1: start from the system's previous dipoles when it has them */
  int use_graphs;             /* 1: replay each chunk as a captured HIP graph */
} thp_serve_opts;

typedef struct {
  uint64_t ticket;
  int system_id;
  int status;                 /* THP_OK / THP_ERR_NOCONV / THP_ERR_NAN / THP_ERR_LIST ... */
  int niter;
  double ep;                  /* polarization energy, kcal/mol */
  double latency_ms;          /* submit -> result */
  double resnrm;
  double queue_ms;            /* submit -> start of the solve */
  int warm;                   /* 1 if started from the previous dipoles */
} thp_result;

typedef struct {
  int64_t solves;             /* results published */
  int64_t chunks;             /* launch chunks run */
  double mean_active;         /* mean solves in flight per chunk */
  int64_t graph_launches;     /* chunks replayed as graphs */
  int64_t evictions;          /* resident systems released for memory */
  int64_t resident_bytes;     /* device bytes held by resident systems */
  int resident_systems;
  int pair_segs;              /* wavefronts per block of the last chunk's pair kernel */
  int64_t batch_builds;       /* iteration-batch layouts built (tables, graph) */
  double host_ms;             /* worker time spent enqueueing and admitting */
  double wait_ms;             /* worker time spent waiting for the device */
  double chunk_ms;            /*   of host_ms: enqueueing the iteration chunks */
  double admit_ms;            /*   of host_ms: admission (residency, coordinates, lists) */
  double setup_ms;            /*   of host_ms: enqueueing the set-up batches */
  double coords_ms;           /*   of admit_ms: coordinate updates */
  double lists_ms;            /*   of admit_ms: list checks and rebuilds */
} thp_serve_stats;

thp_status thp_serve_start(thp_ctx* ctx, const thp_serve_opts* opts);
/* register a system with the loop (host copy; device data is created when
 * it is first solved and stays resident until evicted for memory) */
thp_status thp_serve_add_system(thp_ctx* ctx, const thp_system_desc* desc, int* system_id);
/* request a solve of system_id at new coordinates (NULL: unchanged);
 * requests of one system run in submission order */
/* The grid shapes the serving loop's autotuner picked on this context for
 * a grid (nfft[3]) and precision: the largest spreading brick edge (4..8)
 * and the most FFT lines per workgroup (8, 16 or 32).  The loop measures
 * them with HIP events on a scratch system the first time a request brings
 * that grid (thp_serve_opts.autotune); every system created on the context
 * afterwards takes them, alone or in the loop.  THP_ERR_STATE if that grid
 * has not been tuned.  No reference counterpart: the reference picks its
 * kernel variants from a fixed table (sub_config,
 * GPU/source/MOD_interfaces.f:72-98, mechanic.f:907-975). */
thp_status thp_grid_shape(thp_ctx* ctx, const int* nfft, int precision, int* brick_max, int* fft_lines_max);

/* Local-frame multipoles for a registered serving system, before its first
 * request (THP_ERR_STATE after), in the form thp_system_set_frames takes:
 * axis types THP_AXIS_*, 0-based frame atoms (-1 for none) and pole [n][13]
 * in the local frame.  Every request's set-up then rotates them at
 * the request's coordinates on the device (rotpole and chkpole,
 * v1.2/source/rotpole.f:24-339, chkpole.f:20-98, as epolar1.f:113-116 does
 * before each step's field); thp_serve_submit_multipoles is refused for
 * such a system. */
thp_status thp_serve_set_frames(thp_ctx* ctx, int system_id, const int* axistype, const int* zaxis,
                                const int* xaxis, const int* yaxis, const double* pole);
thp_status thp_serve_submit(thp_ctx* ctx, int system_id, const double* x, const double* y,
                            const double* z, uint64_t* ticket);
/* the same with new global-frame multipoles rpole [n][13] (NULL: unchanged),
 * for requests that follow a trajectory (the multipoles rotate with the
 * molecules) */
thp_status thp_serve_submit_multipoles(thp_ctx* ctx, int system_id, const double* x, const double* y,
                                       const double* z, const double* rpole, uint64_t* ticket);
/* 1 if ready (result filled, ticket released), 0 if not yet; wait_ms < 0
 * blocks until ready.  ticket 0 takes any finished request (the oldest). */
thp_status thp_serve_poll(thp_ctx* ctx, uint64_t ticket, thp_result* result, double wait_ms);
/* dipoles [n][2][3] (original order) of the system's last finished solve;
 * waits for a solve of that system in flight */
thp_status thp_serve_get_dipoles(thp_ctx* ctx, int system_id, double* mu);
thp_status thp_serve_stats_get(thp_ctx* ctx, thp_serve_stats* stats);
/* stops the worker; solves not finished are reported THP_ERR_STATE */
thp_status thp_serve_stop(thp_ctx* ctx);

/* ------------------------------------------------------ synthetic inputs */
/* Seeded water box / water+solute system (host-only, no device use).
 * n atoms in a cubic box of edge L; n_solute atoms form a compact blob of
 * radius blob_radius at the box centre.  Outputs must hold n (x,y,z,
 * polarity, thole), 13n (rpole); scaled pairs: at most 3n rows. */
thp_status thp_make_system(int n, double box, int n_solute, double blob_radius, uint64_t seed,
                           double* x, double* y, double* z, double* rpole, double* polarity,
                           double* thole, int* nscaled, int* scaled_ik, double* scaled_dpu);
/* smallest 2,3,5-smooth even grid size >= minsize from the reference table
 * (v1.2/source/kewald.f:67-73) */
int thp_pick_grid(double minsize);

#ifdef __cplusplus
}
#endif
#endif /* THP_H */
