/*
 * cpfem.h — C ABI of the MI355X-native FASOLT3 hot path.
 *
 * Plain pointers and sizes only; every entry point returns an int that is
 * either 0 or a FASOLT3 error code of the reference's FAFERR table
 * (src/fasolt/Fasolt3.f90:299-323): 903 zero diagonal, 904 SOR not
 * converged, 905/906 rotation/strain increment too large, 910 singular LU,
 * 920 bad element weight, 955 row too long, 987 slip Newton residual > 0.5.
 * Codes below 0 are errors of this library (CPFEM_E_*).  The host keeps
 * the reference behaviour by calling its own FAFERR(code) on a nonzero
 * return.  Arrays follow the reference's Fortran layout (column-major per
 * point, e.g. st(1:6, gp) -> st[6*gp + i]); index arrays are 1-based.
 * The context is not re-entrant; calls are stream-ordered and return when
 * the results the host asked for are in host memory.
 */
#ifndef CPFEM_H
#define CPFEM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPFEM_MAXSS 60 /* fas-com-mod.f90:55 maxss */
#define CPFEM_MAXSV 63 /* fas-com-mod.f90:14 maxsvs */

#define CPFEM_E_ARG (-1)       /* bad argument or unsupported size */
#define CPFEM_E_HIP (-2)       /* HIP runtime error */
#define CPFEM_E_NSS (-3)       /* number of slip systems outside 1..CPFEM_MAXSS */
#define CPFEM_E_NOGPU (-4)     /* no gfx950 device */
#define CPFEM_E_SINGULAR (-6)  /* calibration: a reverse step met a singular Jacobian of the converged equations */
#define CPFEM_E_WIDTH (-5)     /* a point's active set is wider than the compiled kernels: 12 systems
                                  for materials of 12 or 18 systems, 24 for other counts */

/* One material set as held in the .fsd file (Faslt_io_nowin.f90:77-107). */
typedef struct {
    int32_t nss;
    int32_t twinlim;
    double twinprob;
    double e[9];                              /* S11 S22 S33 S23 S31 S12 S44 S55 S66 */
    double ssang[CPFEM_MAXSS];                /* 0 slip, 1 unidirectional, |.|>1.5 twin (deg) */
    double ab[CPFEM_MAXSS][3];                /* ab(1:3,i) slip direction */
    double an[CPFEM_MAXSS][3];                /* an(1:3,i) plane normal */
    double ampar[CPFEM_MAXSS][10];            /* ampar(1:10,i) */
    double smatrx[CPFEM_MAXSS][CPFEM_MAXSS];  /* smatrx[i][j] == Fortran smatrx(j,i) */
} cpfem_material;

/* fafcdf.dat (Fasolt3.f90:216-218) plus the solver settings of FACSLS. */
typedef struct {
    double toler;
    int32_t maxit;
    int32_t sor_maxit;    /* Stif_sol.f90:173 maxitl */
    double alpha, afactor;
    double efix, smalp, smalf;
    double smald, smale, smalr;
    double sor_omega;     /* Stif_sol.f90:173 alphal; in (0, 2), or <= 0: chosen by cpfem_autotune at the first solve */
    double sor_tol;       /* tolerance of the SOR stopping test */
    int32_t sor_kind;     /* 0: |dx|^2/|x|^2 < tol^2 (Stif_sol.f90:208-210); 1: |D^-1(b-Ax)|/|D^-1 b| < tol, D = diag(A) */
    int32_t precision;    /* 0: FP64 throughout.  1: FP32 element products and matrix, FP32 SOR sweeps on the
                             correction equation, residuals accumulated in FP64 from the FP32 matrix (the
                             solve converges to the FP32-rounded system's solution).  2: FP64 element
                             products, matrix and residuals; the SOR sweeps run in FP32 on a rounded copy of
                             the matrix (iterative refinement of the FP64 system: the solve meets the same
                             FP64 residual test as precision 0).  1 and 2 need sor_kind 1. */
    double stall_tol;     /* 0: CPSTRS runs to maxit as the reference does (Cpstrs.f90:139).  > 0: a point
                             also stops when an outer iteration moves its stress by at most stall_tol
                             relative (max norm) while the Newton correction stays above toler: the stress
                             has reached a fixed point of SCALE_IT's cap and the Newton step, which later
                             iterations only move at the rounding level (results then differ from the
                             reference's maxit-th iterate by about stall_tol / 10) */
} cpfem_control;

/*
 * Per-point constitutive update on a batch of independent points.
 * Replaces DSTRSS(iel, jgp, deff, es, de, st, ds, t, r, chi)
 * (src/fasolt/Sig_eps.f90:84), which the reference calls once per point
 * from ELVALS (src/fasolt/Elvals.f90:113), together with the module state
 * it reads and writes: sv/dsv(:,jgp), srate(:,jgp), twinsys(jgp), tflag and
 * COMMON /CPITER/ sumcpiter.
 *   iel[n]            element material code (negative: elastic only)
 *   de[n][6], r[n][3] strain increment (eng. shear) and rotation increment
 *   st[n][6]          stress at the start of the increment
 *   ds[n][6]          in: stress-increment guess, out: stress increment
 *   sv[n][ldsv]       state: Euler angles 1-3, slip resistances 4..3+nss
 *   dsv[n][ldsv]      in/out: state increments
 *   t[n][6]           out: updated stress
 *   chi[n][36]        out: tangent d(sigma)/d(eps), Fortran chi(6,6) order
 *   deff[n]           out: effective plastic strain increment
 *   srate[n][60]      in/out: slip rates (unchanged on the elastic branch)
 *   twinsys[n]        in/out; timestwinned[n] in
 *   tflag, sumcpiter  in/out accumulators (max / sum)
 *   flops[n], err[n]  optional (may be NULL): per-point algorithmic FLOPs and codes
 * Host pointers.  Returns the largest error code over the batch.
 */
int cpfem_dstrss_batch(const cpfem_material *mat, const cpfem_control *ctl, double dt, int32_t n,
                       const int32_t *iel, const double *de, const double *r, const double *st,
                       double *ds, const double *sv, double *dsv, int32_t ldsv,
                       double *t, double *chi, double *deff, double *srate,
                       int32_t *twinsys, const int32_t *timestwinned,
                       int32_t *tflag, double *sumcpiter, int64_t *flops, int32_t *err);


/*
 * cpfem_dstrss_batch without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and error codes as cpfem_dstrss_batch (host pointers),
 * but the staging, the field-major point arrays, the material record and the
 * pass schedule live in one file-scope device array of 40 MiB per GPU, held
 * until the process ends; batches of more than 8192 points are rejected with
 * CPFEM_E_ARG.  It calls no hipMalloc/hipFree and no allocating routine.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * race.
 */
int cpfem_dstrss_batch_noalloc(const cpfem_material *mat, const cpfem_control *ctl, double dt, int32_t n,
                               const int32_t *iel, const double *de, const double *r, const double *st,
                               double *ds, const double *sv, double *dsv, int32_t ldsv,
                               double *t, double *chi, double *deff, double *srate,
                               int32_t *twinsys, const int32_t *timestwinned,
                               int32_t *tflag, double *sumcpiter, int64_t *flops, int32_t *err);

/* ---------------------------------------------------------------------- */
/* Device-resident problem: the NRITER call sequence (Stif_sol.f90:10-88)  */
/* ---------------------------------------------------------------------- */

typedef struct cpfem_ctx cpfem_ctx;

/* Mesh, boundary codes and load tables as read from the .fsd file
 * (Faslt_io_nowin.f90:29-128).  Element types: 12 (8-node brick, one point,
 * SHAPE_12 Shapef.f90:32), 13 (8-node brick, 2x2x2 points), 21 (10-node
 * tetrahedron, 4 points, SHAPE_21 Shapef.f90:104), 22 (20-node brick,
 * SHAPE_22 Shapef.f90:210), 23 (15-node pentahedron, 6 points, SHAPE_23
 * Shapef.f90:360) and the two-node ties 1-6 (Elvals.f90:62-77).
 * nsv >= 3 + nss state variables per point (at most 63); those past 3 + nss
 * are carried unchanged. */
typedef struct {
    int32_t nlmnt, nnod, ngps, nsv;
    const int32_t *le, *lt, *nd, *np;
    const int32_t *ln;       /* [nlmnt][20] ln(1:20,el), 1-based */
    const int32_t *lg;       /* [nlmnt][8]  lg(1:8,el), 1-based */
    const int32_t *ib;       /* [nnod] boundary codes ccbbaa (Bconds.f90:8-12) */
    int32_t ninc1, nxs, nfs;
    const double *tiem;      /* [ninc1] time step of each increment */
    const double *x_spec;    /* [ninc1][99][15] kinematic tables */
    const double *f_spec;    /* [ninc1][99][3] load tables */
} cpfem_mesh;

/* Report values and timings of the last calls. */
typedef struct {
    int32_t nit;              /* NR iterations of the last cpfem_nriter */
    int32_t sor_sweeps;       /* sweeps of the last FACSLS */
    int32_t sor_sweeps_total; /* sweeps summed over the last cpfem_nriter */
    int32_t err;              /* largest FAFERR code seen */
    double sdx, sfc, sumcpiter, alp2, sor_error;
    double t_const_ms;        /* point kernels (kinematics + constitutive) */
    double t_asm_ms;          /* element assembly */
    double t_suvals_ms;       /* boundary conditions */
    double t_sor_ms;          /* SOR solve */
    double t_update_ms;       /* norms, damping, update */
    double t_total_ms;        /* whole call */
    int64_t flops_const;      /* algorithmic FLOPs of the constitutive updates */
    int64_t transc_const;     /* their transcendental calls */
    int64_t gp_updates;       /* points updated (sum over NR iterations) */
    int64_t pts_continued;    /* point updates that needed more than the first pass's outer iterations */
    int64_t pts_wide;         /* point updates whose active set exceeded the register-resident width */
    double t_sweep_ms;        /* device time of the sweep-kernel batches alone (events around each batch) */
    int64_t sweep_launches;   /* colour-sweep kernel launches in those batches */
    double sor_omega;         /* relaxation factor of the last FACSLS (the tuned one when set by cpfem_autotune) */
    double t_tune_ms;         /* time spent in cpfem_autotune during the call (automatic omega) */
    int32_t sor_wpb;          /* waves per workgroup of the colour-sweep launches */
    int32_t tune_sweeps;      /* sweeps run by cpfem_autotune during the call */
    int64_t pts_987;          /* point updates whose slip Newton ended above the 0.5 residual (FAFERR 987,
                                 crys-cons.f90:108), summed over the ELVALS calls */
    int32_t pts_987_last;     /* such points in the last ELVALS (for NRITER: its final iteration) */
    int32_t graph_launches;   /* graph mode: graph launches of the last cpfem_nriter (iteration graphs and their
                                 continuations) */
    double t_const_pass_ms[4];/* device time of the constitutive passes: every point's first outer iterations,
                                 the continuation of the unfinished ones, the 8-wide and the 12-wide (or wider)
                                 active-set tiers (call-by-call iterations only) */
    double t_graph_ms;        /* graph mode: device time of the Newton iterations run as graphs (the per-phase
                                 times above cover the call-by-call iterations only) */
} cpfem_stats;

/* Upload mesh, materials and control; build the block sparsity, colourings
 * and element scatter map.  Replaces SETPTA (Stif_sol.f90:98), which the
 * reference calls once after RDDATA (Fasolt3.f90:24).  mats[0..nmat-1] are the
 * .fsd material sets; a point uses set |le| - 1 of its element (Sig_eps.f90:98). */
int cpfem_create(const cpfem_mesh *mesh, const cpfem_material *mats, int32_t nmat,
                 const cpfem_control *ctl, cpfem_ctx **out);
int cpfem_destroy(cpfem_ctx *ctx);

/* Upload / download the state of the .fsd/.frs records (Faslt_io_nowin.f90:45-65,
 * 207-240): nodal x, dx (3*nnod) and the point arrays in the reference layout
 * (eps, dep [ngps]; sv, dsv [ngps][ldsv]; ep, de, st, ds [ngps][6]; rt, dr [ngps][3]). */
int cpfem_set_state(cpfem_ctx *ctx, const double *x, const double *dx, const double *eps, const double *dep,
                    const double *sv, const double *dsv, int32_t ldsv, const double *ep, const double *de,
                    const double *st, const double *ds, const double *rt, const double *dr,
                    const int32_t *twinsys, const int32_t *timestwinned);
int cpfem_get_state(cpfem_ctx *ctx, double *x, double *dx, double *fc, double *eps, double *dep, double *sv,
                    double *dsv, int32_t ldsv, double *ep, double *de, double *st, double *ds, double *rt,
                    double *dr, double *srate, double *cumslip, int32_t *twinsys, int32_t *timestwinned);

/* REAL(4) binding: the reference keeps its state in single-precision module
 * arrays (FAS_COM / FAS_CPL, fas-com-mod.f90:13-80).  These entry points take
 * those arrays as they are, so a Fortran host passes its module variables
 * with no copies of its own; the library widens them to its FP64 state and
 * narrows results with round-to-nearest.  Layouts as in the module: x(3*nnod),
 * sv(ldsv, ngps) with ldsv = maxsvs = 63, st(6, ngps), srate(60, ngps), ... */
#define CPFEM_MAXMDS 12 /* fas-com-mod.f90:57 maxmds */
typedef struct {
    int32_t nlmnt, nnod, ngps, nsv;
    const int32_t *le, *lt, *nd, *np;
    const int32_t *ln;       /* ln(1:20, nlmnt), 1-based */
    const int32_t *lg;       /* lg(1:8, nlmnt), 1-based */
    const int32_t *ib;       /* ib(nnod) */
    int32_t ninc1, nxs, nfs;
    const float *tiem;       /* tiem(ninc1) */
    const float *x_spec;     /* x_spec(15, 99, ninc1) */
    const float *f_spec;     /* f_spec(3, 99, ninc1) */
} cpfem_mesh_r4;
/* The FAS_CPL material arrays of sets 1..nmatl, leading dimensions maxss = 60:
 * nss(k), e(9,k), ssang(60,k), ab/an(3,60,k), ampar(10,60,k), smatrx(60,60,k);
 * twinlim and twinprob are the module's scalars (one value for every set). */
typedef struct {
    int32_t nmatl;
    int32_t twinlim;
    float twinprob;
    const int32_t *nss;
    const float *e, *ssang, *ab, *an, *ampar, *smatrx;
} cpfem_materials_r4;
/* cpfem_create on the module arrays (RDDATA's, Faslt_io_nowin.f90:29-128). */
int cpfem_create_r4(const cpfem_mesh_r4 *mesh, const cpfem_materials_r4 *mats, const cpfem_control *ctl,
                    cpfem_ctx **out);
/* cpfem_set_state / cpfem_get_state on REAL(4) arrays (the .fsd / .frs point
 * and node records, Faslt_io_nowin.f90:45-65,207-240).  get: entries the
 * library does not own (sv past nsv, srate/cumslip past nss) are left as they are. */
int cpfem_set_state_r4(cpfem_ctx *ctx, const float *x, const float *dx, const float *eps, const float *dep,
                       const float *sv, const float *dsv, int32_t ldsv, const float *ep, const float *de,
                       const float *st, const float *ds, const float *rt, const float *dr,
                       const int32_t *twinsys, const int32_t *timestwinned);
int cpfem_get_state_r4(cpfem_ctx *ctx, float *x, float *dx, float *fc, float *eps, float *dep, float *sv,
                       float *dsv, int32_t ldsv, float *ep, float *de, float *st, float *ds, float *rt,
                       float *dr, float *srate, float *cumslip, int32_t *twinsys, int32_t *timestwinned);

/* UPDATE_NODE (Fasolt3.f90:231): x += dx. */
int cpfem_update_node(cpfem_ctx *ctx);
/* fc = 0, ss = 0 and ELVALS (Stif_sol.f90:38-42, Elvals.f90:11): every point's
 * constitutive update, internal forces and the stiffness matrix. */
int cpfem_elvals(cpfem_ctx *ctx, int32_t inc, double *sumcpiter, int32_t *tflag);
/* SUVALS (Bconds.f90:19) and the free-node force norm of Stif_sol.f90:46-51. */
int cpfem_suvals(cpfem_ctx *ctx, int32_t inc, double *sfc);
/* FACSLS (Stif_sol.f90:161): solve for ddx.  A stopping measure that is not
 * finite (NaN or Inf in the system, or diverging sweeps) ends the solve at
 * the next check with 904 and a NaN *error; cpfem_nriter then ends without
 * applying that correction. */
int cpfem_facsls(cpfem_ctx *ctx, int32_t *sweeps, double *error);
/* Norms, damping and update of Stif_sol.f90:55-74: x -= alp2*ddx, dx -= alp2*ddx. */
int cpfem_nr_update(cpfem_ctx *ctx, double *sdx, double *alp2);
/* The whole of NRITER for increment inc (Stif_sol.f90:10-88), cumulative slip included. */
int cpfem_nriter(cpfem_ctx *ctx, int32_t inc, cpfem_stats *stats);
/* Graph mode of cpfem_nriter and cpfem_increment (on != 0; needs sor_kind 1,
 * else CPFEM_E_ARG).  Replaces the host loop of NRITER's iteration body
 * (Stif_sol.f90:32-76): each Newton iteration runs as one launch of a graph
 * captured once per context (ELVALS, SUVALS, the solve's start, its first
 * sweep batches with their residual checks, the damped update), continued by
 * replays of a second graph of sweep batches when the solve needs more; every
 * stopping decision of the solve is taken on the device.  Same kernels in the
 * same order as the call-by-call path: the iterates are the same.  The host
 * reads one record per iteration (plus one per continuation beyond the
 * predicted number).  Per-phase times are not taken in graph mode
 * (cpfem_stats.t_graph_ms holds the iterations' device time). */
int cpfem_set_graph_mode(cpfem_ctx *ctx, int32_t on);
/* Cumulative slip at the end of NRITER (Stif_sol.f90:80-84): cumslip += srate*tiem(inc). */
int cpfem_cumslip(cpfem_ctx *ctx, int32_t inc);
/* Power of the POWER report (Faslt_io_nowin.f90:353-366): sum(fc*dx). */
int cpfem_power(cpfem_ctx *ctx, double *fdx);
/* UPDATE_GPTV (Fasolt3.f90:254): commit the point increments. */
int cpfem_update_gptv(cpfem_ctx *ctx);
/* DoReorient (Fasolt3.f90:79) if the twin flag is set, then tflag = 0
 * (Fasolt3.f90:43-46); returns the count in *n. */
int cpfem_reorient(cpfem_ctx *ctx, int32_t *n);
/* Volume averages over all points at the current coordinates (the macroscopic
 * stress-strain state that CSVMKR's element means feed, data_out.f90:113-235):
 * out[0] = volume, out[1..6] = mean stress, out[7..12] = mean total strain ep
 * (engineering shears), out[13] = mean effective strain eps.  Weights are the
 * points' integration weights (Gauss weight x detJ). */
int cpfem_volume_average(cpfem_ctx *ctx, double *out);
/* Set the twin flag, the module variable tflag (fas-com-mod.f90:17) that SLPINI
 * raises (crys-cons.f90:171), RDDATA clears (Faslt_io_nowin.f90:133) and the
 * driver tests before DoReorient (Fasolt3.f90:43). */
int cpfem_set_tflag(cpfem_ctx *ctx, int32_t tflag);
/* Stream of the twin draws.  SLPINI draws RANDOM_NUMBER afresh at every twin
 * check with twinprob < 1 (crys-cons.f90:165-173), from the compiler
 * runtime's stream, which cannot be reproduced.  Here the draw of twin system
 * i in outer CPSTRS iteration k of point g is a counter-based hash of
 * (epoch, g, k, i): every check is an independent uniform number.  Each
 * ELVALS of a context takes the next epoch of the context's counter, each
 * cpfem_dstrss_batch call the next of a process-wide counter.  This sets the
 * epoch the next such call uses (ctx NULL: the batch counter), for repeatable
 * runs; with twinprob = 1 no draw is consulted. */
int cpfem_twin_stream(cpfem_ctx *ctx, uint64_t next_epoch);
/* One full increment: UPDATE_NODE, NRITER, UPDATE_GPTV, DoReorient (Fasolt3.f90:38-46). */
int cpfem_increment(cpfem_ctx *ctx, int32_t inc, cpfem_stats *stats);
/* Right-hand side and solution of the last solve (3*nnod each, may be NULL). */
int cpfem_get_system(cpfem_ctx *ctx, double *rhs, double *ddx);
/* Assembled matrix as COO in DOF numbering (0-based); call with NULL arrays to get *nnz. */
int cpfem_get_matrix(cpfem_ctx *ctx, int64_t *nnz, int32_t *rows, int32_t *cols, double *vals);
int cpfem_get_stats(cpfem_ctx *ctx, cpfem_stats *stats);

/* Tuning of FACSLS's sweeps on the system assembled by the last cpfem_suvals.
 * No counterpart in the reference, whose FACSLS is a fixed loop with omega
 * compiled in (Stif_sol.f90:173); for callers that solve many systems of one
 * mesh.  With CPFEM_TUNE_LAUNCH the colour-sweep kernel is timed at 1, 2 and 4
 * waves per workgroup and the fastest kept (the iterates do not depend on it).
 * With CPFEM_TUNE_OMEGA each relaxation factor of a ladder in [1, 1.98] runs
 * two windows of sweeps from a zero start, the number of sweeps to the solve's
 * tolerance is extrapolated from the measured residuals, and the factor with
 * the fewest is kept.  The system and ddx are left as they were; the sweep
 * graphs are rebuilt.  omega and waves_per_block (either may be NULL) return
 * the settings in force.  With flags 0 nothing is tuned and no system is
 * needed: a positive *omega (< 2) and a *waves_per_block of 1, 2 or 4 are
 * applied as given (0 keeps the setting).  A context created with
 * sor_omega <= 0 runs both tunings at its first solve and afterwards moves
 * omega by a coordinate search on the measured convergence rate of each
 * residual-rule solve (a trial factor best +- step on alternate solves, kept
 * when it converges faster per sweep; step 0.01 halving to 0.0025, omega in
 * [1, 1.995]); setting omega through this call ends that.
 * With CPFEM_TUNE_ORDER the order in which a sweep relaxes the node colours is
 * chosen: each candidate (the order in force; with 8 colours also the six
 * binary-reflected Gray codes of the colour numbers, which on a brick mesh
 * numbered x fastest are the parity classes visited one axis flip at a time)
 * runs two windows of sweeps at the current omega and the one with the fewest
 * extrapolated sweeps is kept.  The order changes the iterates, not the
 * solution the solve's test accepts.  The automatic tuning of a context
 * created with sor_omega <= 0 leaves the order alone: its first system is the
 * elastic one, on which the orders differ by a few percent, while after yield
 * the default order (below) needs 20-30 % fewer sweeps than the colour-number
 * order. */
#define CPFEM_TUNE_LAUNCH 1
#define CPFEM_TUNE_OMEGA 2
#define CPFEM_TUNE_ORDER 4
int cpfem_autotune(cpfem_ctx *ctx, int32_t flags, double *omega, int32_t *waves_per_block);

/* Order of the node colours within a sweep.  No counterpart in the reference
 * (its FACSLS sweeps the DOFs in natural order, Stif_sol.f90:187-198).  The
 * default is the colour-number order, except with 8 colours: a reflected Gray
 * code of the colour numbers (see CPFEM_TUNE_ORDER).  With
 * set == 0 the order in force is written to order[0..ncolor-1]; with set != 0
 * order[0..n-1] (a permutation of 0..ncolor-1, n == ncolor) is applied.
 * Returns ncolor, or a negative code (n too small to hold the order, or not a
 * permutation). */
int cpfem_sweep_order(cpfem_ctx *ctx, int32_t *order, int32_t n, int32_t set);

/* Launch shapes of the point kernels.  No counterpart in the reference.  The
 * constitutive kernel runs every point for a first budget of outer CPSTRS
 * iterations, then resumes the unfinished points in later passes sorted by
 * their work; the 8-node assembly packs 1, 2 or 4 elements (one wave each)
 * per workgroup.  Neither changes a result.  With inc > 0 both are timed on
 * ELVALS of increment inc on the current state (budgets 4, 8, 16, 32; the
 * state ELVALS overwrites is restored, the matrix is not: call ELVALS again
 * before the next solve) and the fastest are kept; with inc == 0 the given
 * values (> 0; elements per workgroup 1, 2 or 4) are applied.  Both return
 * the settings in force (defaults 8 and 4). */
int cpfem_autotune_points(cpfem_ctx *ctx, int32_t inc, int32_t *first_iters, int32_t *elems_per_block);

/* Calibration: replay of a recorded load path and the adjoint of its
 * constitutive updates.  No counterpart in the reference, which has no
 * derivatives; for fitting hardening and rate parameters to a measured
 * macroscopic stress-strain curve.
 *
 * cpfem_calib_begin starts a record of ninc increments from the current state
 * (stress and angles kept as the start of the path).  One material set only,
 * without twin systems, with the same ampar(1:6) on every system and a latent
 * matrix of one diagonal and one off-diagonal value (else CPFEM_E_ARG), and at
 * most 24 systems.  toler and maxit are the CPSTRS/SLPFTA tolerance and
 * iteration limit of the replay: the gradient is that of the converged
 * equations, so the replay should converge far below the reference's toler
 * (1e-10 for a 1e-6 finite-difference check); the replay also stops points
 * whose stress is stationary (stall_tol 1e-14 of cpfem_control), which
 * still end above toler and are excluded.  ckpt_every: increments per
 * checkpoint of the reverse sweep, whose segments except the last are run
 * forward again from their checkpoints; ckpt_every <= 0 keeps the whole
 * trajectory from the forward pass (ninc * (15 + 2 nss) doubles per point)
 * and recomputes nothing.
 *
 * cpfem_calib_record(inc) keeps, for every point, the inputs of the last
 * DSTRSS call of increment inc (de, dr, dsv(1:3), ds) and its integration
 * weight; call it after cpfem_nriter and before cpfem_update_gptv.
 *
 * cpfem_calib_run replays the path with params = {tau0, taus, h0, q, m}
 * (tau0: every initial slip resistance; taus = ampar(5), h0 = ampar(2),
 * q = off-diagonal latent ratio, m = ampar(1) of every system): curve[inc-1]
 * = weight-averaged stress component comp (0..5: 11 22 33 23 31 12) after
 * increment inc, loss = sum (curve - target)^2, and with grad != NULL the
 * derivative of the loss with respect to the five parameters.  Returns 0,
 * CPFEM_E_SINGULAR (grad unreliable), or another error code. */
typedef struct {
    int32_t points, increments;
    int32_t unconverged;     /* replayed point updates stopped at maxit (forward pass) */
    int32_t singular;        /* reverse steps with a singular Jacobian */
    int32_t elastic;         /* reverse steps on the elastic branch */
    int32_t equilibrium_solves; /* solves of the equilibrium correction */
    double t_forward_ms;     /* forward replay */
    double t_reverse_ms;     /* segment recomputation and reverse steps */
    double solve_sweeps;     /* SOR sweeps of the equilibrium correction, summed */
} cpfem_calib_stats;
int cpfem_calib_begin(cpfem_ctx *ctx, int32_t ninc, int32_t ckpt_every, double toler, int32_t maxit);
int cpfem_calib_record(cpfem_ctx *ctx, int32_t inc);
/* Recorded inputs of increment inc per point, rec[ngps][19]: de(6), dr(3),
 * dsv(1:3) (6..), ds(6) (12..), curve weight (18; zero for excluded points). */
int cpfem_calib_history(cpfem_ctx *ctx, int32_t inc, double *rec);
int cpfem_calib_run(cpfem_ctx *ctx, const double *params, const double *target, int32_t comp, double *curve,
                    double *loss, double *grad, cpfem_calib_stats *stats);
/* Points whose replay at params stops at maxit in some increment (CPSTRS
 * held at a fixed point of SCALE_IT's cap, Cpstrs.f90:59,75,187-230, which
 * solves no constitutive equation and has no derivative of the converged
 * equations) get zero weight in every increment of the record: the curve is
 * then the average over the other points and its gradient is exact.
 * Exclusions accumulate over calls (e.g. at each point of a finite-difference
 * stencil); *excluded returns the number of points excluded so far. */
int cpfem_calib_exclude(cpfem_ctx *ctx, const double *params, int32_t *excluded);
/* Options of the record, set after cpfem_calib_begin and before the first
 * cpfem_calib_record.  CPFEM_CALIB_EQUILIBRIUM: also keep every increment's
 * coordinates and Newton matrix (FP64 contexts with 8-node bricks only), and
 * add to the gradient the response of the converged displacement increments:
 * per increment of the reverse sweep, one solve with the recorded matrix
 * (the symmetric stiffness NRITER uses, Stif_sol.f90:42-53, standing in for
 * the exact Jacobian of the discrete equilibrium) and a second reverse step
 * with the corrected stress cotangents.  Memory: 3 nnod + the matrix values
 * per increment.  Flags 0 switches the correction off (the record is kept);
 * it can be switched on again on a record made with it. */
#define CPFEM_CALIB_EQUILIBRIUM 1
int cpfem_calib_options(cpfem_ctx *ctx, int32_t flags);
/* Change the checkpoint spacing of the reverse sweep (as ckpt_every of
 * cpfem_calib_begin); the record is kept. */
int cpfem_calib_checkpoints(cpfem_ctx *ctx, int32_t ckpt_every);
int cpfem_calib_end(cpfem_ctx *ctx);

/* Device name and compute units of the current device (diagnostics). */
int cpfem_device_info(char *name, int32_t len, int32_t *cus);

#ifdef __cplusplus
}
#endif
#endif /* CPFEM_H */
