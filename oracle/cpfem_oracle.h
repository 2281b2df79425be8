/*
 * CPU oracle of FASOLT3's per-increment hot path (test infrastructure only).
 *
 * This is a plain-C restatement of the reference Fortran routines, used by the
 * test-suite, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg as
 * the checker.  The product path never links or calls it.
 *
 * The same source is compiled twice: with -DORC_REAL=float (liboracle32.so,
 * operation order of the REAL(4) reference) and -DORC_REAL=double
 * (liboracle64.so, the FP64 oracle the HIP kernels are held to).  Every
 * boundary array below is `double` regardless of build; the float build
 * rounds on entry, so float-representable inputs pass through exactly.
 *
 * Layouts follow the reference modules (src/fasolt/fas-com-mod.f90:9-80):
 * Fortran column-major per point, e.g. st(1:6, gp) -> st[6*gp + i].
 * Index arrays (ln, lg) are 1-based as in the reference.
 */
#ifndef CPFEM_ORACLE_H
#define CPFEM_ORACLE_H
#include <stdint.h>

#define ORC_MAXSS 60 /* fas-com-mod.f90:55 maxss */
#define ORC_MAXSV 63 /* fas-com-mod.f90:14 maxsvs */

/* One material set as stored in the .fsd file (Faslt_io_nowin.f90:77-107). */
typedef struct {
    int32_t nss;
    int32_t twinlim;
    double twinprob;
    double e[9];                        /* S11 S22 S33 S23 S31 S12 S44 S55 S66 */
    double ssang[ORC_MAXSS];            /* 0 slip, 1 unidirectional, |.|>1.5 twin (deg) */
    double ab[ORC_MAXSS][3];            /* ab(1:3,i): slip direction, crystal axes */
    double an[ORC_MAXSS][3];            /* an(1:3,i): plane normal, crystal axes */
    double ampar[ORC_MAXSS][10];        /* ampar(1:10,i) */
    double smatrx[ORC_MAXSS][ORC_MAXSS];/* smatrx[i][j] == Fortran smatrx(j,i) */
} orc_material;

/* fafcdf.dat (Fasolt3.f90:216-218) plus the SOR settings of FACSLS. */
typedef struct {
    double toler;
    int32_t maxit;
    int32_t sor_maxit;     /* Stif_sol.f90:173 maxitl (1000) */
    double alpha, afactor;
    double efix, smalp, smalf;
    double smald, smale, smalr;
    double sor_omega;      /* Stif_sol.f90:173 alphal (1.2) */
    double sor_tol;        /* tolerance of the stopping test; reference: toler */
    int32_t sor_kind;      /* 0: reference |dx|^2/|x|^2 < tol^2; 1: |D^-1(b-Ax)|/|D^-1 b| < tol, D = diag(A) */
    int32_t keep_going;    /* 0: NRITER stops on a fatal code, as the reference's FAFERR does;
                              1: it continues on 987, 910 and 904 and returns the largest code
                              seen, the policy of the device library */
    double stall_tol;      /* 0: CPSTRS runs to maxit (the reference).  > 0: the device library's
                              opt-in stop at a stationary stress (max-norm change of an outer
                              iteration <= stall_tol relative while erepmd > toler), restated so
                              the two follow the same iteration path */
} orc_control;

/* Algorithmic operation counters of one constitutive call. */
typedef struct {
    int64_t flops;      /* add/sub/mul/div/sqrt, FMA = 2 */
    int64_t transc;     /* exp, log, pow(=2), sin, cos */
    int32_t outer;      /* CPSTRS outer stress iterations */
    int32_t inner;      /* SLPFTA inner iterations, summed over outer */
    int32_t max_nsa;    /* largest active set seen */
    int32_t draws;      /* twin draws made (checks with twinprob consulted) */
    int32_t capped;     /* SCALE_IT scaled in the last outer iteration */
    int32_t pad_;
} orc_counters;

/* Module state of FASOLT3 that the hot path reads and writes. */
typedef struct {
    int32_t nlmnt, nnod, ngps, nsv;
    int32_t nmatl, inc, ninc1, nxs, nfs;
    int32_t maxofd;                 /* row capacity of is/ss */
    int32_t tflag, err;             /* global twin flag, max FAFERR code seen */
    int32_t mat_code;               /* last material code (module FAS_CPL) */
    const int32_t *le, *lt, *nd, *np;
    const int32_t *ln;              /* [nlmnt][20] 1-based */
    const int32_t *lg;              /* [nlmnt][8]  1-based */
    const int32_t *ib;              /* [nnod] */
    double *x, *dx, *fc, *ddx;      /* [3*nnod] */
    double *eps, *dep;              /* [ngps] */
    double *st, *ep, *ds, *de;      /* [ngps][6] */
    double *rt, *dr;                /* [ngps][3] */
    double *sv, *dsv;               /* [ngps][ORC_MAXSV] */
    double *srate, *cumslip;        /* [ngps][ORC_MAXSS] */
    int32_t *twinsys, *timestwinned;/* [ngps] */
    const double *tiem;             /* [ninc1] */
    const double *x_spec;           /* [ninc1][99][15] */
    const double *f_spec;           /* [ninc1][99][3] */
    const orc_material *mat;        /* [nmatl], mat_code 1-based */
    const orc_control *ctl;
    int32_t *is;                    /* [3*nnod][maxofd]: is(1,row)=count, then columns (1-based) */
    double *ss;                     /* [3*nnod][maxofd] */
    double sumcpiter;               /* COMMON /CPITER/ */
    int64_t flops, transc;          /* summed constitutive counters */
    int32_t sor_sweeps;             /* sweeps of the last FACSLS */
    double sor_error;               /* last stopping-test value */
    int32_t nit;                    /* NR iterations of the last NRITER */
    double sdx, sfc, alp2;          /* last NR report values */
    uint64_t rng_calls;             /* counter of the twin random draws */
    int32_t n987;                   /* points of the last ELVALS whose slip Newton ended with 987 */
    int32_t n987_capped;            /* ... of which SCALE_IT held in their last outer iteration */
    int32_t nmaxit;                 /* points of the last ELVALS that ran CPSTRS to maxit */
    int32_t pad_;
} orc_model;

#ifdef __cplusplus
extern "C" {
#endif
/* Per-point constitutive update on a batch of independent points. */
int orc_dstrss_batch(const orc_material *mat, int nmat, const orc_control *ctl,
                     double dt, int n, const int32_t *iel,
                     const double *de, const double *r, const double *st,
                     double *ds, double *sv, double *dsv, int ldsv,
                     double *t, double *chi, double *deff, double *srate,
                     int32_t *twinsys, const int32_t *timestwinned,
                     int32_t *tflag, double *sumcpiter, orc_counters *cnt,
                     int32_t *err);
/* Stream epoch of the next batch call's twin draws (each call advances it). */
void orc_set_batch_epoch(uint64_t epoch);
/* The twin draw of system sys (0-based) in outer iteration outer (1-based) of
 * the point stream key = (epoch << 32) ^ point. */
double orc_twin_draw(uint64_t key, int outer, int sys);
/* Building blocks exposed for closed-form tests. */
double orc_spower(double a, double b);
double orc_fj2vom(int lstress, const double *x);
int orc_factor(int n, double *w);
void orc_solver(int n, const double *w, double *x);
void orc_ctmatrx(double ph1, double phi, double ph2, double *a, double *aug);
void orc_elastc(const orc_material *m, double ph1, double phi, double ph2, double *el_comp);
void orc_tens_rot(const double *r, const double *a, double *b);
int orc_invert_3(double *m, double *det);
/* Shape functions of element type lt at point ig (0-based): lt 12, 13 take 8
 * nodes, lt 22 (SHAPE_22) 20; xnode[3*a+k], sfd[3*a+k] = dN_a/dx_k. */
void orc_shape(int lt, int ig, const double *xnode, double *sf, double *sfd, double *weight);
int orc_dstrns(int npl, const double *sfd, const double *dxl, double smale, double smalr,
               double *del, double *rl, double *deff);
/* Assembly, boundary conditions, solver and the increment driver. */
int orc_setpta(orc_model *m);
int orc_elvals(orc_model *m);
int orc_suvals(orc_model *m);
int orc_facsls(orc_model *m, int ndof);
int orc_nriter(orc_model *m);
void orc_update_node(orc_model *m);
void orc_update_gptv(orc_model *m);
int orc_doreorient(orc_model *m);
int orc_real_size(void);
#ifdef __cplusplus
}
#endif
#endif
