/*
 * CPU oracle of FASOLT3's per-increment hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference Fortran (gabortimar/CPFEM_code,
 * src/fasolt/), statement order kept so that the float build follows the
 * REAL(4) reference operation by operation.  Each routine cites the
 * reference lines it restates.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, always as the checker.
 *
 * Compile with -DORC_REAL=float or -DORC_REAL=double.  No -ffast-math and
 * no FMA contraction (-ffp-contract=off): the order of rounding is the point.
 *
 * Parity pins (what checks each routine; see tests/):
 *   SPOWER, FACTOR/SOLVER, FJ2VOM, CTMATRX, BTMATRX, ELASTC, TENS_ROT,
 *   INVERT_3, shape functions, DSTRNS: closed forms and invariants, plus the
 *   compiled reference through oracle/_ref where it builds.
 *   DSTRSS/CPSTRS/SLPFTA/PLASTC/SLPINI/SCALE_IT/ECHROT: fixed-point
 *   invariants (Schmid rate law, strain-increment balance, chi inverse)
 *   and the compiled reference's DSTRSS on seeded batches (tests/golden).
 *   ELVALS/SUVALS/FACSLS/NRITER: compiled reference on 1-GP bricks, symmetry,
 *   rigid-body null space, direct-solve agreement.
 *   Twin draws with twinprob < 1: the reference uses the compiler's
 *   RANDOM_NUMBER afresh at every check; its stream is not reproducible, so
 *   the draw here is the build's own counter-based one (one uniform number
 *   per check), pinned by what the reference's model fixes: uniformity and
 *   independence of the draws, and Wald's identity (twinned points =
 *   twinprob x checks) on large batches (tests/test_twin_draw.py).
 *   twinprob = 1 never consults the draw.
 */
#include "cpfem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#ifdef ORC_TRACE
#include <stdio.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
#define R(v) ((real)(v))

int orc_real_size(void) { return (int)sizeof(real); }

/* ------------------------------------------------------------------------ */
/* material and control converted to the working precision                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nss, twinlim;
    real twinprob;
    real e[9];
    real ssang[ORC_MAXSS];
    real ab[ORC_MAXSS][3], an[ORC_MAXSS][3];
    real ampar[ORC_MAXSS][10];
    real smatrx[ORC_MAXSS][ORC_MAXSS]; /* [i][j] = smatrx(j,i) */
} wmat;

typedef struct {
    real toler, alpha, afactor, efix, smalp, smalf, smald, smale, smalr;
    int maxit;
    real stall_tol;
} wctl;

static void load_mat(wmat *w, const orc_material *m) {
    w->nss = m->nss;
    w->twinlim = m->twinlim;
    w->twinprob = R(m->twinprob);
    for (int i = 0; i < 9; ++i) w->e[i] = R(m->e[i]);
    for (int i = 0; i < ORC_MAXSS; ++i) {
        w->ssang[i] = R(m->ssang[i]);
        for (int k = 0; k < 3; ++k) {
            w->ab[i][k] = R(m->ab[i][k]);
            w->an[i][k] = R(m->an[i][k]);
        }
        for (int k = 0; k < 10; ++k) w->ampar[i][k] = R(m->ampar[i][k]);
        for (int k = 0; k < ORC_MAXSS; ++k) w->smatrx[i][k] = R(m->smatrx[i][k]);
    }
}

static void load_ctl(wctl *w, const orc_control *c) {
    w->toler = R(c->toler);
    w->alpha = R(c->alpha);
    w->afactor = R(c->afactor);
    w->efix = R(c->efix);
    w->smalp = R(c->smalp);
    w->smalf = R(c->smalf);
    w->smald = R(c->smald);
    w->smale = R(c->smale);
    w->smalr = R(c->smalr);
    w->maxit = c->maxit;
    w->stall_tol = R(c->stall_tol);
}

/* Scratch of module FAS_CPL (fas-com-mod.f90:62-66) for one point. */
typedef struct {
    int nsa;
    int isa[ORC_MAXSS];                /* 0-based system indices */
    real tau[ORC_MAXSS], gam[ORC_MAXSS], wks[ORC_MAXSS], s0[ORC_MAXSS], s1[ORC_MAXSS];
    real b[ORC_MAXSS][6];              /* b(1:6,i) */
    real rot[ORC_MAXSS][3];            /* rot(1:3,i) */
    real ah[ORC_MAXSS * ORC_MAXSS];    /* ah(1:nsa,1:nsa) packed column-major, ld nsa */
    const wmat *m;
    const wctl *c;
    orc_counters *cnt;
    int err;
    int32_t *twinsys;
    int32_t timestwinned;
    int32_t *tflag;
    uint64_t rng_key;
    int outer;                         /* CPSTRS outer iteration (1-based), the twin draw's counter */
    int capped;                        /* SCALE_IT scaled in the current outer iteration */
} cpwork;

static void note_err(int *err, int code) {
    if (code > *err) *err = code;
}

static void count(orc_counters *c, int64_t f, int64_t t) {
    if (c) {
        c->flops += f;
        c->transc += t;
    }
}

static int64_t factor_flops(int n) { return (int64_t)n * (n - 1) / 2 + (int64_t)n * (n - 1) * (2 * n - 1) / 3; }
static int64_t solver_flops(int n) { return 2 * (int64_t)n * n - n; }

/* ------------------------------------------------------------------------ */
/* crys-cons.f90                                                            */
/* ------------------------------------------------------------------------ */

/* SPOWER, crys-cons.f90:210-230: signed power with clamps. */
static real spower(real a, real b) {
    const real allm = R(80.0), bllm = R(1.e-35);
    real aa = fabs(a);
    if (aa < bllm) return R(0.0);
    real sparg = b * log(aa);
    if (sparg < -allm) sparg = -allm;
    if (sparg > allm) sparg = allm;
    return copysign(exp(sparg), a);
}

/* FACTOR, crys-cons.f90:244-269: Doolittle LU without pivoting, in place.
 * w is n x n column-major with leading dimension n.  Returns ifail. */
static int factor(int n, real *w) {
#define W(i, j) w[(i) + (size_t)(j) * n]
    for (int k = 0; k < n - 1; ++k)
        for (int i = k + 1; i < n; ++i) {
            if (W(k, k) == R(0.0)) return 1;
            W(i, k) = W(i, k) / W(k, k);
            for (int j = k + 1; j < n; ++j) W(i, j) = W(i, j) - W(i, k) * W(k, j);
        }
    if (W(n - 1, n - 1) == R(0.0)) return 1;
    return 0;
#undef W
}

/* SOLVER, crys-cons.f90:283-309: forward / back substitution. */
static void solver(int n, const real *w, real *x) {
#define W(i, j) w[(i) + (size_t)(j) * n]
    for (int i = 1; i < n; ++i)
        for (int j = 0; j < i; ++j) x[i] = x[i] - W(i, j) * x[j];
    x[n - 1] = x[n - 1] / W(n - 1, n - 1);
    for (int i = n - 2; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) x[i] = x[i] - W(i, j) * x[j];
        x[i] = x[i] / W(i, i);
    }
#undef W
}

/* ------------------------------------------------------------------------ */
/* Sig_eps.f90                                                              */
/* ------------------------------------------------------------------------ */

/* FJ2VOM, Sig_eps.f90:217-241. */
static real fj2vom(int lstress, const real *x) {
    real amult = lstress ? R(6.0) : R(1.5);
    real sum = R(0.0);
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        sum = sum + (x[i] - x[j]) * (x[i] - x[j]) + amult * x[i + 3] * x[i + 3];
    }
    return sqrt(sum / R(4.5));
}

/* TENS_ROT, Sig_eps.f90:166-202: new = rot^T . old . rot. */
static void tens_rot(const real *r, const real *a, real *b) {
    real rot[3][3], old[3][3], tmp[3][3], nw[3][3]; /* [row][col] */
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (j + 1) % 3;
        old[i][i] = a[i];
        old[j][k] = a[i + 3];
        old[k][j] = old[j][k];
        rot[j][k] = r[i];
        rot[k][j] = -r[i];
        rot[i][i] = sqrt(R(1.0) - r[j] * r[j] - r[k] * r[k]);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = R(0.0);
            for (int k = 0; k < 3; ++k) s = s + old[i][k] * rot[k][j];
            tmp[i][j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = R(0.0);
            for (int k = 0; k < 3; ++k) s = s + rot[k][i] * tmp[k][j];
            nw[i][j] = s;
        }
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (j + 1) % 3;
        b[i] = nw[i][i];
        b[i + 3] = nw[j][k];
    }
}

/* ------------------------------------------------------------------------ */
/* crys_math.f90                                                            */
/* ------------------------------------------------------------------------ */

/* CTMATRX, crys_math.f90:25-62: a[row][col] = a(row,col), Bunge angles. */
static void ctmatrx(real ph1, real phi, real ph2, real a[3][3], real aug[3]) {
    real c1 = cos(ph1), s1 = sin(ph1), c2 = cos(ph2), s2 = sin(ph2);
    real c = cos(phi), s = sin(phi);
    real c1c2 = c1 * c2, s1s2 = s1 * s2, c1s2 = c1 * s2, s1c2 = s1 * c2;
    a[0][0] = c1c2 - s1s2 * c;
    a[1][0] = -c1s2 - s1c2 * c;
    a[2][0] = s1 * s;
    a[0][1] = s1c2 + c1s2 * c;
    a[1][1] = -s1s2 + c1c2 * c;
    a[2][1] = -c1 * s;
    a[0][2] = s2 * s;
    a[1][2] = c2 * s;
    a[2][2] = c;
    aug[0] = s2;
    aug[1] = c2;
    aug[2] = s;
}

/* BTMATRX, crys_math.f90:75-115: Schmid (eng. shear) and spin columns. */
static void btmatrx(cpwork *w, real a[3][3]) {
    const wmat *m = w->m;
    for (int i = 0; i < m->nss; ++i) {
        real sdr[3], spn[3];
        for (int j = 0; j < 3; ++j) {
            real sum1 = R(0.0), sum2 = R(0.0);
            for (int l = 0; l < 3; ++l) {
                sum1 = sum1 + a[l][j] * m->ab[i][l];
                sum2 = sum2 + a[l][j] * m->an[i][l];
            }
            sdr[j] = sum1;
            spn[j] = sum2;
        }
        w->b[i][0] = sdr[0] * spn[0];
        w->b[i][1] = sdr[1] * spn[1];
        w->b[i][2] = sdr[2] * spn[2];
        w->b[i][3] = sdr[1] * spn[2] + sdr[2] * spn[1];
        w->b[i][4] = sdr[2] * spn[0] + sdr[0] * spn[2];
        w->b[i][5] = sdr[0] * spn[1] + sdr[1] * spn[0];
        w->rot[i][0] = (sdr[1] * spn[2] - sdr[2] * spn[1]) / R(2.0);
        w->rot[i][1] = (sdr[2] * spn[0] - sdr[0] * spn[2]) / R(2.0);
        w->rot[i][2] = (sdr[0] * spn[1] - sdr[1] * spn[0]) / R(2.0);
    }
}

/* ECHROT, crys_math.f90:137-167: Euler-angle increments from the lattice spin. */
static void echrot(cpwork *w, real a[3][3], const real aug[3], const real *r, real *dp1, real *dph,
                   real *dp2) {
    const real small = R(1.e-5);
    real rtn[3], rx[3];
    for (int i = 0; i < 3; ++i) {
        real s = R(0.0);
        for (int k = 0; k < w->m->nss; ++k) s = s + w->rot[k][i] * w->gam[k];
        rtn[i] = s;
    }
    for (int i = 0; i < 3; ++i) rtn[i] = rtn[i] - r[i];
    for (int i = 0; i < 3; ++i) {
        real s = R(0.0);
        for (int k = 0; k < 3; ++k) s = s + a[i][k] * rtn[k];
        rx[i] = s;
    }
    if (fabs(aug[2]) < small)
        *dp1 = rtn[2];
    else
        *dp1 = (aug[0] * rx[0] + aug[1] * rx[1]) / aug[2];
    *dph = aug[1] * rx[0] - aug[0] * rx[1];
    *dp2 = rx[2] - *dp1 * a[2][2];
}

/* ------------------------------------------------------------------------ */
/* ela-plas.f90                                                             */
/* ------------------------------------------------------------------------ */

/* ELASTC, ela-plas.f90:12-86: compliance in specimen axes (eng. shear). */
static void elastc(const wmat *m, real ph1, real phi, real ph2, real el[6][6], orc_counters *cnt) {
    real a0[3][3], aug[3], a[3][3];
    ctmatrx(ph1, phi, ph2, a0, aug);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = a0[j][i];
    const real *e = m->e;
    for (int iset = 0; iset < 6; ++iset) {
        int i, j;
        real mult;
        if (iset < 3) {
            i = j = iset;
            mult = R(1.0);
        } else {
            i = (iset - 2) % 3; /* 1+MOD(iset-3,3) in 1-based terms */
            j = (i + 1) % 3;
            mult = R(2.0);
        }
        for (int jset = iset; jset < 6; ++jset) {
            int k, l;
            if (jset < 3) {
                k = l = jset;
            } else {
                k = (jset - 2) % 3;
                l = (k + 1) % 3;
            }
            real term1 = a[i][0] * a[j][0] * a[k][0] * a[l][0] * e[0] +
                         a[i][1] * a[j][1] * a[k][1] * a[l][1] * e[1] +
                         a[i][2] * a[j][2] * a[k][2] * a[l][2] * e[2];
            real term2 = a[i][1] * a[j][1] * a[k][2] * a[l][2] * e[3] +
                         a[i][2] * a[j][2] * a[k][1] * a[l][1] * e[3] +
                         a[i][2] * a[j][2] * a[k][0] * a[l][0] * e[4] +
                         a[i][0] * a[j][0] * a[k][2] * a[l][2] * e[4] +
                         a[i][0] * a[j][0] * a[k][1] * a[l][1] * e[5] +
                         a[i][1] * a[j][1] * a[k][0] * a[l][0] * e[5];
            real term3 = a[i][1] * a[j][2] * a[k][1] * a[l][2] * e[6] +
                         a[i][2] * a[j][1] * a[k][2] * a[l][1] * e[6] +
                         a[i][2] * a[j][0] * a[k][2] * a[l][0] * e[7] +
                         a[i][0] * a[j][2] * a[k][0] * a[l][2] * e[7] +
                         a[i][0] * a[j][1] * a[k][0] * a[l][1] * e[8] +
                         a[i][1] * a[j][0] * a[k][1] * a[l][0] * e[8];
            el[iset][jset] = mult * (term1 + term2 + R(0.5) * term3);
        }
    }
    for (int iset = 0; iset < 5; ++iset)
        for (int jset = iset + 1; jset < 6; ++jset) el[jset][iset] = el[iset][jset];
    count(cnt, 1575 + 12, 6);
}

/* PLASTC, ela-plas.f90:110-201: hardening, s1 and the log-hardening matrix
 * ah(i,j) on the active set (packed nsa x nsa, column-major). */
static void plastc(cpwork *w, real dt, real *am) {
    const wmat *m = w->m;
    const int ns = m->nss;
    const real tiny = R(1.e-14);
    real grad_h[ORC_MAXSS];
    for (int i = 0; i < ns; ++i) am[i] = m->ampar[i][0];
    real sumss = R(0.0);
    for (int iss = 0; iss < ns; ++iss) sumss = sumss + fabs(w->gam[iss]);
    for (int iss = 0; iss < ns; ++iss) {
        real sst = m->ampar[iss][4] * pow(sumss / dt, m->ampar[iss][5]);
        real arg;
        if (sst != R(0.0))
            arg = R(1.0) - ((w->s0[iss]) / sst);
        else
            arg = R(0.0);
        if (arg < R(0.0)) arg = R(0.0);
        /* THETA statement function, ela-plas.f90:127-128 */
        grad_h[iss] = m->ampar[iss][2] + m->ampar[iss][1] * copysign(pow(fabs(arg), m->ampar[iss][3]), arg);
    }
    /* gradient(k,i) = smatrx(k,i) * grad_h(i) */
#define GRAD(k, i) (m->smatrx[(i)][(k)] * grad_h[(i)])
    for (int iss = 0; iss < ns; ++iss) {
        real s = w->s0[iss];
        for (int jss = 0; jss < ns; ++jss) s = s + GRAD(iss, jss) * fabs(w->gam[jss]);
        w->s1[iss] = s;
    }
    const int nsa = w->nsa;
    for (int i = 0; i < nsa; ++i) {
        int iss = w->isa[i];
        for (int j = 0; j < nsa; ++j) {
            int jss = w->isa[j];
            real v;
            if (fabs(w->gam[jss]) > tiny) {
                real gam_sign = w->gam[jss] / fabs(w->gam[jss]);
                v = gam_sign * GRAD(iss, jss) / w->s1[iss];
            } else {
                v = R(0.0);
            }
            w->ah[i + (size_t)j * nsa] = v;
        }
    }
#undef GRAD
    count(w->cnt, 3 * (int64_t)ns * ns + 6 * ns + 3 * (int64_t)nsa * nsa, 4 * ns);
}

/* ------------------------------------------------------------------------ */
/* crys-cons.f90 slip routines                                              */
/* ------------------------------------------------------------------------ */

/* Draw for the twin test of system sys (0-based) in outer iteration outer
 * (1-based) of a point with stream key key = (call epoch << 32) ^ point.  The
 * reference calls RANDOM_NUMBER afresh at every check (crys-cons.f90:168-173),
 * a compiler-runtime stream that cannot be reproduced; this counter-based draw
 * is the build's own, shared with the HIP kernel: the splitmix64 finaliser of
 * the key, offset by the check's counter times the golden-ratio increment,
 * mixed again, so every check is an independent uniform draw.  Only matters
 * when twinprob < 1. */
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double twin_draw(uint64_t key, int outer, int sys) {
    const uint64_t h = mix64(key + 0x9E3779B97F4A7C15ull);
    const uint64_t z = mix64(h + ((uint64_t)outer * 64ull + (uint64_t)sys + 1ull) * 0x9E3779B97F4A7C15ull);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* SLPINI, crys-cons.f90:143-197: predictor, twin flagging, active set. */
static void slpini(cpwork *w, real dt, real *am) {
    const wmat *m = w->m;
    const int ns = m->nss;
    const real cofact = R(1.e-6);
    for (int i = 0; i < ns; ++i) am[i] = m->ampar[i][0];
    real gammax = R(0.0);
    int nspow = 0;
    for (int i = 0; i < ns; ++i) {
        if (fabs(m->ssang[i]) > R(1.5)) {
            w->gam[i] = R(0.0);
            if ((w->tau[i] > w->s0[i]) && (w->timestwinned != m->twinlim) && (*w->twinsys == 0)) {
                real twinrn = R(twin_draw(w->rng_key, w->outer, i));
                if (w->cnt) w->cnt->draws += 1;
                if (twinrn < m->twinprob) {
                    *w->tflag = 1;
                    *w->twinsys = i + 1;
                }
            }
        } else if (m->ssang[i] * w->tau[i] / w->s0[i] < R(0.0)) {
            w->gam[i] = R(0.0);
        } else {
            w->gam[i] = dt * spower((w->tau[i] / w->s0[i]), (R(1.0) / am[i]));
            ++nspow;
        }
        if (fabs(w->gam[i]) > gammax) gammax = fabs(w->gam[i]);
    }
    real cutoff = cofact * gammax;
    int nsa = 0;
    for (int i = 0; i < ns; ++i) {
        if (fabs(w->gam[i]) < cutoff)
            w->gam[i] = R(0.0);
        else
            w->isa[nsa++] = i;
    }
    w->nsa = nsa;
    if (w->cnt && nsa > w->cnt->max_nsa) w->cnt->max_nsa = nsa;
    (void)nspow;
    count(w->cnt, 5 * (int64_t)ns, 2 * (int64_t)ns);
}

/* SCALE_IT, Cpstrs.f90:187-230: cap the largest slip at gamcap. */
static void scale_it(cpwork *w, const real *am, real *s, real gamcap, real *scale) {
    const wmat *m = w->m;
    real amax = am[0];
    real gammax = fabs(w->gam[w->isa[0]]);
    for (int i = 1; i < w->nsa; ++i) {
        int ii = w->isa[i];
        if (fabs(w->gam[ii]) > gammax) {
            gammax = fabs(w->gam[ii]);
            amax = am[ii];
        }
    }
    if (gammax > gamcap) {
        w->capped = 1;
        *scale = gamcap / gammax;
        real smult = spower(R(0.9) * *scale, w->c->afactor * amax);
        if (smult == R(0.0)) smult = R(0.004) / amax;
        for (int i = 0; i < 6; ++i) s[i] = smult * s[i];
        for (int i = 0; i < m->nss; ++i) {
            w->gam[i] = *scale * w->gam[i];
            w->tau[i] = smult * w->tau[i];
        }
        count(w->cnt, w->nsa + 2 * (int64_t)m->nss + 6, 2);
    } else {
        *scale = R(1.0);
        count(w->cnt, w->nsa, 0);
    }
}

/* SLPFTA, crys-cons.f90:24-128: Newton on ln(gam) over the active set,
 * leaving ah in the X (visco-plastic) form. */
static int slpfta(cpwork *w, real dt) {
    const wmat *m = w->m;
    const int ns = m->nss;
    const real facmax = R(5.0), tiny = R(1.e-12);
    real am[ORC_MAXSS], errtau[ORC_MAXSS];
    real amx = m->ampar[0][0];
    for (int i = 1; i < ns; ++i)
        if (m->ampar[i][0] > amx) amx = m->ampar[i][0];
    real conv = w->c->toler * amx;
    int iter = 0, l_iter = 1;
    real tauerr = R(0.0);
    const int nsa = w->nsa;
    while (l_iter) {
        ++iter;
        if (w->cnt) w->cnt->inner += 1;
        plastc(w, dt, am);
        for (int i = 0; i < nsa; ++i) {
            int ii = w->isa[i];
            real gamn = w->gam[ii];
            for (int j = 0; j < nsa; ++j) w->ah[j + (size_t)i * nsa] = gamn * w->ah[j + (size_t)i * nsa];
            w->ah[i + (size_t)i * nsa] = w->ah[i + (size_t)i * nsa] + am[ii];
        }
        if (factor(nsa, w->ah)) note_err(&w->err, 910);
        for (int i = 0; i < nsa; ++i) {
            int ii = w->isa[i];
            real taufg = w->s1[ii] * spower((w->gam[ii] / dt), am[ii]);
            if ((w->tau[ii] * taufg) > R(0.0))
                errtau[i] = log(w->tau[ii] / taufg);
            else
                errtau[i] = R(0.0);
        }
        tauerr = R(0.0);
        for (int i = 0; i < nsa; ++i) tauerr = tauerr + errtau[i] * errtau[i];
        tauerr = sqrt(tauerr);
        solver(nsa, w->ah, errtau);
        real errmax = R(0.0);
        for (int i = 0; i < nsa; ++i)
            if (fabs(errtau[i]) > errmax) errmax = fabs(errtau[i]);
        real scale = (errmax > facmax) ? facmax / errmax : R(1.0);
        for (int i = 0; i < nsa; ++i) {
            int ii = w->isa[i];
            w->gam[ii] = w->gam[ii] * exp(scale * errtau[i]);
        }
        l_iter = (iter < w->c->maxit) && (tauerr > conv);
        count(w->cnt, (int64_t)nsa * nsa + 10 * (int64_t)nsa + 1 + factor_flops(nsa) + solver_flops(nsa),
              4 * (int64_t)nsa);
    }
    if (tauerr > R(0.5)) note_err(&w->err, 987);
    plastc(w, dt, am);
    for (int i = 0; i < nsa; ++i) {
        int ii = w->isa[i];
        for (int j = 0; j < nsa; ++j) {
            int jj = w->isa[j];
            w->ah[j + (size_t)i * nsa] = w->tau[jj] * w->ah[j + (size_t)i * nsa];
        }
        if (fabs(w->gam[ii]) > tiny)
            w->ah[i + (size_t)i * nsa] = w->ah[i + (size_t)i * nsa] + am[ii] * w->tau[ii] / w->gam[ii];
        else
            w->ah[i + (size_t)i * nsa] = w->ah[i + (size_t)i * nsa] + am[ii] / tiny;
    }
    count(w->cnt, (int64_t)nsa * nsa + 3 * (int64_t)nsa, 0);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Cpstrs.f90                                                               */
/* ------------------------------------------------------------------------ */

/* CPSTRS, Cpstrs.f90:34-172: outer stress Newton for one point.
 * sv/dsv: this point's state-variable columns.  Returns erepmd. */
static real cpstrs(cpwork *w, real dt, real el_comp[6][6], const real *del, const real *rl, real *s,
                   const real *t, real *epl, real chi[6][6], real *sv, real *dsv, real *srate) {
    const wmat *m = w->m;
    const int ns = m->nss;
    const real capfac = R(6.0);
    const real alpha_cps = w->c->alpha, tolercp = w->c->toler;
    const int maxcpsit = w->c->maxit;
    real a[3][3], aug[3], am[ORC_MAXSS], erreps[6], djac[36];
    real ph1 = sv[0] + dsv[0];
    real phi = sv[1] + dsv[1];
    real ph2 = sv[2] + dsv[2];
    for (int i = 0; i < ns; ++i) {
        w->s0[i] = sv[i + 3];
        w->s1[i] = w->s0[i] + dsv[i + 3];
    }
    real gamcap = capfac * fj2vom(0, del);
    ctmatrx(ph1, phi, ph2, a, aug);
    btmatrx(w, a);
    count(w->cnt, 14 + 1 + 12 + 60 * (int64_t)ns, 6);
    int iter = 0, l_iter = 1;
    real erepmd = R(0.0), scale;
    while (l_iter) {
        ++iter;
        w->outer = iter;
        w->capped = 0;
        if (w->cnt) w->cnt->outer += 1;
        real s_top[6];
        for (int i = 0; i < 6; ++i) s_top[i] = s[i];
        for (int i = 0; i < ns; ++i) {
            real sum = R(0.0);
            for (int k = 0; k < 6; ++k) sum = sum + w->b[i][k] * s[k];
            w->tau[i] = sum;
        }
        slpini(w, dt, am);
        scale_it(w, am, s, gamcap, &scale);
        slpfta(w, dt);
        scale_it(w, am, s, gamcap, &scale);
        for (int i = 0; i < 6; ++i) {
            real sum = R(0.0);
            for (int k = 0; k < ns; ++k) sum = sum + w->b[k][i] * w->gam[k];
            epl[i] = sum;
        }
        {
            real sm[6];
            for (int k = 0; k < 6; ++k) sm[k] = s[k] - t[k];
            for (int i = 0; i < 6; ++i) {
                real sum = R(0.0);
                for (int k = 0; k < 6; ++k) sum = sum + el_comp[i][k] * sm[k];
                erreps[i] = del[i] - epl[i] - sum;
            }
        }
        const int nsa = w->nsa;
        if (factor(nsa, w->ah)) note_err(&w->err, 910);
        /* djac = B_a X^-1 B_a^T.  The reference re-solves column j once per
         * row i (Cpstrs.f90:103-117); the solve depends on j only, so one
         * solve per column gives the same bits. */
        real wcol[6][ORC_MAXSS];
        for (int j = 0; j < 6; ++j) {
            for (int k = 0; k < nsa; ++k) wcol[j][k] = w->b[w->isa[k]][j];
            solver(nsa, w->ah, wcol[j]);
        }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                real sum = R(0.0);
                for (int k = 0; k < nsa; ++k) sum = sum + w->b[w->isa[k]][i] * wcol[j][k];
                djac[i + 6 * j] = sum;
            }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) djac[i + 6 * j] = djac[i + 6 * j] + el_comp[i][j];
        if (factor(6, djac)) note_err(&w->err, 910);
        solver(6, djac, erreps);
        erepmd = fj2vom(1, erreps) / fj2vom(1, s);
        scale = R(1.0) / (R(1.0) + alpha_cps * erepmd);
        for (int i = 0; i < 6; ++i) erreps[i] = scale * erreps[i];
        for (int i = 0; i < 6; ++i) s[i] = s[i] + erreps[i];
        l_iter = (iter < maxcpsit) && (erepmd > tolercp);
        if (l_iter && w->c->stall_tol > R(0.0)) {
            real dmax = R(0.0), smax = R(0.0);
            for (int i = 0; i < 6; ++i) {
                if (fabs(s[i] - s_top[i]) > dmax) dmax = fabs(s[i] - s_top[i]);
                if (fabs(s[i]) > smax) smax = fabs(s[i]);
            }
            if (dmax <= w->c->stall_tol * smax) l_iter = 0;
        }
#ifdef ORC_TRACE
        fprintf(stderr, "outer %d nsa %d erepmd %.3e scale %.3e\n", iter, nsa, (double)erepmd, (double)scale);
#endif
        count(w->cnt,
              12 * (int64_t)ns + 12 * (int64_t)ns + 90 + 72 * (int64_t)nsa + 36 + 44 + factor_flops(nsa) +
                  6 * solver_flops(nsa) + factor_flops(6) + solver_flops(6) + 28,
              0);
    }
    for (int i = 0; i < ns; ++i) srate[i] = fabs(w->gam[i] / dt);
    real dp1, dph, dp2;
    echrot(w, a, aug, rl, &dp1, &dph, &dp2);
    for (int i = 0; i < 6; ++i) {
        for (int k = 0; k < 6; ++k) erreps[k] = R(0.0);
        erreps[i] = R(1.0);
        solver(6, djac, erreps);
        for (int j = 0; j < 6; ++j) chi[j][i] = erreps[j];
    }
    dsv[0] = dp1;
    dsv[1] = dph;
    dsv[2] = dp2;
    for (int i = 0; i < ns; ++i) dsv[i + 3] = w->s1[i] - w->s0[i];
    count(w->cnt, 6 * (int64_t)ns + 25 + 6 * solver_flops(6), 0);
    return erepmd;
}

/* DSTRSS, Sig_eps.f90:84-151: one point.  Returns the FAFERR code (0 ok). */
static int dstrss(const wmat *mats, const wctl *c, real dt, int iel, real *deff, const real *de,
                  const real *st, real *ds, real *t, const real *r, real chi[6][6], real *sv, real *dsv,
                  real *srate, int32_t *twinsys, int32_t timestwinned, int32_t *tflag, double *sumcpiter,
                  orc_counters *cnt, uint64_t rng_key) {
    int mat_code = abs(iel);
    const wmat *m = &mats[mat_code - 1];
    real el_comp[6][6], s[6], ep[6];
    elastc(m, sv[0] + dsv[0], sv[1] + dsv[1], sv[2] + dsv[2], el_comp, cnt);
    tens_rot(r, st, t);
    for (int i = 0; i < 6; ++i) s[i] = t[i] + ds[i];
    real sinit = fj2vom(1, s);
    real einit = fj2vom(0, de);
    count(cnt, 120 + 28, 0);
    int err = 0;
    if (iel > 0 && sinit > R(0.0) && einit > R(0.0)) {
        cpwork w;
        memset(&w, 0, sizeof(w));
        w.m = m;
        w.c = c;
        w.cnt = cnt;
        w.twinsys = twinsys;
        w.timestwinned = timestwinned;
        w.tflag = tflag;
        w.rng_key = rng_key;
        real erepmd = cpstrs(&w, dt, el_comp, de, r, s, t, ep, chi, sv, dsv, srate);
        *sumcpiter = (double)(real)(R(*sumcpiter) + erepmd);
        for (int i = 0; i < 6; ++i) t[i] = s[i];
        *deff = fj2vom(0, ep);
        for (int i = 0; i < 6; ++i)
            for (int j = i; j < 6; ++j) {
                real d = (chi[i][j] + chi[j][i]) / R(2.0);
                chi[i][j] = d;
                chi[j][i] = d;
            }
        err = w.err;
        if (cnt) cnt->capped = w.capped;
        count(cnt, 82, 0);
    } else {
        *deff = R(0.0);
        real lu[36];
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) lu[i + 6 * j] = el_comp[i][j];
        (void)factor(6, lu); /* Sig_eps.f90:134 ignores ifail */
        for (int i = 0; i < 6; ++i) {
            for (int k = 0; k < 6; ++k) ep[k] = R(0.0);
            ep[i] = R(1.0);
            solver(6, lu, ep);
            for (int j = 0; j < 6; ++j) chi[j][i] = ep[j];
        }
        for (int i = 0; i < 6; ++i) {
            real sum = R(0.0);
            for (int k = 0; k < 6; ++k) sum = sum + chi[i][k] * de[k];
            t[i] = sum;
        }
        count(cnt, factor_flops(6) + 6 * solver_flops(6) + 72 + 12, 0);
    }
    for (int i = 0; i < 6; ++i) ds[i] = t[i] - st[i];
    return err;
}

/* Stream epoch of the next orc_dstrss_batch call's twin draws (advanced per
 * call, as the library's batch entry point does). */
static uint64_t g_batch_epoch = 1;
void orc_set_batch_epoch(uint64_t epoch) { g_batch_epoch = epoch; }
double orc_twin_draw(uint64_t key, int outer, int sys) { return twin_draw(key, outer, sys); }

int orc_dstrss_batch(const orc_material *mat, int nmat, const orc_control *ctl, double dt, int n,
                     const int32_t *iel, const double *de, const double *r, const double *st, double *ds,
                     double *sv, double *dsv, int ldsv, double *t, double *chi, double *deff,
                     double *srate, int32_t *twinsys, const int32_t *timestwinned, int32_t *tflag,
                     double *sumcpiter, orc_counters *cnt, int32_t *err) {
    wmat *wm = (wmat *)malloc(sizeof(wmat) * (size_t)nmat);
    if (!wm) return -1;
    for (int i = 0; i < nmat; ++i) load_mat(&wm[i], &mat[i]);
    wctl wc;
    load_ctl(&wc, ctl);
    int worst = 0;
    const uint64_t epoch = g_batch_epoch++;
    for (int g = 0; g < n; ++g) {
        real lde[6], lr[3], lst[6], lds[6], lt[6], lchi[6][6], lsv[ORC_MAXSV], lvd[ORC_MAXSV], lsr[ORC_MAXSS];
        real ldeff = R(0.0);
        int ns = wm[abs(iel[g]) - 1].nss;
        for (int i = 0; i < 6; ++i) {
            lde[i] = R(de[6 * g + i]);
            lst[i] = R(st[6 * g + i]);
            lds[i] = R(ds[6 * g + i]);
        }
        for (int i = 0; i < 3; ++i) lr[i] = R(r[3 * g + i]);
        for (int i = 0; i < 3 + ns; ++i) {
            lsv[i] = R(sv[(size_t)ldsv * g + i]);
            lvd[i] = R(dsv[(size_t)ldsv * g + i]);
        }
        for (int i = 0; i < ns; ++i) lsr[i] = R(srate[(size_t)ORC_MAXSS * g + i]);
        orc_counters *c = cnt ? &cnt[g] : NULL;
        if (c) memset(c, 0, sizeof(*c));
        int e = dstrss(wm, &wc, R(dt), iel[g], &ldeff, lde, lst, lds, lt, lr, lchi, lsv, lvd, lsr,
                       &twinsys[g], timestwinned[g], tflag, sumcpiter, c, (epoch << 32) ^ (uint64_t)g);
        if (err) err[g] = e;
        if (e > worst) worst = e;
        for (int i = 0; i < 6; ++i) {
            ds[6 * g + i] = lds[i];
            t[6 * g + i] = lt[i];
        }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) chi[36 * (size_t)g + i + 6 * j] = lchi[i][j];
        deff[g] = ldeff;
        for (int i = 0; i < 3 + ns; ++i) dsv[(size_t)ldsv * g + i] = lvd[i];
        for (int i = 0; i < ns; ++i) srate[(size_t)ORC_MAXSS * g + i] = lsr[i];
    }
    free(wm);
    return worst;
}

/* ------------------------------------------------------------------------ */
/* Shapef.f90 and DSTRNS                                                    */
/* ------------------------------------------------------------------------ */

/* INVERT_3, Shapef.f90:499-536: adjugate inverse, matrix[col*3+row]. */
static void invert_3(real *mtx, real *det) {
    real loc[9];
#define M(i, j) mtx[(i) + 3 * (j)]
#define L(i, j) loc[(i) + 3 * (j)]
    memcpy(loc, mtx, sizeof(loc));
    for (int i = 0; i < 3; ++i) {
        int ia = (i + 1) % 3, ib = (ia + 1) % 3;
        for (int j = 0; j < 3; ++j) {
            int ja = (j + 1) % 3, jb = (ja + 1) % 3;
            M(j, i) = L(ia, ja) * L(ib, jb) - L(ia, jb) * L(ib, ja);
        }
    }
    *det = M(0, 0) * L(0, 0) + M(0, 1) * L(1, 0) + M(0, 2) * L(2, 0);
    if (*det != R(0.0))
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M(i, j) = M(i, j) / *det;
#undef M
#undef L
}

/* Local-coordinate signs of the 8 brick nodes (fs3pre.f90:193-196 order);
 * the 2x2x2 Gauss points take the same order (Shapef.f90:203-205). */
static const int brick_sign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

/* Shape functions of the 8-node brick.  lt = 12: one point at the centre,
 * weight 8 detJ (SHAPE_12, Shapef.f90:32-87).  lt = 13: the build's
 * 2x2x2-point brick, points at +-1/sqrt(3) in the SHAPE_22 order, weight
 * detJ (not in the reference).  sfd[3*a + k] = dN_a/dx_k. */
static void shape8(int lt, int ig, const real *xnode, real *sf, real *sfd, real *weight) {
    real xi[3], dndn[8][3];
    if (lt == 12) {
        xi[0] = xi[1] = xi[2] = R(0.0);
    } else {
        real g = R(1.0) / sqrt(R(3.0));
        for (int k = 0; k < 3; ++k) xi[k] = R(brick_sign[ig][k]) * g;
    }
    for (int a = 0; a < 8; ++a) {
        real f[3], d[3];
        for (int k = 0; k < 3; ++k) {
            f[k] = R(1.0) + R(brick_sign[a][k]) * xi[k];
            d[k] = R(brick_sign[a][k]);
        }
        sf[a] = R(0.125) * f[0] * f[1] * f[2];
        dndn[a][0] = R(0.125) * d[0] * f[1] * f[2];
        dndn[a][1] = R(0.125) * f[0] * d[1] * f[2];
        dndn[a][2] = R(0.125) * f[0] * f[1] * d[2];
    }
    real jac[9]; /* jac(j,k) = sum_a dndn(j,a) x(k,a), [j + 3k] */
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            real sum = R(0.0);
            for (int a = 0; a < 8; ++a) sum = sum + dndn[a][j] * xnode[3 * a + k];
            jac[j + 3 * k] = sum;
        }
    real det;
    invert_3(jac, &det);
    for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k) {
            real sum = R(0.0);
            for (int j = 0; j < 3; ++j) sum = sum + jac[k + 3 * j] * dndn[a][j];
            sfd[3 * a + k] = sum;
        }
    *weight = (lt == 12) ? R(8.0) * det : det;
}

/* Local coordinates of the 20 nodes of SHAPE_22 (Shapef.f90:191-200): bottom
 * face corners 1,3,5,7 and mid-sides 2,4,6,8 anticlockwise seen from the top,
 * nodes 9-12 halfway up the vertical edges above 1,3,5,7, top face 13-20 like
 * the bottom. */
static const int brick20_loc[20][3] = {
    {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}, {-1, 1, -1}, {-1, 0, -1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},   {-1, 1, 0},
    {-1, -1, 1},  {0, -1, 1},  {1, -1, 1},  {1, 0, 1},  {1, 1, 1},  {0, 1, 1},  {-1, 1, 1},  {-1, 0, 1}};

/* SHAPE_22 tabulates the shape functions and their local derivatives at the
 * 2x2x2 points to 6 decimals (Shapef.f90:227-308).  The tables are the
 * quadratic serendipity functions evaluated at +-g, g = 1/sqrt(3) held in
 * single precision, each value rounded to 6 decimals: that rule reproduces
 * every tabulated entry, so the oracle generates them instead of storing
 * them.  Points in the brick_sign order. */
static int tab22_ready = 0;
static double tab22_n[8][20], tab22_d[8][3][20];

static double round6(double v) { return round(v * 1.0e6) / 1.0e6; }

static void brick20_tables(void) {
    if (tab22_ready) return;
    const double g = (double)(float)(1.0 / sqrt(3.0));
    for (int ig = 0; ig < 8; ++ig) {
        double xi[3];
        for (int k = 0; k < 3; ++k) xi[k] = brick_sign[ig][k] * g;
        for (int a = 0; a < 20; ++a) {
            const int *c = brick20_loc[a];
            double n, d[3];
            if (c[0] != 0 && c[1] != 0 && c[2] != 0) {
                /* corner: (1+xi c)(1+eta c)(1+zeta c)(xi c + eta c + zeta c - 2)/8 */
                double f[3], sum = -2.0;
                for (int k = 0; k < 3; ++k) {
                    f[k] = 1.0 + xi[k] * c[k];
                    sum += xi[k] * c[k];
                }
                n = f[0] * f[1] * f[2] * sum / 8.0;
                for (int k = 0; k < 3; ++k) {
                    const double other = f[(k + 1) % 3] * f[(k + 2) % 3];
                    d[k] = c[k] * other * sum / 8.0 + f[0] * f[1] * f[2] * c[k] / 8.0;
                }
            } else {
                /* mid-side node with c[z] = 0: (1-xi_z^2)(1+xi_p c_p)(1+xi_q c_q)/4 */
                int z = 0;
                while (c[z] != 0) ++z;
                const int p = (z + 1) % 3, q = (z + 2) % 3;
                const double qz = 1.0 - xi[z] * xi[z], fp = 1.0 + xi[p] * c[p], fq = 1.0 + xi[q] * c[q];
                n = qz * fp * fq / 4.0;
                d[z] = -2.0 * xi[z] * fp * fq / 4.0;
                d[p] = qz * c[p] * fq / 4.0;
                d[q] = qz * fp * c[q] / 4.0;
            }
            tab22_n[ig][a] = round6(n);
            for (int k = 0; k < 3; ++k) tab22_d[ig][k][a] = round6(d[k]);
        }
    }
    tab22_ready = 1;
}

/* SHAPE_22, Shapef.f90:210-342: 20-node brick, 2x2x2 points, weight detJ. */
static void shape20(int ig, const real *xnode, real *sf, real *sfd, real *weight) {
    brick20_tables();
    real dndn[20][3];
    for (int a = 0; a < 20; ++a) {
        sf[a] = R(tab22_n[ig][a]);
        for (int j = 0; j < 3; ++j) dndn[a][j] = R(tab22_d[ig][j][a]);
    }
    real jac[9]; /* jac(j,k) = sum_a dndn(j,a) x(k,a), [j + 3k] */
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            real sum = R(0.0);
            for (int a = 0; a < 20; ++a) sum = sum + dndn[a][j] * xnode[3 * a + k];
            jac[j + 3 * k] = sum;
        }
    real det;
    invert_3(jac, &det);
    for (int a = 0; a < 20; ++a)
        for (int k = 0; k < 3; ++k) {
            real sum = R(0.0);
            for (int j = 0; j < 3; ++j) sum = sum + jac[k + 3 * j] * dndn[a][j];
            sfd[3 * a + k] = sum;
        }
    *weight = R(1.0) * det;
}

/* SHAPE_21 (Shapef.f90:104-189) and SHAPE_23 (:360-485) also tabulate their
 * functions and local derivatives to 6 decimals.  Generated here:
 *  - 10-node tetrahedron: volume coordinates of the corners 1,3,5,10 with
 *    local axes xi1 = L10, xi2 = L3, xi3 = L1 and L5 = 1 - xi1 - xi2 - xi3;
 *    corner N = L(2L - 1), mid-edge N = 4 La Lb (2: 1-3, 4: 3-5, 6: 1-5,
 *    7: 1-10, 8: 3-10, 9: 5-10); point k has L = a at the k-th corner (order
 *    1,3,5,10) and b at the others, a = (5 + 3 sqrt 5)/20, b = (5 - sqrt 5)/20.
 *    Rounded to 6 decimals this reproduces every tabulated entry.
 *  - 15-node pentahedron: area coordinates L1 = xi1, L3 = xi2, L5 = 1 - xi1 -
 *    xi2 of the base corners 1,3,5 (top corners 10,12,14 above them) and
 *    zeta = xi3; corner N = L((2L - 1)(1 + zeta zc) - (1 - zeta^2))/2,
 *    triangle-edge N = 2 La Lb (1 + zeta zc), vertical-edge N = L (1 - zeta^2);
 *    points 1-3 at zeta = -1/sqrt 3 nearest corners 1, 3, 5 (area coordinates
 *    2/3, 1/6, 1/6), points 4-6 at +1/sqrt 3.  This reproduces the tabulated
 *    functions; 60 of the 270 tabulated derivatives differ from it by 1e-6 or
 *    2e-6 (the table's own rounding noise), so on 15-node elements the
 *    restatement agrees with the reference to that precision, not bit for bit. */
static int tabq_ready = 0;
static double tab21_n[4][10], tab21_d[4][3][10], tab23_n[6][15], tab23_d[6][3][15];

static void tet10_tables(void) {
    static const int corner_pos[4] = {0, 2, 4, 9};                 /* nodes 1,3,5,10 */
    static const int mids[6][3] = {{1, 0, 1}, {3, 1, 2}, {5, 0, 2}, {6, 0, 3}, {7, 1, 3}, {8, 2, 3}};
    /* d L_c / d xi for corners (1,3,5,10) */
    static const double dl[4][3] = {{0, 0, 1}, {0, 1, 0}, {-1, -1, -1}, {1, 0, 0}};
    const double a = (5.0 + 3.0 * sqrt(5.0)) / 20.0, b = (5.0 - sqrt(5.0)) / 20.0;
    for (int ig = 0; ig < 4; ++ig) {
        double L[4];
        for (int c = 0; c < 4; ++c) L[c] = (c == ig) ? a : b;
        for (int c = 0; c < 4; ++c) {
            const int node = corner_pos[c];
            tab21_n[ig][node] = round6(L[c] * (2.0 * L[c] - 1.0));
            for (int j = 0; j < 3; ++j) tab21_d[ig][j][node] = round6((4.0 * L[c] - 1.0) * dl[c][j]);
        }
        for (int m = 0; m < 6; ++m) {
            const int node = mids[m][0], p = mids[m][1], q = mids[m][2];
            tab21_n[ig][node] = round6(4.0 * L[p] * L[q]);
            for (int j = 0; j < 3; ++j) tab21_d[ig][j][node] = round6(4.0 * (L[q] * dl[p][j] + L[p] * dl[q][j]));
        }
    }
}

static void wedge15_tables(void) {
    /* node -> (kind, corner a, corner b, zc): kind 0 corner, 1 triangle edge, 2 vertical edge;
     * corners 0,1,2 = base corners 1,3,5 (area coordinates L1, L3, L5) */
    static const int nodes[15][4] = {{0, 0, 0, -1}, {1, 0, 1, -1}, {0, 1, 1, -1}, {1, 1, 2, -1}, {0, 2, 2, -1},
                                     {1, 2, 0, -1}, {2, 0, 0, 0},   {2, 1, 1, 0},   {2, 2, 2, 0},   {0, 0, 0, 1},
                                     {1, 0, 1, 1},  {0, 1, 1, 1},   {1, 1, 2, 1},   {0, 2, 2, 1},   {1, 2, 0, 1}};
    static const double dl[3][2] = {{1, 0}, {0, 1}, {-1, -1}};
    const double z0 = 1.0 / sqrt(3.0);
    for (int ig = 0; ig < 6; ++ig) {
        double L[3];
        for (int c = 0; c < 3; ++c) L[c] = (c == ig % 3) ? 2.0 / 3.0 : 1.0 / 6.0;
        const double z = ig < 3 ? -z0 : z0;
        for (int a = 0; a < 15; ++a) {
            const int kind = nodes[a][0], p = nodes[a][1], q = nodes[a][2];
            const double zc = nodes[a][3];
            double n, d[3];
            if (kind == 0) {
                const double l = L[p], f = (2.0 * l - 1.0) * (1.0 + z * zc) - (1.0 - z * z);
                n = 0.5 * l * f;
                const double dldL = 0.5 * ((4.0 * l - 1.0) * (1.0 + z * zc) - (1.0 - z * z));
                d[0] = dldL * dl[p][0];
                d[1] = dldL * dl[p][1];
                d[2] = 0.5 * l * ((2.0 * l - 1.0) * zc + 2.0 * z);
            } else if (kind == 1) {
                const double f = 1.0 + z * zc;
                n = 2.0 * L[p] * L[q] * f;
                for (int j = 0; j < 2; ++j) d[j] = 2.0 * f * (L[q] * dl[p][j] + L[p] * dl[q][j]);
                d[2] = 2.0 * L[p] * L[q] * zc;
            } else {
                n = L[p] * (1.0 - z * z);
                for (int j = 0; j < 2; ++j) d[j] = (1.0 - z * z) * dl[p][j];
                d[2] = -2.0 * z * L[p];
            }
            tab23_n[ig][a] = round6(n);
            for (int j = 0; j < 3; ++j) tab23_d[ig][j][a] = round6(d[j]);
        }
    }
}

/* The Jacobian, its inverse and the global derivatives of a tabulated
 * element at one point (the loops of Shapef.f90:163-187 and 459-483);
 * returns detJ. */
static real shape_tab(int nn, const double *tn, const double *td, const real *xnode, real *sf, real *sfd) {
    real dndn[20][3];
    for (int a = 0; a < nn; ++a) {
        sf[a] = R(tn[a]);
        for (int j = 0; j < 3; ++j) dndn[a][j] = R(td[j * nn + a]);
    }
    real jac[9];
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            real sum = R(0.0);
            for (int a = 0; a < nn; ++a) sum = sum + dndn[a][j] * xnode[3 * a + k];
            jac[j + 3 * k] = sum;
        }
    real det;
    invert_3(jac, &det);
    for (int a = 0; a < nn; ++a)
        for (int k = 0; k < 3; ++k) {
            real sum = R(0.0);
            for (int j = 0; j < 3; ++j) sum = sum + jac[k + 3 * j] * dndn[a][j];
            sfd[3 * a + k] = sum;
        }
    return det;
}

/* Shape functions of element type lt at point ig (0-based); returns the
 * node count, 0 for a type this restatement does not cover. */
static int shape_el(int lt, int ig, const real *xnode, real *sf, real *sfd, real *weight) {
    if (lt == 12 || lt == 13) {
        shape8(lt, ig, xnode, sf, sfd, weight);
        return 8;
    }
    if (lt == 22) {
        shape20(ig, xnode, sf, sfd, weight);
        return 20;
    }
    if (lt == 21 || lt == 23) {
        if (!tabq_ready) {
            tet10_tables();
            wedge15_tables();
            tabq_ready = 1;
        }
        if (lt == 21) {
            /* SHAPE_21: weight = -0.0416667 detJ (Shapef.f90:187) */
            *weight = -(R(0.0416667) * shape_tab(10, tab21_n[ig], &tab21_d[ig][0][0], xnode, sf, sfd));
            return 10;
        }
        /* SHAPE_23: weight = 0.16666667 detJ (Shapef.f90:483) */
        *weight = R(0.16666667) * shape_tab(15, tab23_n[ig], &tab23_d[ig][0][0], xnode, sf, sfd);
        return 15;
    }
    return 0;
}

/* DSTRNS, Sig_eps.f90:25-61: strain / rotation increments at a point. */
static int dstrns(int npl, const real *sfd, const real *dxl, real smale, real smalr, real *del, real *rl,
                  real *deff) {
    real ddxl[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = R(0.0);
            for (int n = 0; n < npl; ++n) s = s + sfd[3 * n + i] * dxl[3 * n + j];
            ddxl[i][j] = s;
        }
    real sum = R(0.0);
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (j + 1) % 3;
        del[i] = ddxl[i][i];
        del[i + 3] = (ddxl[k][j] + ddxl[j][k]);
        rl[i] = (ddxl[k][j] - ddxl[j][k]) / R(2.0);
        sum = sum + rl[i] * rl[i];
    }
    *deff = fj2vom(0, del);
    int err = 0;
    if (sum > smalr) err = 905;
    if (*deff > smale) err = err > 906 ? err : 906;
    return err;
}

/* ------------------------------------------------------------------------ */
/* Stif_sol.f90 / Elvals.f90 / Bconds.f90 / Fasolt3.f90                     */
/* ------------------------------------------------------------------------ */

#define IS(k, row) m->is[(size_t)(row) * m->maxofd + (k)] /* k, row 0-based slots */
#define SS(k, row) m->ss[(size_t)(row) * m->maxofd + (k)]

/* SETPTA, Stif_sol.f90:98-149: first-seen off-diagonal column lists. */
int orc_setpta(orc_model *m) {
    int ndof = 3 * m->nnod;
    memset(m->is, 0, sizeof(int32_t) * (size_t)ndof * m->maxofd);
    for (int el = 0; el < m->nlmnt; ++el) {
        int loc[60];
        int nn = m->nd[el];
        for (int i = 0; i < nn; ++i)
            for (int j = 0; j < 3; ++j) loc[3 * i + j] = 3 * m->ln[20 * el + i] - 2 + j;
        for (int i = 0; i < 3 * nn; ++i) {
            int row = loc[i];
            for (int j = 0; j < 3 * nn; ++j) {
                int col = loc[j];
                if (row != col) {
                    int isnew = 1;
                    for (int k = 1; k <= IS(0, row - 1); ++k)
                        if (IS(k, row - 1) == col) isnew = 0;
                    if (isnew) {
                        int noff = IS(0, row - 1) + 1;
                        if (noff + 1 > m->maxofd) return 955;
                        IS(0, row - 1) = noff;
                        IS(noff, row - 1) = col;
                    }
                }
            }
        }
    }
    return 0;
}

/* ELVALS, Elvals.f90:11-275: element loop, constitutive call per point,
 * internal force and stiffness into ss.  Returns the worst FAFERR code. */
int orc_elvals(orc_model *m) {
    wmat *wm = (wmat *)malloc(sizeof(wmat) * (size_t)m->nmatl);
    if (!wm) return -1;
    for (int i = 0; i < m->nmatl; ++i) load_mat(&wm[i], &m->mat[i]);
    wctl wc;
    load_ctl(&wc, m->ctl);
    const real efix = wc.efix, small = R(1.e-7);
    const real dt = R(m->tiem[m->inc - 1]);
    int worst = 0;
    m->n987 = m->n987_capped = m->nmaxit = 0;
    for (int el = 0; el < m->nlmnt; ++el) {
        int iel = m->le[el], itp = m->lt[el], nn = m->nd[el], ng = m->np[el];
        int in2[60];
        real xl[60], dxl[60], fl[60], esm[60][60];
        for (int i = 0; i < nn; ++i)
            for (int j = 0; j < 3; ++j) in2[3 * i + j] = 3 * m->ln[20 * el + i] - 2 + j;
        for (int i = 0; i < 3 * nn; ++i) {
            fl[i] = R(0.0);
            for (int j = 0; j < 3 * nn; ++j) esm[i][j] = R(0.0);
        }
        for (int i = 0; i < nn; ++i) {
            int ks = 3 * m->ln[20 * el + i] - 3;
            for (int j = 0; j < 3; ++j) {
                xl[3 * i + j] = R(m->x[ks + j]);
                dxl[3 * i + j] = R(m->dx[ks + j]);
            }
        }
        if (itp < 9) {
            /* tie element, Elvals.f90:62-77 */
            if (itp >= 1 && itp <= 3) {
                esm[itp - 1][itp - 1] = efix;
                esm[itp - 1][itp + 2] = -efix;
            } else if (itp >= 4 && itp <= 6) {
                int ii = (itp - 3) % 3; /* 1+MOD(itp-3,3), 0-based */
                esm[ii][ii] = efix;
                esm[ii][ii + 3] = -efix;
                ii = (ii + 1) % 3;
                esm[ii][ii] = efix;
                esm[ii][ii + 3] = -efix;
            }
        } else {
            for (int ig = 0; ig < ng; ++ig) {
                real sf[20], sfd[60], weight;
                if (shape_el(itp, ig, xl, sf, sfd, &weight) != nn) {
                    free(wm);
                    return -2; /* element type not covered */
                }
                if (weight < small) note_err(&worst, 920);
                int j = m->lg[8 * el + ig] - 1;
                real epl[6], stl[6], dsl[6], del[6], r[3], s[6], d[6][6], defl;
                for (int k = 0; k < 6; ++k) {
                    epl[k] = R(m->ep[6 * j + k]);
                    stl[k] = R(m->st[6 * j + k]);
                    dsl[k] = R(m->ds[6 * j + k]);
                }
                int e = dstrns(nn, sfd, dxl, wc.smale, wc.smalr, del, r, &defl);
                note_err(&worst, e);
                /* DSTRSS on this point's state columns */
                int mc = abs(iel);
                int ns = wm[mc - 1].nss;
                real lsv[ORC_MAXSV], ldsv[ORC_MAXSV], lsr[ORC_MAXSS];
                for (int k = 0; k < 3 + ns; ++k) {
                    lsv[k] = R(m->sv[(size_t)ORC_MAXSV * j + k]);
                    ldsv[k] = R(m->dsv[(size_t)ORC_MAXSV * j + k]);
                }
                for (int k = 0; k < ns; ++k) lsr[k] = R(m->srate[(size_t)ORC_MAXSS * j + k]);
                orc_counters c;
                memset(&c, 0, sizeof(c));
                e = dstrss(wm, &wc, dt, iel, &defl, del, stl, dsl, s, r, d, lsv, ldsv, lsr, &m->twinsys[j],
                           m->timestwinned[j], &m->tflag, &m->sumcpiter, &c,
                           ((uint64_t)m->rng_calls << 32) ^ (uint64_t)j);
                m->mat_code = mc;
                m->flops += c.flops;
                m->transc += c.transc;
                if (e == 987) {
                    m->n987 += 1;
                    if (c.capped) m->n987_capped += 1;
                }
                if (c.outer >= wc.maxit) m->nmaxit += 1;
                note_err(&worst, e);
                for (int k = 0; k < 3 + ns; ++k) m->dsv[(size_t)ORC_MAXSV * j + k] = ldsv[k];
                for (int k = 0; k < ns; ++k) m->srate[(size_t)ORC_MAXSS * j + k] = lsr[k];
                m->dep[j] = defl;
                for (int k = 0; k < 6; ++k) {
                    m->de[6 * j + k] = del[k];
                    m->ds[6 * j + k] = dsl[k];
                }
                for (int k = 0; k < 3; ++k) m->dr[3 * j + k] = r[k];
                /* nodal forces, Elvals.f90:122-126 */
                for (int i = 0; i < nn; ++i) {
                    const real *g = &sfd[3 * i];
                    fl[3 * i + 0] = fl[3 * i + 0] + weight * (s[0] * g[0] + s[5] * g[1] + s[4] * g[2]);
                    fl[3 * i + 1] = fl[3 * i + 1] + weight * (s[1] * g[1] + s[3] * g[2] + s[5] * g[0]);
                    fl[3 * i + 2] = fl[3 * i + 2] + weight * (s[2] * g[2] + s[4] * g[0] + s[3] * g[1]);
                }
                /* node-pair stiffness, Elvals.f90:129-237 (d 1-based in the comments) */
#define D(p, q) d[(p)-1][(q)-1]
#define S(p) s[(p)-1]
                for (int i = 0; i < nn; ++i)
                    for (int jn = 0; jn < nn; ++jn) {
                        const real a1 = sfd[3 * i], a2 = sfd[3 * i + 1], a3 = sfd[3 * i + 2];
                        const real b1 = sfd[3 * jn], b2 = sfd[3 * jn + 1], b3 = sfd[3 * jn + 2];
                        real gm[3][3];
                        gm[0][0] = (a1 * b1 * (D(1, 1) - S(1)) + a2 * b2 * (D(6, 6) - (S(1) - S(2)) / R(2.)) +
                                    a3 * b3 * (D(5, 5) - (S(1) - S(3)) / R(2.)) + a2 * b3 * (D(5, 6) + S(4) / R(2.)) +
                                    a3 * b1 * (D(1, 5)) + a1 * b2 * (D(1, 6)) + b2 * a3 * (D(5, 6) + S(4) / R(2.)) +
                                    b3 * a1 * (D(1, 5)) + b1 * a2 * (D(1, 6)));
                        gm[1][0] = (a1 * b1 * (D(1, 6) - S(6)) + a2 * b2 * (D(2, 6) - S(6)) +
                                    a3 * b3 * (D(4, 5) - S(6) / R(2.)) + a2 * b3 * (D(2, 5) - S(5) / R(4.)) +
                                    a3 * b1 * (D(1, 4) - S(4) / R(4.)) + a1 * b2 * (D(6, 6) - (S(1) + S(2)) / R(4.)) +
                                    b2 * a3 * (D(4, 6) - S(5) / R(4.)) + b3 * a1 * (D(5, 6) - S(4) / R(4.)) +
                                    b1 * a2 * (D(1, 2) - (S(1) + S(2)) / R(4.)));
                        gm[2][0] = (a1 * b1 * (D(1, 5) - S(5)) + a2 * b2 * (D(4, 6) - S(5) / R(2.)) +
                                    a3 * b3 * (D(3, 5) - S(5)) + a2 * b3 * (D(4, 5) - S(6) / R(4.)) +
                                    a3 * b1 * (D(1, 3) - (S(1) + S(3)) / R(4.)) + a1 * b2 * (D(5, 6) - S(4) / R(4.)) +
                                    b2 * a3 * (D(3, 6) - S(6) / R(4.)) + b3 * a1 * (D(5, 5) - (S(1) + S(3)) / R(4.)) +
                                    b1 * a2 * (D(1, 4) - S(4) / R(4.)));
                        gm[0][1] = (b1 * a1 * (D(1, 6) - S(6)) + b2 * a2 * (D(2, 6) - S(6)) +
                                    b3 * a3 * (D(4, 5) - S(6) / R(2.)) + b2 * a3 * (D(2, 5) - S(5) / R(4.)) +
                                    b3 * a1 * (D(1, 4) - S(4) / R(4.)) + b1 * a2 * (D(6, 6) - (S(1) + S(2)) / R(4.)) +
                                    a2 * b3 * (D(4, 6) - S(5) / R(4.)) + a3 * b1 * (D(5, 6) - S(4) / R(4.)) +
                                    a1 * b2 * (D(1, 2) - (S(1) + S(2)) / R(4.)));
                        gm[1][1] = (a1 * b1 * (D(6, 6) - (S(2) - S(1)) / R(2.)) + a2 * b2 * (D(2, 2) - S(2)) +
                                    a3 * b3 * (D(4, 4) - (S(2) - S(3)) / R(2.)) + a2 * b3 * (D(2, 4)) +
                                    a3 * b1 * (D(4, 6) + S(5) / R(2.)) + a1 * b2 * (D(2, 6)) + b2 * a3 * (D(2, 4)) +
                                    b3 * a1 * (D(4, 6) + S(5) / R(2.)) + b1 * a2 * (D(2, 6)));
                        gm[2][1] = (a1 * b1 * (D(5, 6) - S(4) / R(2.)) + a2 * b2 * (D(2, 4) - S(4)) +
                                    a3 * b3 * (D(3, 4) - S(4)) + a2 * b3 * (D(4, 4) - (S(2) + S(3)) / R(4.)) +
                                    a3 * b1 * (D(3, 6) - S(6) / R(4.)) + a1 * b2 * (D(2, 5) - S(5) / R(4.)) +
                                    b2 * a3 * (D(2, 3) - (S(2) + S(3)) / R(4.)) + b3 * a1 * (D(4, 5) - S(6) / R(4.)) +
                                    b1 * a2 * (D(4, 6) - S(5) / R(4.)));
                        gm[0][2] = (b1 * a1 * (D(1, 5) - S(5)) + b2 * a2 * (D(4, 6) - S(5) / R(2.)) +
                                    b3 * a3 * (D(3, 5) - S(5)) + b2 * a3 * (D(4, 5) - S(6) / R(4.)) +
                                    b3 * a1 * (D(1, 3) - (S(1) + S(3)) / R(4.)) + b1 * a2 * (D(5, 6) - S(4) / R(4.)) +
                                    a2 * b3 * (D(3, 6) - S(6) / R(4.)) + a3 * b1 * (D(5, 5) - (S(1) + S(3)) / R(4.)) +
                                    a1 * b2 * (D(1, 4) - S(4) / R(4.)));
                        gm[1][2] = (b1 * a1 * (D(5, 6) - S(4) / R(2.)) + b2 * a2 * (D(2, 4) - S(4)) +
                                    b3 * a3 * (D(3, 4) - S(4)) + b2 * a3 * (D(4, 4) - (S(2) + S(3)) / R(4.)) +
                                    b3 * a1 * (D(3, 6) - S(6) / R(4.)) + b1 * a2 * (D(2, 5) - S(5) / R(4.)) +
                                    a2 * b3 * (D(2, 3) - (S(2) + S(3)) / R(4.)) + a3 * b1 * (D(4, 5) - S(6) / R(4.)) +
                                    a1 * b2 * (D(4, 6) - S(5) / R(4.)));
                        gm[2][2] = (a1 * b1 * (D(5, 5) - (S(3) - S(1)) / R(2.)) +
                                    a2 * b2 * (D(4, 4) - (S(3) - S(2)) / R(2.)) + a3 * b3 * (D(3, 3) - S(3)) +
                                    a2 * b3 * (D(3, 4)) + a3 * b1 * (D(3, 5)) + a1 * b2 * (D(4, 5) + S(6) / R(2.)) +
                                    b2 * a3 * (D(3, 4)) + b3 * a1 * (D(3, 5)) + b1 * a2 * (D(4, 5) + S(6) / R(2.)));
                        for (int p = 0; p < 3; ++p)
                            for (int q = 0; q < 3; ++q) esm[3 * i + p][3 * jn + q] = esm[3 * i + p][3 * jn + q] + weight * gm[p][q];
                    }
#undef D
#undef S
            }
        }
        /* scatter, Elvals.f90:244-262 */
        for (int i = 0; i < 3 * nn; ++i) {
            int ii = in2[i];
            for (int j = 0; j < 3 * nn; ++j) {
                int jj = in2[j];
                int k;
                if (jj != ii) {
                    k = 1;
                    while (jj != IS(k, ii - 1)) ++k;
                } else {
                    k = 0;
                }
                SS(k, ii - 1) = (double)(R(SS(k, ii - 1)) + esm[i][j]);
            }
        }
        for (int i = 0; i < nn; ++i) {
            int k = 3 * m->ln[20 * el + i] - 3;
            for (int j = 0; j < 3; ++j) m->fc[k + j] = (double)(R(m->fc[k + j]) + fl[3 * i + j]);
        }
    }
    free(wm);
    if (worst > m->err) m->err = worst;
    return worst;
}

/* SUVALS, Bconds.f90:19-113: penalty constraints and loads. */
int orc_suvals(orc_model *m) {
    static const int point[3][3] = {{0, 1, 3}, {1, 2, 4}, {3, 4, 5}}; /* Bconds.f90:33 */
    const real efix = R(m->ctl->efix);
    const int inc = m->inc;
    int ndof = 3 * m->nnod;
    for (int i = 0; i < ndof; ++i) m->ddx[i] = (double)R(m->fc[i]);
    for (int inod = 0; inod < m->nnod; ++inod) {
        int ibc = m->ib[inod];
        if (ibc < 100) continue;
        int idof = 3 * inod;
        int i_kin = (ibc / 100) % 100;
        if (i_kin > 0) {
            real xsloc[15], xspec[3];
            const double *xs = &m->x_spec[((size_t)(inc - 1) * 99 + (i_kin - 1)) * 15];
            for (int k = 0; k < 15; ++k) xsloc[k] = R(xs[k]);
            if (i_kin < 50) {
                xspec[0] = xsloc[6];
                xspec[1] = xsloc[10];
                xspec[2] = xsloc[14];
            } else {
                real d_spec[3][3];
                int l = 6;
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) d_spec[j][k] = xsloc[l++];
                for (int j = 0; j < 3; ++j) {
                    real s = R(0.0);
                    for (int k = 0; k < 3; ++k) s = s + d_spec[j][k] * (R(m->x[idof + k]) + R(m->dx[idof + k]));
                    xspec[j] = s;
                }
            }
            for (int i = 0; i < 3; ++i) {
                int row = idof + i;
                for (int j = 0; j < 3; ++j) {
                    int k;
                    if (j != i) {
                        k = 1;
                        while (idof + j + 1 != IS(k, row)) ++k;
                    } else {
                        k = 0;
                    }
                    SS(k, row) = (double)(R(SS(k, row)) + (efix * xsloc[point[i][j]]));
                }
                real sum = R(0.0);
                for (int j = 0; j < 3; ++j) sum = sum + efix * xsloc[point[i][j]] * (R(m->dx[idof + j]) - xspec[j]);
                m->ddx[row] = (double)(R(m->ddx[row]) + sum);
            }
        }
        int i_spf = (ibc / 10000) % 100;
        if (i_spf > 0) {
            const double *fs = &m->f_spec[((size_t)(inc - 1) * 99 + (i_spf - 1)) * 3];
            for (int i = 0; i < 3; ++i) m->ddx[idof + i] = (double)(R(m->ddx[idof + i]) + R(fs[i]));
        }
    }
    return 0;
}

/* FACSLS, Stif_sol.f90:161-215: natural-order SOR on the row storage. */
int orc_facsls(orc_model *m, int ndof) {
    const real alphal = R(m->ctl->sor_omega);
    const int maxitl = m->ctl->sor_maxit;
    const real tolerl = R(m->ctl->sor_tol) * R(m->ctl->sor_tol);
    real *y = (real *)malloc(sizeof(real) * (size_t)ndof);
    real *xo = (real *)malloc(sizeof(real) * (size_t)ndof);
    real *xv = (real *)malloc(sizeof(real) * (size_t)ndof);
    if (!y || !xo || !xv) return -1;
    for (int i = 0; i < ndof; ++i) {
        y[i] = R(m->ddx[i]);
        xv[i] = R(0.0);
    }
    int err = 0;
    for (int i = 0; i < ndof; ++i)
        if (fabs(R(SS(0, i))) == R(0.0)) err = 903;
    /* kind 1: relative Jacobi-scaled residual |D^-1 (b - A x)| / |D^-1 b| (build's own
     * criterion; the reference only has the relative-update test) */
    double bnorm = 0.0;
    if (m->ctl->sor_kind == 1)
        for (int i = 0; i < ndof; ++i) {
            double yi = (double)y[i] / (double)R(SS(0, i));
            bnorm += yi * yi;
        }
    int iter = 0, conv = 0;
    real error = R(0.0);
    double rel = 0.0;
    while (!conv && iter < maxitl && !err) {
        ++iter;
        for (int i = 0; i < ndof; ++i) xo[i] = xv[i];
        for (int i = 0; i < ndof; ++i) {
            real sum = R(0.0);
            for (int j = 1; j <= IS(0, i); ++j) sum = sum + R(SS(j, i)) * xv[IS(j, i) - 1];
            sum = (y[i] - sum) / R(SS(0, i));
            xv[i] = xv[i] + alphal * (sum - xv[i]);
        }
        if (m->ctl->sor_kind == 0) {
            real num = R(0.0), den = R(0.0);
            for (int i = 0; i < ndof; ++i) {
                xo[i] = xv[i] - xo[i];
                num = num + xo[i] * xo[i];
            }
            for (int i = 0; i < ndof; ++i) den = den + xv[i] * xv[i];
            error = num / den;
            conv = (error < tolerl);
        } else {
            /* relative Jacobi-scaled residual, accumulated in double */
            double rr = 0.0;
            for (int i = 0; i < ndof; ++i) {
                double sum = (double)R(SS(0, i)) * (double)xv[i];
                for (int j = 1; j <= IS(0, i); ++j) sum += (double)R(SS(j, i)) * (double)xv[IS(j, i) - 1];
                double ri = ((double)y[i] - sum) / (double)R(SS(0, i));
                rr += ri * ri;
            }
            rel = bnorm > 0.0 ? sqrt(rr / bnorm) : 0.0;
            conv = rel < m->ctl->sor_tol;
        }
    }
    for (int i = 0; i < ndof; ++i) m->ddx[i] = (double)xv[i];
    m->sor_sweeps = iter;
    if (m->ctl->sor_kind == 0) {
        m->sor_error = (double)error;
        if (!err && error > R(10.0) * tolerl) err = 904;
    } else {
        m->sor_error = rel;
        if (!err && rel > 10.0 * m->ctl->sor_tol) err = 904;
    }
    free(y);
    free(xo);
    free(xv);
    if (err > m->err) m->err = err;
    return err;
}

/* NRITER, Stif_sol.f90:10-88: damped Newton for one increment. */
int orc_nriter(orc_model *m) {
    const real toler = R(m->ctl->toler), alpha = R(m->ctl->alpha);
    const int ndof = 3 * m->nnod;
    int nit = 0, l_conv = 1, err = 0, worst = 0;
    real sdx = R(0.0), sfc, alp2 = R(1.0);
    /* codes that end the loop: all fatal ones, or with keep_going those other
     * than 987 (slip Newton), 910 (singular slip matrix) and 904 (SOR) */
#define ORC_STOPS(e) ((e) > 500 && !(m->ctl->keep_going && ((e) == 987 || (e) == 910 || (e) == 904)))
    while (l_conv) {
        ++nit;
        m->sumcpiter = 0.0;
        for (int i = 0; i < ndof; ++i) m->fc[i] = 0.0;
        for (size_t i = 0; i < (size_t)ndof * m->maxofd; ++i) m->ss[i] = 0.0;
        m->rng_calls += 1;
        err = orc_elvals(m);
        if (err > worst) worst = err;
        if (ORC_STOPS(err)) break;
        orc_suvals(m);
        sfc = R(0.0);
        for (int i = 0; i < m->nnod; ++i)
            if (m->ib[i] < 10) {
                real a = R(m->ddx[3 * i]), b = R(m->ddx[3 * i + 1]), c = R(m->ddx[3 * i + 2]);
                sfc = sfc + a * a + b * b + c * c;
            }
        sfc = sqrt(sfc);
        err = orc_facsls(m, ndof);
        if (err > worst) worst = err;
        if (ORC_STOPS(err)) break;
        real dd = R(0.0), ee = R(0.0);
        for (int i = 0; i < ndof; ++i) dd = dd + R(m->dx[i]) * R(m->dx[i]);
        for (int i = 0; i < ndof; ++i) ee = ee + R(m->ddx[i]) * R(m->ddx[i]);
        real errdx = sqrt(dd), erddx = sqrt(ee);
        sdx = (errdx == R(0.0)) ? toler : erddx / errdx;
        alp2 = R(1.0) / (R(1.0) + alpha * sdx);
        m->sumcpiter = (double)(R(m->sumcpiter) / (real)m->ngps);
        for (int i = 0; i < ndof; ++i) {
            real xc = alp2 * R(m->ddx[i]);
            m->x[i] = (double)(R(m->x[i]) - xc);
            m->dx[i] = (double)(R(m->dx[i]) - xc);
        }
        m->sfc = sfc;
        l_conv = (sdx >= toler) && (nit < m->ctl->maxit);
    }
    m->nit = nit;
    m->sdx = sdx;
    m->alp2 = alp2;
    const int stopped = ORC_STOPS(err);
#undef ORC_STOPS
    err = worst;
    if (!stopped) {
        /* cumulative slip, Stif_sol.f90:80-84 */
        int ns = m->mat[(m->mat_code > 0 ? m->mat_code : 1) - 1].nss;
        real dt = R(m->tiem[m->inc - 1]);
        for (int g = 0; g < m->ngps; ++g)
            for (int k = 0; k < ns; ++k) {
                size_t o = (size_t)ORC_MAXSS * g + k;
                m->cumslip[o] = (double)(R(m->cumslip[o]) + (R(m->srate[o]) * dt));
            }
    }
    return err;
}

/* UPDATE_NODE, Fasolt3.f90:231-244: predictor x += dx. */
void orc_update_node(orc_model *m) {
    for (int i = 0; i < 3 * m->nnod; ++i) m->x[i] = (double)(R(m->x[i]) + R(m->dx[i]));
}

/* UPDATE_GPTV, Fasolt3.f90:254-284: commit the point increments. */
void orc_update_gptv(orc_model *m) {
    for (int i = 0; i < m->ngps; ++i) {
        m->eps[i] = (double)(R(m->eps[i]) + R(m->dep[i]));
        for (int j = 0; j < 3; ++j) m->rt[3 * i + j] = (double)(R(m->rt[3 * i + j]) + R(m->dr[3 * i + j]));
        for (int j = 0; j < 6; ++j) {
            m->ep[6 * i + j] = (double)(R(m->ep[6 * i + j]) + R(m->de[6 * i + j]));
            m->st[6 * i + j] = (double)(R(m->st[6 * i + j]) + R(m->ds[6 * i + j]));
        }
        for (int j = 0; j < m->nsv; ++j) {
            size_t o = (size_t)ORC_MAXSV * i + j;
            m->sv[o] = (double)(R(m->sv[o]) + R(m->dsv[o]));
        }
    }
}

/* DoReorient, Fasolt3.f90:79-178: twin reorientation (material 1 tables). */
int orc_doreorient(orc_model *m) {
    const orc_material *mt = &m->mat[0];
    int n = 0;
    for (int j = 0; j < m->ngps; ++j) {
        if (m->twinsys[j] == 0) continue;
        int i = m->twinsys[j] - 1;
        real *svj = NULL;
        real e1 = R(m->sv[(size_t)ORC_MAXSV * j + 0]);
        real e2 = R(m->sv[(size_t)ORC_MAXSV * j + 1]);
        real e3 = R(m->sv[(size_t)ORC_MAXSV * j + 2]);
        (void)svj;
        real ang = R(mt->ssang[i]) * R(3.14159) / R(180.);
        real c1 = cos(e1), c2 = cos(e2), c3 = cos(e3), ss1 = sin(e1), s2 = sin(e2), s3 = sin(e3);
        real rm[3][3];
        rm[0][0] = c1 * c3 - c2 * ss1 * s3;
        rm[0][1] = -c1 * s3 - c2 * c3 * ss1;
        rm[0][2] = ss1 * s2;
        rm[1][0] = c3 * ss1 + c1 * c2 * s3;
        rm[1][1] = c1 * c2 * c3 - ss1 * s3;
        rm[1][2] = -c1 * s2;
        rm[2][0] = s2 * s3;
        rm[2][1] = c3 * s2;
        rm[2][2] = c2;
        real nv[3], dv[3];
        for (int p = 0; p < 3; ++p) {
            real sn = R(0.0), sd = R(0.0);
            for (int q = 0; q < 3; ++q) {
                sn = sn + rm[p][q] * R(mt->an[i][q]);
                sd = sd + rm[p][q] * R(mt->ab[i][q]);
            }
            nv[p] = sn;
            dv[p] = sd;
        }
        real u1 = nv[1] * dv[2] - nv[2] * dv[1];
        real u2 = nv[2] * dv[0] - nv[0] * dv[2];
        real u3 = nv[0] * dv[1] - nv[1] * dv[0];
        real c = cos(ang), s = sin(ang);
        real erm[3][3];
        erm[0][0] = c + u1 * u1 * (R(1.) - c);
        erm[0][1] = u1 * u2 * (R(1.) - c) - u3 * s;
        erm[0][2] = u1 * u3 * (R(1.) - c) + u2 * s;
        erm[1][0] = u1 * u2 * (R(1.) - c) + u3 * s;
        erm[1][1] = c + u2 * u2 * (R(1.) - c);
        erm[1][2] = u2 * u3 * (R(1.) - c) - u1 * s;
        erm[2][0] = u1 * u3 * (R(1.) - c) - u2 * s;
        erm[2][1] = u2 * u3 * (R(1.) - c) + u1 * s;
        erm[2][2] = c + u3 * u3 * (R(1.) - c);
        real frm[3][3];
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) {
                real sum = R(0.0);
                for (int k = 0; k < 3; ++k) sum = sum + erm[p][k] * rm[k][q];
                frm[p][q] = sum;
            }
        real e1n = atan2(frm[0][2], -frm[1][2]);
        if (fabs(frm[2][2]) > R(1.)) frm[2][2] = frm[2][2] / fabs(frm[2][2]);
        real e2n = acos(frm[2][2]);
        real e3n = atan2(frm[2][0], frm[2][1]);
        m->sv[(size_t)ORC_MAXSV * j + 0] = e1n;
        m->sv[(size_t)ORC_MAXSV * j + 1] = e2n;
        m->sv[(size_t)ORC_MAXSV * j + 2] = e3n;
        m->twinsys[j] = 0;
        m->timestwinned[j] += 1;
        ++n;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* thin double-boundary wrappers for closed-form tests                      */
/* ------------------------------------------------------------------------ */
double orc_spower(double a, double b) { return (double)spower(R(a), R(b)); }
double orc_fj2vom(int lstress, const double *x) {
    real v[6];
    for (int i = 0; i < 6; ++i) v[i] = R(x[i]);
    return (double)fj2vom(lstress, v);
}
int orc_factor(int n, double *w) {
    real *v = (real *)malloc(sizeof(real) * (size_t)n * n);
    for (int i = 0; i < n * n; ++i) v[i] = R(w[i]);
    int f = factor(n, v);
    for (int i = 0; i < n * n; ++i) w[i] = v[i];
    free(v);
    return f;
}
void orc_solver(int n, const double *w, double *x) {
    real *v = (real *)malloc(sizeof(real) * (size_t)n * n);
    real *b = (real *)malloc(sizeof(real) * (size_t)n);
    for (int i = 0; i < n * n; ++i) v[i] = R(w[i]);
    for (int i = 0; i < n; ++i) b[i] = R(x[i]);
    solver(n, v, b);
    for (int i = 0; i < n; ++i) x[i] = b[i];
    free(v);
    free(b);
}
void orc_ctmatrx(double ph1, double phi, double ph2, double *a, double *aug) {
    real la[3][3], lg[3];
    ctmatrx(R(ph1), R(phi), R(ph2), la, lg);
    for (int i = 0; i < 3; ++i) {
        aug[i] = lg[i];
        for (int j = 0; j < 3; ++j) a[i + 3 * j] = la[i][j];
    }
}
void orc_elastc(const orc_material *mt, double ph1, double phi, double ph2, double *el_comp) {
    wmat *w = (wmat *)malloc(sizeof(wmat));
    load_mat(w, mt);
    real el[6][6];
    elastc(w, R(ph1), R(phi), R(ph2), el, NULL);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) el_comp[i + 6 * j] = el[i][j];
    free(w);
}
void orc_tens_rot(const double *r, const double *a, double *b) {
    real lr[3], la[6], lb[6];
    for (int i = 0; i < 3; ++i) lr[i] = R(r[i]);
    for (int i = 0; i < 6; ++i) la[i] = R(a[i]);
    tens_rot(lr, la, lb);
    for (int i = 0; i < 6; ++i) b[i] = lb[i];
}
int orc_invert_3(double *mm, double *det) {
    real v[9], d;
    for (int i = 0; i < 9; ++i) v[i] = R(mm[i]);
    invert_3(v, &d);
    for (int i = 0; i < 9; ++i) mm[i] = v[i];
    *det = d;
    return d != R(0.0);
}
void orc_shape(int lt, int ig, const double *xnode, double *sf, double *sfd, double *weight) {
    const int nn = (lt == 22) ? 20 : (lt == 21) ? 10 : (lt == 23) ? 15 : 8;
    real x[60], f[20], d[60], w = R(0.0);
    for (int i = 0; i < 3 * nn; ++i) x[i] = R(xnode[i]);
    shape_el(lt, ig, x, f, d, &w);
    for (int i = 0; i < nn; ++i) sf[i] = f[i];
    for (int i = 0; i < 3 * nn; ++i) sfd[i] = d[i];
    *weight = w;
}
int orc_dstrns(int npl, const double *sfd, const double *dxl, double smale, double smalr, double *del,
               double *rl, double *deff) {
    real s[60], d[60], e[6], r[3], f;
    for (int i = 0; i < 3 * npl; ++i) {
        s[i] = R(sfd[i]);
        d[i] = R(dxl[i]);
    }
    int err = dstrns(npl, s, d, R(smale), R(smalr), e, r, &f);
    for (int i = 0; i < 6; ++i) del[i] = e[i];
    for (int i = 0; i < 3; ++i) rl[i] = r[i];
    *deff = f;
    return err;
}
