// Per-point rate-dependent crystal-plasticity update (FP64), device side.
//
// One integration point per lane.  The routines restate the reference's
// constitutive chain (gabortimar/CPFEM_code src/fasolt/):
//   DSTRSS  Sig_eps.f90:84-151      ELASTC  ela-plas.f90:12-86
//   CPSTRS  Cpstrs.f90:34-172       PLASTC  ela-plas.f90:110-201
//   SLPINI  crys-cons.f90:143-197   SLPFTA  crys-cons.f90:24-128
//   SCALE_IT Cpstrs.f90:187-230     SPOWER  crys-cons.f90:210-230
//   FACTOR/SOLVER crys-cons.f90:244-309
//   CTMATRX/BTMATRX/ECHROT crys_math.f90:25-167  TENS_ROT Sig_eps.f90:166-202
//
// Register-resident layout: every per-system quantity is a fixed array over
// the N systems of the material, indexed by compile-time constants only.  The
// reference compacts the active set into an nsa x nsa matrix; here the active
// set is a mask and inactive rows/columns of the slip matrix are identity
// rows/columns.  Doolittle elimination and the triangular solves then perform
// exactly the reference's operations on the active entries (the extra ones
// subtract exact zeros), so results match the compacted form bit for bit up
// to contraction of a*b+c into FMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpfem {

// Material tables of one material set, FP64, in the layout the kernel reads.
template <int N>
struct DevMaterial {
    int nss, twinlim;
    double twinprob;
    double e[9];
    double ssang[N];
    double ab[N][3], an[N][3];
    double am[N];        // ampar(1): rate sensitivity m
    double th0[N];       // ampar(2): theta0
    double th4[N];       // ampar(3): thetaIV
    double alp[N];       // ampar(4): alpha index
    double taus[N];      // ampar(5): saturation stress (rate 1)
    double nsat[N];      // ampar(6): saturation rate sensitivity
    double smatrx[N][N]; // smatrx[r][c] == Fortran smatrx(r,c): H_rc = smatrx(r,c) theta_c
    double conv;         // toler * max(m)
};

struct DevControl {
    double toler, alpha, afactor;
    int maxit;
    int pad_;
};

// Counters of one point: algorithmic FLOPs and transcendentals of the
// reference chain, evaluated from its own iteration path.
struct Count {
    int64_t flops, transc;
    int outer, inner;
    __device__ void add(int64_t f, int64_t t) { flops += f; transc += t; }
};

__device__ __forceinline__ int64_t factor_flops(int n) {
    return (int64_t)n * (n - 1) / 2 + (int64_t)n * (n - 1) * (2 * n - 1) / 3;
}
__device__ __forceinline__ int64_t solver_flops(int n) { return 2 * (int64_t)n * n - n; }

// x**y with the exact special cases of pow written out (pow(x,0)=1,
// pow(x,1)=x): the hardening exponents are 0 or 1 for Voce-type materials.
__device__ __forceinline__ double pow_fast(double x, double y) {
    if (y == 0.0) return 1.0;
    if (y == 1.0) return x;
    return pow(x, y);
}

// SPOWER: signed power with clamps.
__device__ __forceinline__ double spower(double a, double b) {
    double aa = fabs(a);
    if (aa < 1.e-35) return 0.0;
    double sparg = b * log(aa);
    if (sparg < -80.0) sparg = -80.0;
    if (sparg > 80.0) sparg = 80.0;
    return copysign(exp(sparg), a);
}

__device__ __forceinline__ double fj2vom(bool lstress, const double x[6]) {
    const double amult = lstress ? 6.0 : 1.5;
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        sum = sum + (x[i] - x[j]) * (x[i] - x[j]) + amult * x[i + 3] * x[i + 3];
    }
    return sqrt(sum / 4.5);
}

// Doolittle LU without pivoting on an n x n register matrix w[row][col].
template <int n>
__device__ __forceinline__ bool factor(double (&w)[n][n]) {
    bool fail = false;
#pragma unroll
    for (int k = 0; k < n - 1; ++k) {
        fail = fail || (w[k][k] == 0.0);
        const double piv = w[k][k];
#pragma unroll
        for (int i = k + 1; i < n; ++i) {
            w[i][k] = w[i][k] / piv;
#pragma unroll
            for (int j = k + 1; j < n; ++j) w[i][j] = w[i][j] - w[i][k] * w[k][j];
        }
    }
    fail = fail || (w[n - 1][n - 1] == 0.0);
    return fail;
}

template <int n>
__device__ __forceinline__ void solver(const double (&w)[n][n], double (&x)[n]) {
#pragma unroll
    for (int i = 1; i < n; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) x[i] = x[i] - w[i][j] * x[j];
    x[n - 1] = x[n - 1] / w[n - 1][n - 1];
#pragma unroll
    for (int i = n - 2; i >= 0; --i) {
#pragma unroll
        for (int j = i + 1; j < n; ++j) x[i] = x[i] - w[i][j] * x[j];
        x[i] = x[i] / w[i][i];
    }
}

// CTMATRX: orientation matrix a[row][col] and the Euler-rate helper terms.
__device__ __forceinline__ void ctmatrx(double ph1, double phi, double ph2, double a[3][3], double aug[3]) {
    double s1, c1, s2, c2, s, c;
    sincos(ph1, &s1, &c1);
    sincos(ph2, &s2, &c2);
    sincos(phi, &s, &c);
    const double c1c2 = c1 * c2, s1s2 = s1 * s2, c1s2 = c1 * s2, s1c2 = s1 * c2;
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

// ELASTC: compliance in specimen axes; el[i][j] full symmetric 6x6.
template <int N>
__device__ __forceinline__ void elastc(const DevMaterial<N>& M, double ph1, double phi, double ph2, double el[6][6]) {
    double a0[3][3], aug[3], a[3][3];
    ctmatrx(ph1, phi, ph2, a0, aug);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) a[i][j] = a0[j][i];
    const double* e = M.e;
#pragma unroll
    for (int iset = 0; iset < 6; ++iset) {
        const int i = iset < 3 ? iset : (iset - 2) % 3;
        const int j = iset < 3 ? iset : (i + 1) % 3;
        const double mult = iset < 3 ? 1.0 : 2.0;
#pragma unroll
        for (int jset = iset; jset < 6; ++jset) {
            const int k = jset < 3 ? jset : (jset - 2) % 3;
            const int l = jset < 3 ? jset : (k + 1) % 3;
            const double term1 = a[i][0] * a[j][0] * a[k][0] * a[l][0] * e[0] +
                                 a[i][1] * a[j][1] * a[k][1] * a[l][1] * e[1] +
                                 a[i][2] * a[j][2] * a[k][2] * a[l][2] * e[2];
            const double term2 = a[i][1] * a[j][1] * a[k][2] * a[l][2] * e[3] +
                                 a[i][2] * a[j][2] * a[k][1] * a[l][1] * e[3] +
                                 a[i][2] * a[j][2] * a[k][0] * a[l][0] * e[4] +
                                 a[i][0] * a[j][0] * a[k][2] * a[l][2] * e[4] +
                                 a[i][0] * a[j][0] * a[k][1] * a[l][1] * e[5] +
                                 a[i][1] * a[j][1] * a[k][0] * a[l][0] * e[5];
            const double term3 = a[i][1] * a[j][2] * a[k][1] * a[l][2] * e[6] +
                                 a[i][2] * a[j][1] * a[k][2] * a[l][1] * e[6] +
                                 a[i][2] * a[j][0] * a[k][2] * a[l][0] * e[7] +
                                 a[i][0] * a[j][2] * a[k][0] * a[l][2] * e[7] +
                                 a[i][0] * a[j][1] * a[k][0] * a[l][1] * e[8] +
                                 a[i][1] * a[j][0] * a[k][1] * a[l][0] * e[8];
            el[iset][jset] = mult * (term1 + term2 + 0.5 * term3);
        }
    }
#pragma unroll
    for (int iset = 0; iset < 5; ++iset)
#pragma unroll
        for (int jset = iset + 1; jset < 6; ++jset) el[jset][iset] = el[iset][jset];
}

// TENS_ROT: rotate a symmetric tensor (tensor shears) by the small rotation r.
__device__ __forceinline__ void tens_rot(const double r[3], const double a[6], double b[6]) {
    double rot[3][3], old[3][3], tmp[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        old[i][i] = a[i];
        old[j][k] = a[i + 3];
        old[k][j] = a[i + 3];
        rot[j][k] = r[i];
        rot[k][j] = -r[i];
        rot[i][i] = sqrt(1.0 - r[j] * r[j] - r[k] * r[k]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s = s + old[i][k] * rot[k][j];
            tmp[i][j] = s;
        }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        double sii = 0.0, sjk = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) sii = sii + rot[q][i] * tmp[q][i];
#pragma unroll
        for (int q = 0; q < 3; ++q) sjk = sjk + rot[q][j] * tmp[q][k];
        b[i] = sii;
        b[i + 3] = sjk;
    }
}

// Slip direction / normal of system i in specimen axes (BTMATRX inner sums).
template <int N>
__device__ __forceinline__ void slip_vectors(const DevMaterial<N>& M, const double a[3][3], int i, double sdr[3],
                                             double spn[3]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            s1 = s1 + a[l][j] * M.ab[i][l];
            s2 = s2 + a[l][j] * M.an[i][l];
        }
        sdr[j] = s1;
        spn[j] = s2;
    }
}

// Splitmix64 draw for the twin test (only consulted when twinprob < 1).
__device__ __forceinline__ double twin_draw(uint64_t key) {
    uint64_t z = key + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Inputs and outputs of one point.
struct PointIO {
    int iel;
    double de[6], r[3], st[6], ds[6];   // ds: guess in, increment out
    double sv0[3], dsv0[3];             // Euler angles and their increments
    double t[6], chi[6][6], deff, erepmd;
    int twinsys, timestwinned;
    int tflag, err;
    uint64_t rng_key;
};

// Slip-system scratch of one point (module FAS_CPL scratch of the reference).
template <int N>
struct SlipState {
    double b[N][6];
    double gam[N], tau[N], s0[N], s1[N];
    bool act[N];
    int nsa;
    double ah[N][N];   // ah[row][col], active mask embedded as identity rows/cols
};

// PLASTC: hardening on all systems, s1, and ah on the active pairs.
template <int N>
__device__ __forceinline__ void plastc(const DevMaterial<N>& M, SlipState<N>& S, double dt, Count& cnt) {
    double grad_h[N];
    double sumss = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) sumss = sumss + fabs(S.gam[i]);
    const double rate = sumss / dt;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double sst = M.taus[i] * pow_fast(rate, M.nsat[i]);
        double arg = (sst != 0.0) ? 1.0 - (S.s0[i] / sst) : 0.0;
        if (arg < 0.0) arg = 0.0;
        grad_h[i] = M.th4[i] + M.th0[i] * copysign(pow_fast(fabs(arg), M.alp[i]), arg);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double s = S.s0[i];
#pragma unroll
        for (int j = 0; j < N; ++j) s = s + (M.smatrx[i][j] * grad_h[j]) * fabs(S.gam[j]);
        S.s1[i] = s;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double inv = S.s1[i];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double v;
            if (S.act[i] && S.act[j]) {
                const double ag = fabs(S.gam[j]);
                v = (ag > 1.e-14) ? ((S.gam[j] / ag) * (M.smatrx[i][j] * grad_h[j])) / inv : 0.0;
            } else {
                v = (i == j) ? 1.0 : 0.0;
            }
            S.ah[i][j] = v;
        }
    }
    cnt.add(3 * (int64_t)N * N + 6 * N + 3 * (int64_t)S.nsa * S.nsa, 4 * N);
}

// SCALE_IT: cap the largest active slip at gamcap; scales s, gam and tau.
template <int N>
__device__ __forceinline__ void scale_it(const DevMaterial<N>& M, SlipState<N>& S, double s[6], double gamcap,
                                         double afactor, Count& cnt) {
    // the reference seeds amax with am(1), the first system (Cpstrs.f90:200)
    double amax = M.am[0];
    double gammax = -1.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (S.act[i]) {
            if (gammax < 0.0) {
                gammax = fabs(S.gam[i]);
            } else if (fabs(S.gam[i]) > gammax) {
                gammax = fabs(S.gam[i]);
                amax = M.am[i];
            }
        }
    }
    if (gammax > gamcap) {
        const double scale = gamcap / gammax;
        double smult = spower(0.9 * scale, afactor * amax);
        if (smult == 0.0) smult = 0.004 / amax;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = smult * s[i];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            S.gam[i] = scale * S.gam[i];
            S.tau[i] = smult * S.tau[i];
        }
        cnt.add(S.nsa + 2 * N + 6, 2);
    } else {
        cnt.add(S.nsa, 0);
    }
}

// SLPINI: power-law predictor, twin flagging and the active set.
template <int N>
__device__ __forceinline__ void slpini(const DevMaterial<N>& M, SlipState<N>& S, double dt, PointIO& io,
                                       Count& cnt) {
    double gammax = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double g;
        if (fabs(M.ssang[i]) > 1.5) {
            g = 0.0;
            if ((S.tau[i] > S.s0[i]) && (io.timestwinned != M.twinlim) && (io.twinsys == 0)) {
                const double rn = twin_draw(io.rng_key);
                if (rn < M.twinprob) {
                    io.tflag = 1;
                    io.twinsys = i + 1;
                }
            }
        } else if (M.ssang[i] * S.tau[i] / S.s0[i] < 0.0) {
            g = 0.0;
        } else {
            g = dt * spower((S.tau[i] / S.s0[i]), (1.0 / M.am[i]));
        }
        S.gam[i] = g;
        if (fabs(g) > gammax) gammax = fabs(g);
    }
    const double cutoff = 1.e-6 * gammax;
    int nsa = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool a = !(fabs(S.gam[i]) < cutoff);
        if (!a) S.gam[i] = 0.0;
        S.act[i] = a;
        nsa += a ? 1 : 0;
    }
    S.nsa = nsa;
    cnt.add(5 * N, 2 * N);
}

// SLPFTA: Newton on ln(gam) over the active set; leaves ah in X form.
template <int N>
__device__ __forceinline__ void slpfta(const DevMaterial<N>& M, const DevControl& C, SlipState<N>& S, double dt,
                                       int& err, Count& cnt) {
    int iter = 0;
    bool more = true;
    double tauerr = 0.0;
    while (more) {
        ++iter;
        cnt.inner += 1;
        plastc(M, S, dt, cnt);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if (S.act[i]) {
                const double gamn = S.gam[i];
#pragma unroll
                for (int j = 0; j < N; ++j)
                    if (S.act[j]) S.ah[j][i] = gamn * S.ah[j][i];
                S.ah[i][i] = S.ah[i][i] + M.am[i];
            }
        }
        if (factor<N>(S.ah)) err = err > 910 ? err : 910;
        double errtau[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double e = 0.0;
            if (S.act[i]) {
                const double taufg = S.s1[i] * spower((S.gam[i] / dt), M.am[i]);
                if ((S.tau[i] * taufg) > 0.0) e = log(S.tau[i] / taufg);
            }
            errtau[i] = e;
        }
        tauerr = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (S.act[i]) tauerr = tauerr + errtau[i] * errtau[i];
        tauerr = sqrt(tauerr);
        solver<N>(S.ah, errtau);
        double errmax = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (S.act[i] && fabs(errtau[i]) > errmax) errmax = fabs(errtau[i]);
        const double scale = (errmax > 5.0) ? 5.0 / errmax : 1.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (S.act[i]) S.gam[i] = S.gam[i] * exp(scale * errtau[i]);
        more = (iter < C.maxit) && (tauerr > M.conv);
        const int a = S.nsa;
        cnt.add((int64_t)a * a + 10 * a + 1 + factor_flops(a) + solver_flops(a), 4 * a);
    }
    if (tauerr > 0.5) err = err > 987 ? err : 987;
    plastc(M, S, dt, cnt);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (S.act[i]) {
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (S.act[j]) S.ah[j][i] = S.tau[j] * S.ah[j][i];
            const double ag = fabs(S.gam[i]);
            S.ah[i][i] = S.ah[i][i] + ((ag > 1.e-12) ? M.am[i] * S.tau[i] / S.gam[i] : M.am[i] / 1.e-12);
        }
    }
    cnt.add((int64_t)S.nsa * S.nsa + 3 * S.nsa, 0);
}

// CPSTRS + the plastic branch of DSTRSS for one point.
template <int N>
__device__ __forceinline__ void cpstrs(const DevMaterial<N>& M, const DevControl& C, double dt, double el[6][6],
                                       double s[6], const double t[6], PointIO& io, double* s0_io, double* dsv_io,
                                       double* srate_out, Count& cnt) {
    SlipState<N> S;
    double a[3][3], aug[3];
    const double ph1 = io.sv0[0] + io.dsv0[0];
    const double phi = io.sv0[1] + io.dsv0[1];
    const double ph2 = io.sv0[2] + io.dsv0[2];
#pragma unroll
    for (int i = 0; i < N; ++i) S.s0[i] = s0_io[i];
    const double gamcap = 6.0 * fj2vom(false, io.de);
    ctmatrx(ph1, phi, ph2, a, aug);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double sdr[3], spn[3];
        slip_vectors(M, a, i, sdr, spn);
        S.b[i][0] = sdr[0] * spn[0];
        S.b[i][1] = sdr[1] * spn[1];
        S.b[i][2] = sdr[2] * spn[2];
        S.b[i][3] = sdr[1] * spn[2] + sdr[2] * spn[1];
        S.b[i][4] = sdr[2] * spn[0] + sdr[0] * spn[2];
        S.b[i][5] = sdr[0] * spn[1] + sdr[1] * spn[0];
    }
    cnt.add(14 + 1 + 12 + 60 * (int64_t)N, 6);
    double djac[6][6];
    double epl[6];
    double erepmd = 0.0;
    int iter = 0;
    bool more = true;
    while (more) {
        ++iter;
        cnt.outer += 1;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) sum = sum + S.b[i][k] * s[k];
            S.tau[i] = sum;
        }
        slpini(M, S, dt, io, cnt);
        scale_it(M, S, s, gamcap, C.afactor, cnt);
        slpfta(M, C, S, dt, io.err, cnt);
        scale_it(M, S, s, gamcap, C.afactor, cnt);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) sum = sum + S.b[k][i] * S.gam[k];
            epl[i] = sum;
        }
        double erreps[6];
        {
            double sm[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) sm[k] = s[k] - t[k];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                double sum = 0.0;
#pragma unroll
                for (int k = 0; k < 6; ++k) sum = sum + el[i][k] * sm[k];
                erreps[i] = io.de[i] - epl[i] - sum;
            }
        }
        if (factor<N>(S.ah)) io.err = io.err > 910 ? io.err : 910;
        // djac = B_a X^-1 B_a^T + S, one solve per column j
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double w[N];
#pragma unroll
            for (int k = 0; k < N; ++k) w[k] = S.act[k] ? S.b[k][j] : 0.0;
            solver<N>(S.ah, w);
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                double sum = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k)
                    if (S.act[k]) sum = sum + S.b[k][i] * w[k];
                djac[i][j] = sum;
            }
        }
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) djac[i][j] = djac[i][j] + el[i][j];
        if (factor<6>(djac)) io.err = io.err > 910 ? io.err : 910;
        solver<6>(djac, erreps);
        erepmd = fj2vom(true, erreps) / fj2vom(true, s);
        const double scale = 1.0 / (1.0 + C.alpha * erepmd);
#pragma unroll
        for (int i = 0; i < 6; ++i) erreps[i] = scale * erreps[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = s[i] + erreps[i];
        more = (iter < C.maxit) && (erepmd > C.toler);
        const int na = S.nsa;
        cnt.add(24 * (int64_t)N + 90 + 72 * (int64_t)na + 36 + 44 + factor_flops(na) + 6 * solver_flops(na) +
                    factor_flops(6) + solver_flops(6) + 28,
                0);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) srate_out[i] = fabs(S.gam[i] / dt);
    io.erepmd = erepmd;
    io.deff = fj2vom(false, epl);   // Sig_eps.f90:119, from the last pass's epl
    // ECHROT: lattice spin of the slips minus the rigid rotation, to Euler rates
    double rtn[3];
    {
        double rs[N][3];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double sdr[3], spn[3];
            slip_vectors(M, a, i, sdr, spn);
            rs[i][0] = (sdr[1] * spn[2] - sdr[2] * spn[1]) / 2.0;
            rs[i][1] = (sdr[2] * spn[0] - sdr[0] * spn[2]) / 2.0;
            rs[i][2] = (sdr[0] * spn[1] - sdr[1] * spn[0]) / 2.0;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) sum = sum + rs[k][q] * S.gam[k];
            rtn[q] = sum - io.r[q];
        }
    }
    double rx[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) sum = sum + a[i][k] * rtn[k];
        rx[i] = sum;
    }
    double dp1 = (fabs(aug[2]) < 1.e-5) ? rtn[2] : (aug[0] * rx[0] + aug[1] * rx[1]) / aug[2];
    const double dph = aug[1] * rx[0] - aug[0] * rx[1];
    const double dp2 = rx[2] - dp1 * a[2][2];
    // chi = djac^-1 from the last factors
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double c[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) c[k] = (k == i) ? 1.0 : 0.0;
        solver<6>(djac, c);
#pragma unroll
        for (int j = 0; j < 6; ++j) io.chi[j][i] = c[j];
    }
    dsv_io[0] = dp1;
    dsv_io[1] = dph;
    dsv_io[2] = dp2;
#pragma unroll
    for (int i = 0; i < N; ++i) dsv_io[3 + i] = S.s1[i] - S.s0[i];
    cnt.add(6 * (int64_t)N + 25 + 6 * solver_flops(6), 0);
}

// DSTRSS for one point.  s0_io: the N slip resistances sv(4:3+N);
// dsv_io: dsv(1:3+N) in/out; srate_io: N slip rates (kept when elastic).
template <int N>
__device__ __forceinline__ void dstrss(const DevMaterial<N>& M, const DevControl& C, double dt, PointIO& io,
                                       double* s0_io, double* dsv_io, double* srate_io, Count& cnt) {
    double el[6][6], s[6], t[6];
    elastc(M, io.sv0[0] + io.dsv0[0], io.sv0[1] + io.dsv0[1], io.sv0[2] + io.dsv0[2], el);
    cnt.add(1575 + 12, 6);
    tens_rot(io.r, io.st, t);
#pragma unroll
    for (int i = 0; i < 6; ++i) s[i] = t[i] + io.ds[i];
    const double sinit = fj2vom(true, s);
    const double einit = fj2vom(false, io.de);
    cnt.add(120 + 28, 0);
    io.erepmd = 0.0;
    if (io.iel > 0 && sinit > 0.0 && einit > 0.0) {
        cpstrs<N>(M, C, dt, el, s, t, io, s0_io, dsv_io, srate_io, cnt);
#pragma unroll
        for (int i = 0; i < 6; ++i) t[i] = s[i];
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = i; j < 6; ++j) {
                const double d = (io.chi[i][j] + io.chi[j][i]) / 2.0;
                io.chi[i][j] = d;
                io.chi[j][i] = d;
            }
        cnt.add(82, 0);
    } else {
        io.deff = 0.0;
        double lu[6][6];
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j) lu[i][j] = el[i][j];
        (void)factor<6>(lu);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double c[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) c[k] = (k == i) ? 1.0 : 0.0;
            solver<6>(lu, c);
#pragma unroll
            for (int j = 0; j < 6; ++j) io.chi[j][i] = c[j];
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) sum = sum + io.chi[i][k] * io.de[k];
            t[i] = sum;
        }
        cnt.add(factor_flops(6) + 6 * solver_flops(6) + 72 + 12, 0);
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        io.t[i] = t[i];
        io.ds[i] = t[i] - io.st[i];
    }
}

}  // namespace cpfem
