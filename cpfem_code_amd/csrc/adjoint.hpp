// Reverse-mode derivative (adjoint) of one point's constitutive update, FP64.
//
// The forward map is DSTRSS (Sig_eps.f90:84-151) for one load increment,
// followed by the commit of UPDATE_GPTV (Fasolt3.f90:254-284):
//   phi   = phi_prev + dg                 angles used by the call (sv + dsv)
//   S, B  = ELASTC(phi), BTMATRX(CTMATRX(phi))
//   t     = TENS_ROT(r, st_prev)
//   (s, gamma_A) solves the converged CPSTRS equations (Cpstrs.f90:98,
//   crys-cons.f90:69-77) on the active set A:
//     E1 = de - B^T gamma - S (s - t)                               = 0  (6)
//     E2_i = ln|tau_i| - ln s1_i(gamma) - m_i ln|gamma_i / dt|      = 0  (i in A)
//   with tau = B s and the hardening of PLASTC (ela-plas.f90:110-183)
//     s1_i = s0_i + sum_k L_ik theta_k |gamma_k|,
//     theta_k = th4 + th0 f(arg_k),  arg_k = max(0, 1 - s0_k / sst_k),
//     f(x) = x^alp,  sst_k = taus (sum|gamma| / dt)^nsat.
//   Outputs committed: stress st = s, slip resistances s0' = s1, angles
//   phi' = phi_prev + ECHROT(phi, gamma, r) (crys_math.f90:137-167).
// The elastic branch (Sig_eps.f90:130-145) gives s = S^-1 de and leaves the
// angles and slip resistances at phi_prev + dg and s0.
//
// The reverse step takes the cotangents of (st', phi', s0') and returns those
// of (st_prev, phi_prev, s0) and of the material parameters
//   pb[1] taus, pb[2] th0 (h0), pb[3] latent ratio q (off-diagonal of L),
//   pb[4] rate sensitivity m        (pb[0], tau0, is the sum of s0's cotangents
//                                    at the start of the load path).
// The converged equations are differentiated (implicit function theorem), so
// the result is the derivative of the solution, not of the iteration that
// reached it: the forward solve must be converged tightly for the two to agree.
// Orientation dependence (S, B, the spin terms of ECHROT) is carried by
// three-direction dual numbers.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "constitutive.hpp"

namespace cpfem {
namespace adj {

constexpr int kParams = 5;   // tau0, taus, h0, q, m

// value and derivatives along the three Euler angles
struct D3 {
    double v, d[3];
};
__host__ __device__ inline D3 dconst(double v) { return D3{v, {0.0, 0.0, 0.0}}; }
__host__ __device__ inline D3 operator+(const D3& a, const D3& b) {
    return D3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__host__ __device__ inline D3 operator-(const D3& a, const D3& b) {
    return D3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__host__ __device__ inline D3 operator-(const D3& a) { return D3{-a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
__host__ __device__ inline D3 operator*(const D3& a, const D3& b) {
    return D3{a.v * b.v,
              {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__host__ __device__ inline D3 operator*(double s, const D3& a) { return D3{s * a.v, {s * a.d[0], s * a.d[1], s * a.d[2]}}; }
__host__ __device__ inline D3 operator/(const D3& a, const D3& b) {
    const double q = a.v / b.v, ib = 1.0 / b.v;
    return D3{q, {(a.d[0] - q * b.d[0]) * ib, (a.d[1] - q * b.d[1]) * ib, (a.d[2] - q * b.d[2]) * ib}};
}
__host__ __device__ inline D3 dsin(const D3& a) {
    const double c = cos(a.v);
    return D3{sin(a.v), {c * a.d[0], c * a.d[1], c * a.d[2]}};
}
__host__ __device__ inline D3 dsqrt(const D3& a) {
    const double r = sqrt(a.v), h = 0.5 / r;
    return D3{r, {h * a.d[0], h * a.d[1], h * a.d[2]}};
}
__host__ __device__ inline D3 dcos(const D3& a) {
    const double s = -sin(a.v);
    return D3{cos(a.v), {s * a.d[0], s * a.d[1], s * a.d[2]}};
}
__host__ __device__ inline double val(double x) { return x; }
__host__ __device__ inline double val(const D3& x) { return x.v; }
__host__ __device__ inline double tsin(double x) { return sin(x); }
__host__ __device__ inline double tcos(double x) { return cos(x); }
__host__ __device__ inline D3 tsin(const D3& x) { return dsin(x); }
__host__ __device__ inline D3 tcos(const D3& x) { return dcos(x); }

// CTMATRX (crys_math.f90:25-62): a(3,3) crystal <- specimen rows, aug.
template <class T>
__host__ __device__ void ctm(const T& p1, const T& ph, const T& p2, T a[3][3], T aug[3]) {
    const T c1 = tcos(p1), s1 = tsin(p1), c2 = tcos(p2), s2 = tsin(p2), c = tcos(ph), s = tsin(ph);
    a[0][0] = c1 * c2 - s1 * s2 * c;
    a[1][0] = -(c1 * s2) - s1 * c2 * c;
    a[2][0] = s1 * s;
    a[0][1] = s1 * c2 + c1 * s2 * c;
    a[1][1] = -(s1 * s2) + c1 * c2 * c;
    a[2][1] = -(c1 * s);
    a[0][2] = s2 * s;
    a[1][2] = c2 * s;
    a[2][2] = c;
    aug[0] = s2;
    aug[1] = c2;
    aug[2] = s;
}

// ELASTC (ela-plas.f90:12-86): compliance in specimen axes, engineering
// shears, from the crystal compliances e(1:9) and a0 = CTMATRX.
template <class T>
__host__ __device__ void elastc_t(const double e[9], const T a0[3][3], T S[6][6]) {
    for (int iset = 0; iset < 6; ++iset) {
        const int i = iset < 3 ? iset : (iset - 2) % 3, j = iset < 3 ? iset : (i + 1) % 3;
        const double mult = iset < 3 ? 1.0 : 2.0;
        for (int jset = iset; jset < 6; ++jset) {
            const int k = jset < 3 ? jset : (jset - 2) % 3, l = jset < 3 ? jset : (k + 1) % 3;
            // a[p][q] = a0[q][p]
            T acc{};
            for (int q = 0; q < 3; ++q)   // diagonal crystal compliances S11 S22 S33
                acc = acc + e[q] * (a0[q][i] * a0[q][j] * a0[q][k] * a0[q][l]);
            // S23 (e4), S31 (e5), S12 (e6) couple normal components q != p
            const int pq[3][2] = {{1, 2}, {2, 0}, {0, 1}};
            for (int c = 0; c < 3; ++c) {
                const int p = pq[c][0], q = pq[c][1];
                acc = acc + e[3 + c] * (a0[p][i] * a0[p][j] * a0[q][k] * a0[q][l] +
                                        a0[q][i] * a0[q][j] * a0[p][k] * a0[p][l]);
            }
            // shear compliances S44 (23), S55 (31), S66 (12), half weight
            for (int c = 0; c < 3; ++c) {
                const int p = pq[c][0], q = pq[c][1];
                acc = acc + (0.5 * e[6 + c]) * (a0[p][i] * a0[q][j] * a0[p][k] * a0[q][l] +
                                                a0[q][i] * a0[p][j] * a0[q][k] * a0[p][l]);
            }
            S[iset][jset] = mult * acc;
        }
    }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < i; ++j) S[i][j] = S[j][i];
}

// BTMATRX (crys_math.f90:75-115): Schmid columns (engineering shears) and
// spin columns of every system from a0 = CTMATRX.
template <class T, int N, int NM>
__host__ __device__ void btm(const DevMaterial<N>& M, int nss, const T a0[3][3], T b[NM][6], T rot[NM][3]) {
    for (int i = 0; i < nss; ++i) {
        T d[3], n[3];
        for (int j = 0; j < 3; ++j) {
            d[j] = M.ab[i][0] * a0[0][j] + M.ab[i][1] * a0[1][j] + M.ab[i][2] * a0[2][j];
            n[j] = M.an[i][0] * a0[0][j] + M.an[i][1] * a0[1][j] + M.an[i][2] * a0[2][j];
        }
        b[i][0] = d[0] * n[0];
        b[i][1] = d[1] * n[1];
        b[i][2] = d[2] * n[2];
        b[i][3] = d[1] * n[2] + d[2] * n[1];
        b[i][4] = d[2] * n[0] + d[0] * n[2];
        b[i][5] = d[0] * n[1] + d[1] * n[0];
        rot[i][0] = 0.5 * (d[1] * n[2] - d[2] * n[1]);
        rot[i][1] = 0.5 * (d[2] * n[0] - d[0] * n[2]);
        rot[i][2] = 0.5 * (d[0] * n[1] - d[1] * n[0]);
    }
}

// ECHROT's map from the lattice spin rtn (specimen axes) to Euler-angle
// increments (crys_math.f90:152-164), linear in rtn for fixed angles.
template <class T>
__host__ __device__ void echrot_core(const T a[3][3], const T aug[3], const T rtn[3], T out[3]) {
    T rx[3];
    for (int i = 0; i < 3; ++i) rx[i] = a[i][0] * rtn[0] + a[i][1] * rtn[1] + a[i][2] * rtn[2];
    T dp1;
    if (fabs(val(aug[2])) < 1.e-5)
        dp1 = rtn[2];
    else
        dp1 = (aug[0] * rx[0] + aug[1] * rx[1]) / aug[2];
    out[0] = dp1;
    out[1] = aug[1] * rx[0] - aug[0] * rx[1];
    out[2] = rx[2] - dp1 * a[2][2];
}

// TENS_ROT (Sig_eps.f90:166-202): first-order rotation of a symmetric tensor
// (tensor shears) by the small rotation r, new = rot^T old rot.
__host__ __device__ inline void tens_rot_h(const double r[3], const double a[6], double b[6]) {
    double rm[3][3];
    rm[0][0] = sqrt(1.0 - r[1] * r[1] - r[2] * r[2]);
    rm[1][1] = sqrt(1.0 - r[0] * r[0] - r[2] * r[2]);
    rm[2][2] = sqrt(1.0 - r[0] * r[0] - r[1] * r[1]);
    rm[0][1] = r[2];
    rm[1][0] = -r[2];
    rm[0][2] = -r[1];
    rm[2][0] = r[1];
    rm[1][2] = r[0];
    rm[2][1] = -r[0];
    const double o[3][3] = {{a[0], a[5], a[4]}, {a[5], a[1], a[3]}, {a[4], a[3], a[2]}};
    double w[3][3], nw[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) w[i][j] = o[i][0] * rm[0][j] + o[i][1] * rm[1][j] + o[i][2] * rm[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) nw[i][j] = rm[0][i] * w[0][j] + rm[1][i] * w[1][j] + rm[2][i] * w[2][j];
    b[0] = nw[0][0];
    b[1] = nw[1][1];
    b[2] = nw[2][2];
    b[3] = nw[1][2];
    b[4] = nw[2][0];
    b[5] = nw[0][1];
}

// TENS_ROT with the rotation carried as dual numbers (derivative in r).
__host__ __device__ inline void tens_rot_d(const D3 r[3], const double a[6], D3 b[6]) {
    D3 rm[3][3];
    const D3 one = dconst(1.0);
    rm[0][0] = dsqrt(one - r[1] * r[1] - r[2] * r[2]);
    rm[1][1] = dsqrt(one - r[0] * r[0] - r[2] * r[2]);
    rm[2][2] = dsqrt(one - r[0] * r[0] - r[1] * r[1]);
    rm[0][1] = r[2];
    rm[1][0] = -r[2];
    rm[0][2] = -r[1];
    rm[2][0] = r[1];
    rm[1][2] = r[0];
    rm[2][1] = -r[0];
    const double o[3][3] = {{a[0], a[5], a[4]}, {a[5], a[1], a[3]}, {a[4], a[3], a[2]}};
    D3 w[3][3], nw[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) w[i][j] = o[i][0] * rm[0][j] + o[i][1] * rm[1][j] + o[i][2] * rm[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) nw[i][j] = rm[0][i] * w[0][j] + rm[1][i] * w[1][j] + rm[2][i] * w[2][j];
    b[0] = nw[0][0];
    b[1] = nw[1][1];
    b[2] = nw[2][2];
    b[3] = nw[1][2];
    b[4] = nw[2][0];
    b[5] = nw[0][1];
}

__host__ __device__ inline double fj2vom_h(bool lstress, const double x[6]) {
    const double amult = lstress ? 6.0 : 1.5;
    double sum = 0.0;
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        sum = sum + (x[i] - x[j]) * (x[i] - x[j]) + amult * x[i + 3] * x[i + 3];
    }
    return sqrt(sum / 4.5);
}

// Solve A^T x = b in place (b -> x) by LU with partial pivoting; A is n x n
// in a[ld][ld] and is destroyed.  Returns 1 on a zero pivot.
template <int LD>
__host__ __device__ int solve_transposed(int n, double a[LD][LD], double b[LD]) {
    // transpose in place, then an ordinary pivoted elimination
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double t = a[i][j];
            a[i][j] = a[j][i];
            a[j][i] = t;
        }
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(a[k][k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[i][k]) > best) {
                best = fabs(a[i][k]);
                p = i;
            }
        if (best == 0.0) return 1;
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                const double t = a[k][j];
                a[k][j] = a[p][j];
                a[p][j] = t;
            }
            const double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        const double inv = 1.0 / a[k][k];
        for (int i = k + 1; i < n; ++i) {
            const double f = a[i][k] * inv;
            if (f == 0.0) continue;
            for (int j = k + 1; j < n; ++j) a[i][j] -= f * a[k][j];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= a[i][j] * b[j];
        b[i] = s / a[i][i];
    }
    return 0;
}

// Hardening quantities of PLASTC at (s0, gamma): theta, f(arg), f'(arg), sst,
// the factor of sst in taus, and d sst / d sum|gamma|.
template <int N, int NM>
struct Hard {
    double theta[NM], f[NM], fp[NM], sst[NM], pw[NM], dsum[NM];
    double s1[NM];
};

template <int N, int NM>
__host__ __device__ void hardening(const DevMaterial<N>& M, int nss, double dt, const double s0[NM],
                                   const double gam[NM], Hard<N, NM>& H) {
    double sumg = 0.0;
    for (int k = 0; k < nss; ++k) sumg += fabs(gam[k]);
    const double x = sumg / dt;
    for (int k = 0; k < nss; ++k) {
        const double a6 = M.nsat[k];
        const double pw = a6 == 0.0 ? 1.0 : pow(x, a6);
        const double sst = M.taus[k] * pw;
        double arg = sst != 0.0 ? 1.0 - s0[k] / sst : 0.0;
        const bool live = arg > 0.0;
        if (!live) arg = 0.0;
        const double a4 = M.alp[k];
        H.f[k] = live ? (a4 == 1.0 ? arg : pow(arg, a4)) : 0.0;
        H.fp[k] = live ? (a4 == 1.0 ? 1.0 : a4 * pow(arg, a4 - 1.0)) : 0.0;
        H.theta[k] = M.th4[k] + M.th0[k] * H.f[k];
        H.sst[k] = sst;
        H.pw[k] = pw;
        H.dsum[k] = (a6 != 0.0 && x > 0.0) ? M.taus[k] * a6 * pow(x, a6 - 1.0) / dt : 0.0;
    }
    for (int i = 0; i < nss; ++i) {
        double s = s0[i];
        for (int k = 0; k < nss; ++k) s += M.smatrx[i][k] * H.theta[k] * fabs(gam[k]);
        H.s1[i] = s;
    }
}

// Pull a cotangent cot[i] on s1_i back to s0 (added to s0b), to the
// parameters (taus, h0, q) and, when gb is given, to gamma.
template <int N, int NM>
__host__ __device__ void pull_s1(const DevMaterial<N>& M, int nss, const double s0[NM], const double gam[NM],
                                 const Hard<N, NM>& H, const double cot[NM], double s0b[NM], double pb[kParams],
                                 double* gb) {
    double sumgb = 0.0;
    for (int k = 0; k < nss; ++k) {
        s0b[k] += cot[k];
        double c = 0.0;
        for (int i = 0; i < nss; ++i) c += cot[i] * M.smatrx[i][k];
        const double ag = fabs(gam[k]);
        const double thb = c * ag;
        if (gb && gam[k] != 0.0) gb[k] += c * H.theta[k] * (gam[k] > 0.0 ? 1.0 : -1.0);
        pb[2] += thb * H.f[k];
        const double argb = thb * M.th0[k] * H.fp[k];
        if (H.fp[k] != 0.0 && H.sst[k] != 0.0) {
            s0b[k] += -argb / H.sst[k];
            const double sstb = argb * s0[k] / (H.sst[k] * H.sst[k]);
            pb[1] += sstb * H.pw[k];
            sumgb += sstb * H.dsum[k];
        }
    }
    if (gb)
        for (int k = 0; k < nss; ++k)
            if (gam[k] != 0.0) gb[k] += sumgb * (gam[k] > 0.0 ? 1.0 : -1.0);
    // latent ratio: d s1_i / d q = sum_{k != i} theta_k |gamma_k|
    double tot = 0.0;
    for (int k = 0; k < nss; ++k) tot += H.theta[k] * fabs(gam[k]);
    for (int i = 0; i < nss; ++i) pb[3] += cot[i] * (tot - H.theta[i] * fabs(gam[i]));
}

// Inputs of one point's increment (start state, recorded data, forward outputs).
template <int NM>
struct PointIn {
    double st[6], php[3], s0[NM];     // start of the increment
    double de[6], r[3], dg[3], dsg[6];// strain and rotation increments, angle and stress guesses
    double so[6], srate[NM];          // forward outputs: new stress, slip rates |gamma|/dt
    int iel;                          // material code of the element (<= 0: elastic)
};

// One reverse step.  sb (cotangent of the new stress) is consumed; phb and s0b
// hold the cotangents of the new angles / slip resistances on entry and of the
// previous ones on exit; stb receives the cotangent of the start stress, deb
// and rb those of the strain and rotation increments; pb accumulates the
// parameters.  Returns 1 if the plastic branch ran, 0 for the elastic branch,
// -1 on a singular Jacobian.
template <int N, int NM>
__host__ __device__ int point_reverse(const DevMaterial<N>& M, int nss, double dt, const PointIn<NM>& P,
                                      const double sb_in[6], double phb[3], double s0b[NM], double stb[6],
                                      double pb[kParams], double deb[6], double rb[3]) {
    // geometry at the angles of the call, with their Euler-angle derivatives
    const double phi[3] = {P.php[0] + P.dg[0], P.php[1] + P.dg[1], P.php[2] + P.dg[2]};
    const D3 p1{phi[0], {1.0, 0.0, 0.0}}, p2{phi[1], {0.0, 1.0, 0.0}}, p3{phi[2], {0.0, 0.0, 1.0}};
    D3 a[3][3], aug[3], S[6][6];
    ctm(p1, p2, p3, a, aug);
    elastc_t(M.e, a, S);
    double t[6];
    tens_rot_h(P.r, P.st, t);
    double sinit[6];
    for (int i = 0; i < 6; ++i) sinit[i] = t[i] + P.dsg[i];
    const bool plastic = P.iel > 0 && fj2vom_h(true, sinit) > 0.0 && fj2vom_h(false, P.de) > 0.0;
    for (int i = 0; i < 6; ++i) stb[i] = deb[i] = 0.0;
    for (int i = 0; i < 3; ++i) rb[i] = 0.0;
    double phg[3] = {0.0, 0.0, 0.0};   // cotangent of the angles of the call
    if (!plastic) {
        // s = S^-1 de: d s / d phi_k = -S^-1 (dS/dphi_k) s
        double Sm[6][6], u[6];
        for (int i = 0; i < 6; ++i) {
            u[i] = sb_in[i];
            for (int j = 0; j < 6; ++j) Sm[i][j] = S[i][j].v;
        }
        if (solve_transposed<6>(6, Sm, u)) return -1;
        for (int i = 0; i < 6; ++i) deb[i] = u[i];   // s = S^-1 de
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) acc += u[i] * S[i][j].d[k] * P.so[j];
            phg[k] = -acc;
        }
        for (int k = 0; k < 3; ++k) phb[k] += phg[k];
        return 0;
    }
    D3 b[NM][6], rot[NM][3];
    btm<D3, N, NM>(M, nss, a, b, rot);
    double tau[NM], gam[NM], gb[NM];
    int act[NM];
    int na = 0;
    for (int i = 0; i < nss; ++i) {
        double s = 0.0;
        for (int c = 0; c < 6; ++c) s += b[i][c].v * P.so[c];
        tau[i] = s;
        gb[i] = 0.0;
        if (P.srate[i] > 0.0) {
            gam[i] = copysign(P.srate[i] * dt, s);
            act[na++] = i;
        } else {
            gam[i] = 0.0;
        }
    }
    Hard<N, NM> H;
    hardening<N, NM>(M, nss, dt, P.s0, gam, H);
    double s0b_out[NM];
    for (int k = 0; k < nss; ++k) s0b_out[k] = 0.0;
    // (1) angles: phi' = phi_prev + ECHROT(phi, gamma, r)
    const double dpb[3] = {phb[0], phb[1], phb[2]};
    {
        D3 rtn[3], out[3];
        for (int c = 0; c < 3; ++c) {
            D3 s = dconst(-P.r[c]);
            for (int k = 0; k < na; ++k) s = s + gam[act[k]] * rot[act[k]][c];
            rtn[c] = s;
        }
        echrot_core(a, aug, rtn, out);
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) phg[k] += dpb[i] * out[i].d[k];
        // linear part in the spin: rtn_bar = J^T dpb
        double av[3][3], augv[3], rtb[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < 3; ++i) {
            augv[i] = aug[i].v;
            for (int j = 0; j < 3; ++j) av[i][j] = a[i][j].v;
        }
        for (int j = 0; j < 3; ++j) {
            double e[3] = {0.0, 0.0, 0.0}, o[3];
            e[j] = 1.0;
            echrot_core<double>(av, augv, e, o);
            rtb[j] = dpb[0] * o[0] + dpb[1] * o[1] + dpb[2] * o[2];
        }
        for (int k = 0; k < na; ++k) {
            const int i = act[k];
            gb[i] += rot[i][0].v * rtb[0] + rot[i][1].v * rtb[1] + rot[i][2].v * rtb[2];
        }
        for (int c = 0; c < 3; ++c) rb[c] -= rtb[c];   // rtn = rot gamma - r
    }
    // (2) slip resistances s0' = s1(s0, gamma)
    pull_s1<N, NM>(M, nss, P.s0, gam, H, s0b, s0b_out, pb, gb);
    // (3) the converged equations: F_y^T lambda = ybar, y = (s, gamma_A)
    constexpr int LD = NM + 6;
    double F[LD][LD], lam[LD];
    const int ny = 6 + na;
    for (int r = 0; r < 6; ++r) {
        for (int c = 0; c < 6; ++c) F[r][c] = -S[r][c].v;
        for (int jj = 0; jj < na; ++jj) F[r][6 + jj] = -b[act[jj]][r].v;
        lam[r] = sb_in[r];
    }
    for (int ii = 0; ii < na; ++ii) {
        const int i = act[ii];
        for (int c = 0; c < 6; ++c) F[6 + ii][c] = b[i][c].v / tau[i];
        // d s1_i / d sum|gamma| through sst (zero for a rate-independent saturation)
        double hs = 0.0;
        for (int k = 0; k < nss; ++k)
            if (H.fp[k] != 0.0 && H.sst[k] != 0.0)
                hs += M.smatrx[i][k] * fabs(gam[k]) * M.th0[k] * H.fp[k] * P.s0[k] / (H.sst[k] * H.sst[k]) * H.dsum[k];
        for (int jj = 0; jj < na; ++jj) {
            const int j = act[jj];
            const double sg = gam[j] > 0.0 ? 1.0 : -1.0;
            const double G = M.smatrx[i][j] * H.theta[j] * sg + hs * sg;
            F[6 + ii][6 + jj] = -G / H.s1[i] - (ii == jj ? M.am[i] / gam[i] : 0.0);
        }
        lam[6 + ii] = gb[i];
    }
    if (solve_transposed<LD>(ny, F, lam)) return -1;
    // (4) parameters and inputs of F: xbar = -(dF/dx)^T lambda
    double tb[6];
    for (int c = 0; c < 6; ++c) {
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += S[r][c].v * lam[r];
        tb[c] = -s;
    }
    double sm[6];
    for (int c = 0; c < 6; ++c) sm[c] = P.so[c] - t[c];
    for (int k = 0; k < 3; ++k) {
        double acc = 0.0;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) acc += lam[r] * sm[c] * S[r][c].d[k];
        // Schmid columns: E1 through -b_j gamma_j, E2_i through ln|b_i s|
        for (int jj = 0; jj < na; ++jj) {
            const int j = act[jj];
            const double mu = lam[6 + jj] / tau[j];
            for (int c = 0; c < 6; ++c) acc += (lam[c] * gam[j] - mu * P.so[c]) * b[j][c].d[k];
        }
        phg[k] += acc;
    }
    double cot[NM];
    for (int k = 0; k < nss; ++k) cot[k] = 0.0;
    for (int ii = 0; ii < na; ++ii) {
        const int i = act[ii];
        cot[i] = lam[6 + ii] / H.s1[i];
        pb[4] += lam[6 + ii] * log(fabs(gam[i] / dt));
    }
    pull_s1<N, NM>(M, nss, P.s0, gam, H, cot, s0b_out, pb, nullptr);
    // strain increment: E1 = de - ...
    for (int i = 0; i < 6; ++i) deb[i] = -lam[i];
    // rotation increment through t = TENS_ROT(r, st)
    {
        const D3 rd[3] = {D3{P.r[0], {1.0, 0.0, 0.0}}, D3{P.r[1], {0.0, 1.0, 0.0}}, D3{P.r[2], {0.0, 0.0, 1.0}}};
        D3 td[6];
        tens_rot_d(rd, P.st, td);
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            for (int c = 0; c < 6; ++c) acc += tb[c] * td[c].d[k];
            rb[k] += acc;
        }
    }
    // (5) start stress through t = TENS_ROT(r, st)
    for (int j = 0; j < 6; ++j) {
        double e[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, o[6];
        e[j] = 1.0;
        tens_rot_h(P.r, e, o);
        double s = 0.0;
        for (int c = 0; c < 6; ++c) s += tb[c] * o[c];
        stb[j] = s;
    }
    for (int k = 0; k < nss; ++k) s0b[k] = s0b_out[k];
    for (int k = 0; k < 3; ++k) phb[k] += phg[k];
    return 1;
}

}  // namespace adj
}  // namespace cpfem
