// Element-level device routines shared by the point and assembly kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace cpfem {

// Local-coordinate signs of the brick nodes (node order of the reference
// preprocessor, src/prepro/fs3pre.f90:193-196); the 2x2x2 points use the same
// order (src/fasolt/Shapef.f90:203-205).
__device__ __constant__ static const signed char kBrick[8][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};

// INVERT_3 (Shapef.f90:499-536): adjugate inverse in place, m[row][col].
__device__ __forceinline__ double invert_3(double m[3][3]) {
    double l[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) l[i][j] = m[i][j];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int ia = (i + 1) % 3, ib = (ia + 1) % 3;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int ja = (j + 1) % 3, jb = (ja + 1) % 3;
            m[j][i] = l[ia][ja] * l[ib][jb] - l[ia][jb] * l[ib][ja];
        }
    }
    const double det = m[0][0] * l[0][0] + m[0][1] * l[1][0] + m[0][2] * l[2][0];
    if (det != 0.0) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) m[i][j] = m[i][j] / det;
    }
    return det;
}

// Shape-function derivatives of the 8-node brick at point ig.  lt = 12:
// one centre point, weight 8 detJ (SHAPE_12, Shapef.f90:32-87); lt = 13:
// 2x2x2 points at +-1/sqrt(3), weight detJ.  sfd[a][k] = dN_a/dx_k.
__device__ __forceinline__ double brick_shape(int lt, int ig, const double xl[8][3], double sfd[8][3]) {
    const double g = 1.0 / sqrt(3.0);
    double xi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) xi[k] = (lt == 12) ? 0.0 : (double)kBrick[ig][k] * g;
    double dndn[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double f[3], d[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            f[k] = 1.0 + (double)kBrick[a][k] * xi[k];
            d[k] = (double)kBrick[a][k];
        }
        dndn[a][0] = 0.125 * d[0] * f[1] * f[2];
        dndn[a][1] = 0.125 * f[0] * d[1] * f[2];
        dndn[a][2] = 0.125 * f[0] * f[1] * d[2];
    }
    double jac[3][3];  // jac[j][k] = sum_a dndn(j,a) x(k,a)
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < 8; ++a) s = s + dndn[a][j] * xl[a][k];
            jac[j][k] = s;
        }
    const double det = invert_3(jac);
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) s = s + jac[k][j] * dndn[a][j];
            sfd[a][k] = s;
        }
    return (lt == 12) ? 8.0 * det : det;
}

// DSTRNS (Sig_eps.f90:25-61): strain (eng. shear) and rotation increments.
// Returns the FAFERR code of the magnitude traps (905 rotation, 906 strain).
template <int NN>
__device__ __forceinline__ int dstrns(const double sfd[NN][3], const double dxl[NN][3], double smale, double smalr,
                                      double del[6], double rl[3], double& deff) {
    double d[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int n = 0; n < NN; ++n) s = s + sfd[n][i] * dxl[n][j];
            d[i][j] = s;
        }
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        del[i] = d[i][i];
        del[i + 3] = d[k][j] + d[j][k];
        rl[i] = (d[k][j] - d[j][k]) / 2.0;
        sum = sum + rl[i] * rl[i];
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        s = s + (del[i] - del[j]) * (del[i] - del[j]) + 1.5 * del[i + 3] * del[i + 3];
    }
    deff = sqrt(s / 4.5);
    int err = 0;
    if (sum > smalr) err = 905;
    if (deff > smale) err = 906;
    return err;
}

// Node-pair stiffness block of one point (Elvals.f90:132-221): BiT D Bj
// plus the stress terms, for shape derivatives a (node i), b (node j).
template <typename T>
__device__ __forceinline__ void pair_block(const T a[3], const T b[3], const T* D, const T s[6], T gm[3][3]) {
#define Dm(p, q) D[((q)-1) * 6 + ((p)-1)]
#define Sm(p) s[(p)-1]
    const T a1 = a[0], a2 = a[1], a3 = a[2], b1 = b[0], b2 = b[1], b3 = b[2];
    const T h = (T)2, q = (T)4;
    gm[0][0] = (a1 * b1 * (Dm(1, 1) - Sm(1)) + a2 * b2 * (Dm(6, 6) - (Sm(1) - Sm(2)) / h) +
                a3 * b3 * (Dm(5, 5) - (Sm(1) - Sm(3)) / h) + a2 * b3 * (Dm(5, 6) + Sm(4) / h) + a3 * b1 * (Dm(1, 5)) +
                a1 * b2 * (Dm(1, 6)) + b2 * a3 * (Dm(5, 6) + Sm(4) / h) + b3 * a1 * (Dm(1, 5)) + b1 * a2 * (Dm(1, 6)));
    gm[1][0] = (a1 * b1 * (Dm(1, 6) - Sm(6)) + a2 * b2 * (Dm(2, 6) - Sm(6)) + a3 * b3 * (Dm(4, 5) - Sm(6) / h) +
                a2 * b3 * (Dm(2, 5) - Sm(5) / q) + a3 * b1 * (Dm(1, 4) - Sm(4) / q) +
                a1 * b2 * (Dm(6, 6) - (Sm(1) + Sm(2)) / q) + b2 * a3 * (Dm(4, 6) - Sm(5) / q) +
                b3 * a1 * (Dm(5, 6) - Sm(4) / q) + b1 * a2 * (Dm(1, 2) - (Sm(1) + Sm(2)) / q));
    gm[2][0] = (a1 * b1 * (Dm(1, 5) - Sm(5)) + a2 * b2 * (Dm(4, 6) - Sm(5) / h) + a3 * b3 * (Dm(3, 5) - Sm(5)) +
                a2 * b3 * (Dm(4, 5) - Sm(6) / q) + a3 * b1 * (Dm(1, 3) - (Sm(1) + Sm(3)) / q) +
                a1 * b2 * (Dm(5, 6) - Sm(4) / q) + b2 * a3 * (Dm(3, 6) - Sm(6) / q) +
                b3 * a1 * (Dm(5, 5) - (Sm(1) + Sm(3)) / q) + b1 * a2 * (Dm(1, 4) - Sm(4) / q));
    gm[0][1] = (b1 * a1 * (Dm(1, 6) - Sm(6)) + b2 * a2 * (Dm(2, 6) - Sm(6)) + b3 * a3 * (Dm(4, 5) - Sm(6) / h) +
                b2 * a3 * (Dm(2, 5) - Sm(5) / q) + b3 * a1 * (Dm(1, 4) - Sm(4) / q) +
                b1 * a2 * (Dm(6, 6) - (Sm(1) + Sm(2)) / q) + a2 * b3 * (Dm(4, 6) - Sm(5) / q) +
                a3 * b1 * (Dm(5, 6) - Sm(4) / q) + a1 * b2 * (Dm(1, 2) - (Sm(1) + Sm(2)) / q));
    gm[1][1] = (a1 * b1 * (Dm(6, 6) - (Sm(2) - Sm(1)) / h) + a2 * b2 * (Dm(2, 2) - Sm(2)) +
                a3 * b3 * (Dm(4, 4) - (Sm(2) - Sm(3)) / h) + a2 * b3 * (Dm(2, 4)) + a3 * b1 * (Dm(4, 6) + Sm(5) / h) +
                a1 * b2 * (Dm(2, 6)) + b2 * a3 * (Dm(2, 4)) + b3 * a1 * (Dm(4, 6) + Sm(5) / h) + b1 * a2 * (Dm(2, 6)));
    gm[2][1] = (a1 * b1 * (Dm(5, 6) - Sm(4) / h) + a2 * b2 * (Dm(2, 4) - Sm(4)) + a3 * b3 * (Dm(3, 4) - Sm(4)) +
                a2 * b3 * (Dm(4, 4) - (Sm(2) + Sm(3)) / q) + a3 * b1 * (Dm(3, 6) - Sm(6) / q) +
                a1 * b2 * (Dm(2, 5) - Sm(5) / q) + b2 * a3 * (Dm(2, 3) - (Sm(2) + Sm(3)) / q) +
                b3 * a1 * (Dm(4, 5) - Sm(6) / q) + b1 * a2 * (Dm(4, 6) - Sm(5) / q));
    gm[0][2] = (b1 * a1 * (Dm(1, 5) - Sm(5)) + b2 * a2 * (Dm(4, 6) - Sm(5) / h) + b3 * a3 * (Dm(3, 5) - Sm(5)) +
                b2 * a3 * (Dm(4, 5) - Sm(6) / q) + b3 * a1 * (Dm(1, 3) - (Sm(1) + Sm(3)) / q) +
                b1 * a2 * (Dm(5, 6) - Sm(4) / q) + a2 * b3 * (Dm(3, 6) - Sm(6) / q) +
                a3 * b1 * (Dm(5, 5) - (Sm(1) + Sm(3)) / q) + a1 * b2 * (Dm(1, 4) - Sm(4) / q));
    gm[1][2] = (b1 * a1 * (Dm(5, 6) - Sm(4) / h) + b2 * a2 * (Dm(2, 4) - Sm(4)) + b3 * a3 * (Dm(3, 4) - Sm(4)) +
                b2 * a3 * (Dm(4, 4) - (Sm(2) + Sm(3)) / q) + b3 * a1 * (Dm(3, 6) - Sm(6) / q) +
                b1 * a2 * (Dm(2, 5) - Sm(5) / q) + a2 * b3 * (Dm(2, 3) - (Sm(2) + Sm(3)) / q) +
                a3 * b1 * (Dm(4, 5) - Sm(6) / q) + a1 * b2 * (Dm(4, 6) - Sm(5) / q));
    gm[2][2] = (a1 * b1 * (Dm(5, 5) - (Sm(3) - Sm(1)) / h) + a2 * b2 * (Dm(4, 4) - (Sm(3) - Sm(2)) / h) +
                a3 * b3 * (Dm(3, 3) - Sm(3)) + a2 * b3 * (Dm(3, 4)) + a3 * b1 * (Dm(3, 5)) +
                a1 * b2 * (Dm(4, 5) + Sm(6) / h) + b2 * a3 * (Dm(3, 4)) + b3 * a1 * (Dm(3, 5)) +
                b1 * a2 * (Dm(4, 5) + Sm(6) / h));
#undef Dm
#undef Sm
}

}  // namespace cpfem
