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
// Layout.  The slip-system matrices live in registers in the reference's own
// compacted form: the active systems isa(1:nsa) in index order, padded to a
// compile-time AMAX with identity rows, so the elimination and the solves do
// exactly the reference's operations on the active block.  Sums over all
// systems are taken over the active ones only (the others contribute exact
// zeros in the reference).  The Schmid columns of the active systems and the
// compliance go to LDS, one lane-interleaved slice per point.
//
// Scheduling.  Everything a CPSTRS outer iteration uses is rebuilt from the
// current stress s and fixed inputs, so a point can stop after any outer
// iteration with (s, iteration count) saved and resume later with identical
// results.  The kernel runs each point for at most iter_limit outer
// iterations; unfinished points go to a continuation list and points whose
// active set exceeds AMAX to an overflow list, served by later launches.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpfem {

// Material tables of one material set, FP64, in the layout the kernel reads.
// Slip-system counts with kernels of their own; every other count up to
// CPFEM_MAXSS runs in the kernels compiled for kNGeneric systems, whose loops
// over the systems take the material's nss at run time.
constexpr int kNGeneric = 60;

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
    int uniform_par;     // 1: all systems share ampar(1:6) (values of system 1 used)
    int two_valued;      // 1: smatrx = hd on the diagonal, hq off it
    int any_twin;        // 1: some |ssang| > 1.5
    int any_unidir;      // 1: some ssang != 0 (sign test needed)
    double hd, hq;
};

// Systems of a material: N, or nss for the generic kernels.
template <int N>
__device__ __forceinline__ int nsys(const DevMaterial<N>& M) {
    return N == kNGeneric ? M.nss : N;
}

struct DevControl {
    double toler, alpha, afactor;
    int maxit;
    int pad_;
    double stall_tol;   // 0: to maxit; > 0: stop at a stationary stress (cpfem_control::stall_tol)
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
    const double aa = fabs(a);
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

// Unroll count of loops over an active set of width A: whole for the register
// tiers, none for the wide generic tier, whose arrays live in scratch memory
// anyway and whose unrolled code would be far larger than the instruction cache.
template <int A>
constexpr int kUnrollA = A > 12 ? 1 : A;

// Doolittle LU without pivoting on an n x n register matrix w[row][col].
template <int n>
__device__ __forceinline__ bool factor(double (&w)[n][n]) {
    bool fail = false;
#pragma unroll kUnrollA<n>
    for (int k = 0; k < n - 1; ++k) {
        fail = fail || (w[k][k] == 0.0);
        const double piv = w[k][k];
#pragma unroll kUnrollA<n>
        for (int i = k + 1; i < n; ++i) {
            w[i][k] = w[i][k] / piv;
#pragma unroll kUnrollA<n>
            for (int j = k + 1; j < n; ++j) w[i][j] = w[i][j] - w[i][k] * w[k][j];
        }
    }
    fail = fail || (w[n - 1][n - 1] == 0.0);
    return fail;
}

template <int n>
__device__ __forceinline__ void solver(const double (&w)[n][n], double (&x)[n]) {
#pragma unroll kUnrollA<n>
    for (int i = 1; i < n; ++i)
#pragma unroll kUnrollA<n>
        for (int j = 0; j < i; ++j) x[i] = x[i] - w[i][j] * x[j];
    x[n - 1] = x[n - 1] / w[n - 1][n - 1];
#pragma unroll kUnrollA<n>
    for (int i = n - 2; i >= 0; --i) {
#pragma unroll kUnrollA<n>
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

// ELASTC: compliance in specimen axes, upper triangle el21[] in (i<=j) order.
template <int N>
__device__ __forceinline__ void elastc(const DevMaterial<N>& M, double ph1, double phi, double ph2, double* el21,
                                       int stride) {
    double a0[3][3], aug[3], a[3][3];
    ctmatrx(ph1, phi, ph2, a0, aug);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) a[i][j] = a0[j][i];
    const double* e = M.e;
    int q = 0;
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
            el21[(q++) * stride] = mult * (term1 + term2 + 0.5 * term3);
        }
    }
}

// index of (i, j) in the upper-triangle order of elastc
__device__ __forceinline__ constexpr int sym21(int i, int j) {
    return (i <= j) ? (i * 6 - i * (i - 1) / 2 + (j - i)) : (j * 6 - j * (j - 1) / 2 + (i - j));
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

// Slip direction and normal of system i in specimen axes (BTMATRX inner sums).
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

__device__ __forceinline__ void schmid(const double sdr[3], const double spn[3], double b[6]) {
    b[0] = sdr[0] * spn[0];
    b[1] = sdr[1] * spn[1];
    b[2] = sdr[2] * spn[2];
    b[3] = sdr[1] * spn[2] + sdr[2] * spn[1];
    b[4] = sdr[2] * spn[0] + sdr[0] * spn[2];
    b[5] = sdr[0] * spn[1] + sdr[1] * spn[0];
}

// Draw for the twin test of system sys in outer iteration outer (1-based) of
// a point whose stream key is key = (call epoch << 32) ^ point.  The
// reference draws RANDOM_NUMBER afresh at every check (crys-cons.f90:168-173);
// here every check consumes its own counter: the splitmix64 finaliser of the
// key, offset by the check's counter times the golden-ratio increment, mixed
// again.  Only consulted when twinprob < 1.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double twin_draw(uint64_t key, int outer, int sys) {
    const uint64_t h = mix64(key + 0x9E3779B97F4A7C15ull);
    const uint64_t z = mix64(h + ((uint64_t)outer * 64ull + (uint64_t)sys + 1ull) * 0x9E3779B97F4A7C15ull);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// per-point kernel arguments and the multi-pass schedule
// ---------------------------------------------------------------------------
struct ConstArgs {
    int n;                     // points
    int mesh;                  // 1: mesh points (material code of the element, error code
                               //    of the kinematics kernel in ck_kerr); 0: batch of points
    const int* gp_elem;        // [n] element of each point (mesh mode)
    const int* elem_le;        // [nel] material code (mesh mode)
    // per-point arrays, field-major with stride n
    const int* iel;            // batch mode material code
    const double* de;          // [6][n] strain increments (engineering shears)
    const double* r;           // [3][n] rotation increments
    double* dep;               // [n] deff out
    const double* st;          // [6][n]
    double* ds;                // [6][n] guess in, increment out
    const double* sv;          // [3+nss][n]
    double* dsv;               // [3+nss][n]
    double* t;                 // [6][n]
    double* chi;               // [36][n]
    double* erepmd;            // [n]
    double* srate;             // [nss][n]
    int* twinsys;              // [n]
    const int* timestwinned;   // [n]
    long long* flops;          // [n] accumulated over passes
    long long* transc;         // [n]
    int* err;                  // [n]
    int* tflag;                // scalar
    int* errmax;               // scalar
    uint64_t rng_epoch;
    const uint64_t* rng_epoch_dev;   // non-null: the epoch read on the device (graph capture)
    const double* dt_dev;            // non-null: the time step read on the device (graph capture)
    // schedule
    double* el;                // [21][n] compliance in specimen axes (work array)
    double* gp;                // [width][n] slips of the last hardening update (work array)
    double* ck_s;              // [6][n] stress at the next outer iteration
    int* ck_iter;              // [n] outer iterations done
    int* ck_kerr;              // [n] error code of the kinematics / earlier iterations
    int* ck_inner;             // [n] slip-Newton iterations so far
    int first;                 // 1: the first pass (every point, from the increment's inputs)
    const int* in_list;        // points of this launch (first pass: their order, null: 0..n-1)
    const int* in_count;       // device count of in_list
    int* cont_list;            // unfinished points
    int* cont_count;
    int* ovf_list;             // points whose active set exceeds the kernel's width
    int* ovf_count;
    int iter_limit;            // outer iterations allowed in total after this launch
    int* queue;                // non-null: persistent launch, next chunk of the list at *queue
};

// wave-aggregated append of this lane's point to a list
__device__ __forceinline__ void list_append(bool want, int g, int* list, int* count) {
    const unsigned long long m = __ballot(want);
    if (m == 0ull) return;
    const int nw = __popcll(m);
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, nw);
    base = __shfl(base, leader, 64);
    const int off = __popcll(m & ((1ull << lane) - 1ull));
    if (want) list[base + off] = g;
}

}  // namespace cpfem
