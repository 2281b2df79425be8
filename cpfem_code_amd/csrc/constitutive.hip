// Constitutive kernels, their multi-pass schedule and the batch entry point.
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "constitutive.hpp"
#include "constitutive_launch.hpp"
#include "kernels.hpp"

namespace cpfem {

template <int N>
__device__ __forceinline__ double rate_m(const DevMaterial<N>& M, int i) {
    return M.am[M.uniform_par ? 0 : i];
}

template <int N>
__device__ __forceinline__ double latent(const DevMaterial<N>& M, int i, int j) {
    if (M.two_valued) return (i == j) ? M.hd : M.hq;
    return M.smatrx[i][j];
}

// Compacted slip state of one point (module FAS_CPL scratch of the reference).
template <int AMAX>
struct Compact {
    int nsa;
    int isa[AMAX];
    double gam[AMAX], tau[AMAX], s0[AMAX], s1[AMAX];
    double ah[AMAX][AMAX];
};

// PLASTC on the active set: theta at s0, s1 and the log-hardening matrix.
// Slots >= nsa of ah are identity rows / columns.  th[] receives theta.
template <int N, int AMAX>
__device__ __forceinline__ void plastc_c(const DevMaterial<N>& M, Compact<AMAX>& S, double dt, double th[AMAX],
                                         int64_t& flops, int64_t& transc) {
    const int nsa = S.nsa;
    double sumss = 0.0;
#pragma unroll kUnrollA<AMAX>
    for (int k = 0; k < AMAX; ++k)
        if (k < nsa) sumss = sumss + fabs(S.gam[k]);
    const double rate = sumss / dt;
#pragma unroll kUnrollA<AMAX>
    for (int k = 0; k < AMAX; ++k) {
        if (k < nsa) {
            const int j = M.uniform_par ? 0 : S.isa[k];
            const double sst = M.taus[j] * pow_fast(rate, M.nsat[j]);
            double arg = (sst != 0.0) ? 1.0 - (S.s0[k] / sst) : 0.0;
            if (arg < 0.0) arg = 0.0;
            th[k] = M.th4[j] + M.th0[j] * copysign(pow_fast(fabs(arg), M.alp[j]), arg);
        } else {
            th[k] = 0.0;
        }
    }
#pragma unroll kUnrollA<AMAX>
    for (int k = 0; k < AMAX; ++k) {
        if (k < nsa) {
            double s = S.s0[k];
#pragma unroll kUnrollA<AMAX>
            for (int l = 0; l < AMAX; ++l)
                if (l < nsa) s = s + (latent(M, S.isa[k], S.isa[l]) * th[l]) * fabs(S.gam[l]);
            S.s1[k] = s;
        }
    }
#pragma unroll kUnrollA<AMAX>
    for (int k = 0; k < AMAX; ++k) {
#pragma unroll kUnrollA<AMAX>
        for (int l = 0; l < AMAX; ++l) {
            double v = (k == l) ? 1.0 : 0.0;
            if (k < nsa && l < nsa) {
                // gam/|gam| of ela-plas.f90:186-197 is exactly +-1: copysign gives the same value
                const double ag = fabs(S.gam[l]);
                v = (ag > 1.e-14) ? (copysign(1.0, S.gam[l]) * (latent(M, S.isa[k], S.isa[l]) * th[l])) / S.s1[k]
                                  : 0.0;
            }
            S.ah[k][l] = v;
        }
    }
    const int ns = nsys(M);
    flops += 3 * (int64_t)ns * ns + 6 * ns + 3 * (int64_t)nsa * nsa;
    transc += 4 * ns;
}

// SCALE_IT on the active set; amax starts at am of system 1 (Cpstrs.f90:200).
template <int N, int AMAX>
__device__ __forceinline__ void scale_it_c(const DevMaterial<N>& M, Compact<AMAX>& S, double s[6], double gamcap,
                                           double afactor, int64_t& flops, int64_t& transc) {
    double amax = M.am[0];
    double gammax = fabs(S.gam[0]);
#pragma unroll kUnrollA<AMAX>
    for (int k = 1; k < AMAX; ++k) {
        if (k < S.nsa && fabs(S.gam[k]) > gammax) {
            gammax = fabs(S.gam[k]);
            amax = rate_m(M, S.isa[k]);
        }
    }
    if (gammax > gamcap) {
        const double scale = gamcap / gammax;
        double smult = spower(0.9 * scale, afactor * amax);
        if (smult == 0.0) smult = 0.004 / amax;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = smult * s[i];
#pragma unroll kUnrollA<AMAX>
        for (int k = 0; k < AMAX; ++k) {
            S.gam[k] = scale * S.gam[k];
            S.tau[k] = smult * S.tau[k];
        }
        flops += S.nsa + 2 * nsys(M) + 6;
        transc += 2;
    } else {
        flops += S.nsa;
    }
}

// SLPFTA on the active set; leaves ah in X form and gpre = slips of the last PLASTC.
template <int N, int AMAX>
__device__ __forceinline__ int slpfta_c(const DevMaterial<N>& M, const DevControl& C, Compact<AMAX>& S, double dt,
                                        double* gpre, size_t gstride, int& err, int64_t& flops, int64_t& transc) {
    const int nsa = S.nsa;
    double th[AMAX];
    int it = 0;
    bool more = true;
    double tauerr = 0.0;
    while (more) {
        ++it;
        plastc_c<N, AMAX>(M, S, dt, th, flops, transc);
#pragma unroll kUnrollA<AMAX>
        for (int i = 0; i < AMAX; ++i) {
            if (i < nsa) {
                const double gamn = S.gam[i];
#pragma unroll kUnrollA<AMAX>
                for (int j = 0; j < AMAX; ++j)
                    if (j < nsa) S.ah[j][i] = gamn * S.ah[j][i];
                S.ah[i][i] = S.ah[i][i] + rate_m(M, S.isa[i]);
            }
        }
        if (factor<AMAX>(S.ah)) err = err > 910 ? err : 910;
        double e[AMAX];
#pragma unroll kUnrollA<AMAX>
        for (int k = 0; k < AMAX; ++k) {
            double v = 0.0;
            if (k < nsa) {
                const double taufg = S.s1[k] * spower((S.gam[k] / dt), rate_m(M, S.isa[k]));
                if ((S.tau[k] * taufg) > 0.0) v = log(S.tau[k] / taufg);
            }
            e[k] = v;
        }
        tauerr = 0.0;
#pragma unroll kUnrollA<AMAX>
        for (int k = 0; k < AMAX; ++k)
            if (k < nsa) tauerr = tauerr + e[k] * e[k];
        tauerr = sqrt(tauerr);
        solver<AMAX>(S.ah, e);
        double errmax = 0.0;
#pragma unroll kUnrollA<AMAX>
        for (int k = 0; k < AMAX; ++k)
            if (k < nsa && fabs(e[k]) > errmax) errmax = fabs(e[k]);
        const double scale = (errmax > 5.0) ? 5.0 / errmax : 1.0;
#pragma unroll kUnrollA<AMAX>
        for (int k = 0; k < AMAX; ++k)
            if (k < nsa) S.gam[k] = S.gam[k] * exp(scale * e[k]);
        more = (it < C.maxit) && (tauerr > M.conv);
        flops += (int64_t)nsa * nsa + 10 * nsa + 1 + factor_flops(nsa) + solver_flops(nsa);
        transc += 4 * nsa;
    }
    if (tauerr > 0.5) err = err > 987 ? err : 987;
    plastc_c<N, AMAX>(M, S, dt, th, flops, transc);
#pragma unroll kUnrollA<AMAX>
    for (int k = 0; k < AMAX; ++k) gpre[k * gstride] = S.gam[k];
#pragma unroll kUnrollA<AMAX>
    for (int i = 0; i < AMAX; ++i) {
        if (i < nsa) {
#pragma unroll kUnrollA<AMAX>
            for (int j = 0; j < AMAX; ++j)
                if (j < nsa) S.ah[j][i] = S.tau[j] * S.ah[j][i];
            const double ag = fabs(S.gam[i]);
            const double am = rate_m(M, S.isa[i]);
            S.ah[i][i] = S.ah[i][i] + ((ag > 1.e-12) ? am * S.tau[i] / S.gam[i] : am / 1.e-12);
        }
    }
    flops += (int64_t)nsa * nsa + 3 * nsa;
    return it;
}

// The per-point kernel.  One lane per point, 64 lanes per block.
template <int N, int AMAX>
__global__ __launch_bounds__(64) void k_const(const DevMaterial<N>* __restrict__ Mp, const DevControl C,
                                              const double dt_arg, const ConstArgs A) {
    const double dt = A.dt_dev ? *A.dt_dev : dt_arg;
    // LDS: the Schmid columns of the active systems (also the staging area of
    // the active-set compaction) and the active system numbers, interleaved by
    // lane; the compliance and the last slips live in per-point work arrays
    constexpr int kSlots = 6 * AMAX > 2 * N ? 6 * AMAX : 2 * N;
    __shared__ double sh_b[kSlots][64];
    __shared__ int sh_isa[AMAX][64];
    const int lane = threadIdx.x;
    const bool first = A.first != 0;
    // Points of the launch in chunks of 64, one per lane: the workgroup's own
    // chunk (blockIdx) and, when the launch is persistent (A.queue non-null,
    // a grid of resident workgroups), every further chunk taken from the
    // queue counter in list order as soon as the wave has finished its last.
    int base = blockIdx.x * 64;
    for (;;) {
        const int idx = base + lane;
        bool live;
        int g = 0;
        if (first) {
            // every point, in the order of in_list when one is given
            live = idx < A.n;
            if (live) g = A.in_list ? A.in_list[idx] : idx;
        } else {
            live = idx < *A.in_count;
            if (live) g = A.in_list[idx];
        }
        bool want_cont = false, want_ovf = false;
        if (live) {
            // material record of the point: |material code| (checked on the host).
            // The kernels of a fixed count serve one material set, whose record is
            // read through wave-uniform (scalar) loads; the generic ones index it
            // per lane.
            const int iel = A.mesh ? A.elem_le[A.gp_elem[g]] : A.iel[g];
            const DevMaterial<N>& M = N == kNGeneric ? Mp[abs(iel) - 1] : Mp[0];
            const int NS = nsys(M);
            const size_t n = (size_t)A.n;
            double* el = A.el + g;
            double* bc = &sh_b[0][lane];
            double* gpre = A.gp + g;
            double de[6], r[3], st[6], t[6], s[6];
            // strain and rotation increments come from the kinematics kernel (mesh
            // mode) or the caller (batch mode); in mesh mode that kernel's error
            // code (920, 905, 906) seeds the point's code
            int kerr = 0;
    #pragma unroll
            for (int i = 0; i < 6; ++i) de[i] = A.de[i * n + g];
    #pragma unroll
            for (int i = 0; i < 3; ++i) r[i] = A.r[i * n + g];
            if (!first || A.mesh) kerr = A.ck_kerr[g];
    #pragma unroll
            for (int i = 0; i < 6; ++i) st[i] = A.st[i * n + g];
            double ph[3];
    #pragma unroll
            for (int i = 0; i < 3; ++i) ph[i] = A.sv[i * n + g] + A.dsv[i * n + g];
            elastc<N>(M, ph[0], ph[1], ph[2], el, (int)n);
            tens_rot(r, st, t);
            int64_t flops = 0, transc = 0;
            int iter = 0, inner = 0;   // outer and slip-Newton iterations so far
            bool elastic = false;
            if (first) {
                flops = 1575 + 12 + 120 + 28;
                transc = 6;
    #pragma unroll
                for (int i = 0; i < 6; ++i) s[i] = t[i] + A.ds[i * n + g];
                const double sinit = fj2vom(true, s);
                const double einit = fj2vom(false, de);
                elastic = !(iel > 0 && sinit > 0.0 && einit > 0.0);
            } else {
    #pragma unroll
                for (int i = 0; i < 6; ++i) s[i] = A.ck_s[i * n + g];
                iter = A.ck_iter[g];
                inner = A.ck_inner[g];
                flops = A.flops[g];
                transc = A.transc[g];
            }
            if (elastic) {
                // Sig_eps.f90:130-145: chi = S^-1, t = chi de (the start stress is not used)
                double lu[6][6], chi[6][6];
    #pragma unroll
                for (int i = 0; i < 6; ++i)
    #pragma unroll
                    for (int j = 0; j < 6; ++j) lu[i][j] = el[sym21(i, j) * n];
                (void)factor<6>(lu);
    #pragma unroll
                for (int i = 0; i < 6; ++i) {
                    double c[6];
    #pragma unroll
                    for (int k = 0; k < 6; ++k) c[k] = (k == i) ? 1.0 : 0.0;
                    solver<6>(lu, c);
    #pragma unroll
                    for (int j = 0; j < 6; ++j) chi[j][i] = c[j];
                }
    #pragma unroll
                for (int i = 0; i < 6; ++i) {
                    double sum = 0.0;
    #pragma unroll
                    for (int k = 0; k < 6; ++k) sum = sum + chi[i][k] * de[k];
                    A.t[i * n + g] = sum;
                    A.ds[i * n + g] = sum - st[i];
                }
    #pragma unroll
                for (int j = 0; j < 6; ++j)
    #pragma unroll
                    for (int i = 0; i < 6; ++i) A.chi[(i + 6 * j) * n + g] = chi[i][j];
                A.dep[g] = 0.0;
                A.erepmd[g] = 0.0;
                flops += factor_flops(6) + 6 * solver_flops(6) + 72 + 12;
                A.flops[g] = flops;
                A.transc[g] = transc;
                A.err[g] = kerr;
                if (kerr) atomicMax(A.errmax, kerr);
            } else {
                // ---------------- CPSTRS (Cpstrs.f90:34-172) ----------------
                double a[3][3], aug[3];
                ctmatrx(ph[0], ph[1], ph[2], a, aug);
                const double gamcap = 6.0 * fj2vom(false, de);
                if (first) {
                    flops += 14 + 1 + 12 + 60 * (int64_t)NS;
                    transc += 6;
                }
                int twinsys = A.twinsys[g];
                const int timestwinned = A.timestwinned[g];
                const uint64_t rng_key = ((A.rng_epoch_dev ? *A.rng_epoch_dev : A.rng_epoch) << 32) ^ (uint64_t)g;
                int err = kerr, tflag = 0;
                Compact<AMAX> S;
                double djac[6][6], epl[6];
                double erepmd = 0.0;
                bool done = false, ovf = false;
                while (!done && iter < A.iter_limit) {
                    double s_top[6];   // stress at the top of this outer iteration (stall test)
    #pragma unroll
                    for (int i = 0; i < 6; ++i) s_top[i] = s[i];
                    // ---- SLPINI (crys-cons.f90:143-197) over all systems.  The
                    // loop stays rolled (one power per system) and stages each
                    // system's slip and resolved shear in the LDS slice of the
                    // Schmid columns, slots 2i and 2i+1; the active ones are then
                    // compacted in place (write slot <= read slot) to 2k, 2k+1.
                    double gammax = 0.0;
    #pragma unroll 1
                    for (int i = 0; i < NS; ++i) {
                        double sdr[3], spn[3], b[6];
                        slip_vectors(M, a, i, sdr, spn);
                        schmid(sdr, spn, b);
                        double sum = 0.0;
    #pragma unroll
                        for (int k = 0; k < 6; ++k) sum = sum + b[k] * s[k];
                        const double s0i = A.sv[(3 + i) * n + g];
                        double gi;
                        if (M.any_twin && fabs(M.ssang[i]) > 1.5) {
                            gi = 0.0;
                            if ((sum > s0i) && (timestwinned != M.twinlim) && (twinsys == 0)) {
                                if (twin_draw(rng_key, iter + 1, i) < M.twinprob) {
                                    tflag = 1;
                                    twinsys = i + 1;
                                }
                            }
                        } else if (M.any_unidir && M.ssang[i] * sum / s0i < 0.0) {
                            gi = 0.0;
                        } else {
                            gi = dt * spower((sum / s0i), (1.0 / M.am[i]));
                        }
                        bc[(2 * i + 0) * 64] = gi;
                        bc[(2 * i + 1) * 64] = sum;
                        if (fabs(gi) > gammax) gammax = fabs(gi);
                    }
                    const double cutoff = 1.e-6 * gammax;
                    int nsa = 0;
    #pragma unroll 1
                    for (int i = 0; i < NS; ++i) {
                        const double gi = bc[(2 * i + 0) * 64];
                        if (!(fabs(gi) < cutoff)) {
                            if (nsa < AMAX) {
                                const double ti = bc[(2 * i + 1) * 64];
                                bc[(2 * nsa + 0) * 64] = gi;
                                bc[(2 * nsa + 1) * 64] = ti;
                                sh_isa[nsa][lane] = i;
                            }
                            ++nsa;
                        }
                    }
                    if (nsa > AMAX) {
                        ovf = true;
                        break;
                    }
                    ++iter;
                    S.nsa = nsa;
    #pragma unroll kUnrollA<AMAX>
                    for (int k = 0; k < AMAX; ++k) {
                        if (k < nsa) {
                            const int i = sh_isa[k][lane];
                            S.isa[k] = i;
                            S.gam[k] = bc[(2 * k + 0) * 64];
                            S.tau[k] = bc[(2 * k + 1) * 64];
                            S.s0[k] = A.sv[(3 + i) * n + g];
                        } else {
                            S.isa[k] = 0;
                            S.gam[k] = 0.0;
                            S.tau[k] = 0.0;
                            S.s0[k] = 1.0;
                        }
                    }
                    flops += 5 * (int64_t)NS;
                    transc += 2 * (int64_t)NS;
                    // Schmid columns of the active systems (BTMATRX, crys_math.f90:75-115)
    #pragma unroll 1
                    for (int k = 0; k < nsa; ++k) {
                        double sdr[3], spn[3], b[6];
                        slip_vectors(M, a, sh_isa[k][lane], sdr, spn);
                        schmid(sdr, spn, b);
    #pragma unroll
                        for (int q = 0; q < 6; ++q) bc[(6 * k + q) * 64] = b[q];
                    }
                    scale_it_c<N, AMAX>(M, S, s, gamcap, C.afactor, flops, transc);
                    inner += slpfta_c<N, AMAX>(M, C, S, dt, gpre, n, err, flops, transc);
                    scale_it_c<N, AMAX>(M, S, s, gamcap, C.afactor, flops, transc);
                    // plastic strain increment and strain-increment error (Cpstrs.f90:96-98)
    #pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        double sum = 0.0;
    #pragma unroll kUnrollA<AMAX>
                        for (int k = 0; k < AMAX; ++k)
                            if (k < nsa) sum = sum + bc[(6 * k + i) * 64] * S.gam[k];
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
                            for (int k = 0; k < 6; ++k) sum = sum + el[sym21(i, k) * n] * sm[k];
                            erreps[i] = A.de[i * n + g] - epl[i] - sum;
                        }
                    }
                    if (factor<AMAX>(S.ah)) err = err > 910 ? err : 910;
                    // djac = B_a X^-1 B_a^T + S: one solve per column (Cpstrs.f90:101-119)
    #pragma unroll
                    for (int j = 0; j < 6; ++j) {
                        double w[AMAX];
    #pragma unroll kUnrollA<AMAX>
                        for (int k = 0; k < AMAX; ++k) w[k] = (k < nsa) ? bc[(6 * k + j) * 64] : 0.0;
                        solver<AMAX>(S.ah, w);
    #pragma unroll
                        for (int i = 0; i < 6; ++i) {
                            double sum = 0.0;
    #pragma unroll kUnrollA<AMAX>
                            for (int k = 0; k < AMAX; ++k)
                                if (k < nsa) sum = sum + bc[(6 * k + i) * 64] * w[k];
                            djac[i][j] = sum + el[sym21(i, j) * n];
                        }
                    }
                    if (factor<6>(djac)) err = err > 910 ? err : 910;
                    solver<6>(djac, erreps);
                    erepmd = fj2vom(true, erreps) / fj2vom(true, s);
                    const double scale = 1.0 / (1.0 + C.alpha * erepmd);
    #pragma unroll
                    for (int i = 0; i < 6; ++i) erreps[i] = scale * erreps[i];
    #pragma unroll
                    for (int i = 0; i < 6; ++i) s[i] = s[i] + erreps[i];
                    done = !((iter < C.maxit) && (erepmd > C.toler));
                    if (!done && C.stall_tol > 0.0) {
                        // opt-in: the stress no longer moves (a fixed point of SCALE_IT's
                        // cap and the Newton step); the remaining iterations would only
                        // change it at the rounding level
                        double dmax = 0.0, smax = 0.0;
    #pragma unroll
                        for (int i = 0; i < 6; ++i) {
                            if (fabs(s[i] - s_top[i]) > dmax) dmax = fabs(s[i] - s_top[i]);
                            if (fabs(s[i]) > smax) smax = fabs(s[i]);
                        }
                        if (dmax <= C.stall_tol * smax) done = true;
                    }
                    flops += 24 * (int64_t)NS + 90 + 72 * (int64_t)nsa + 36 + 44 + factor_flops(nsa) +
                             6 * solver_flops(nsa) + factor_flops(6) + solver_flops(6) + 28;
                }
                if (tflag) atomicOr(A.tflag, 1);
                if (done) {
                    // ---------------- finish CPSTRS and DSTRSS ----------------
                    const int nsa = S.nsa;
    #pragma unroll
                    for (int i = 0; i < NS; ++i) A.srate[i * n + g] = 0.0;
    #pragma unroll kUnrollA<AMAX>
                    for (int k = 0; k < AMAX; ++k)
                        if (k < nsa) A.srate[S.isa[k] * n + g] = fabs(S.gam[k] / dt);
                    // ECHROT (crys_math.f90:137-167)
                    double rtn[3] = {0.0, 0.0, 0.0};
    #pragma unroll kUnrollA<AMAX>
                    for (int k = 0; k < AMAX; ++k) {
                        if (k < nsa) {
                            double sdr[3], spn[3];
                            slip_vectors(M, a, S.isa[k], sdr, spn);
                            rtn[0] = rtn[0] + ((sdr[1] * spn[2] - sdr[2] * spn[1]) / 2.0) * S.gam[k];
                            rtn[1] = rtn[1] + ((sdr[2] * spn[0] - sdr[0] * spn[2]) / 2.0) * S.gam[k];
                            rtn[2] = rtn[2] + ((sdr[0] * spn[1] - sdr[1] * spn[0]) / 2.0) * S.gam[k];
                        }
                    }
    #pragma unroll
                    for (int q = 0; q < 3; ++q) rtn[q] = rtn[q] - A.r[q * n + g];
                    double rx[3];
    #pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        double sum = 0.0;
    #pragma unroll
                        for (int k = 0; k < 3; ++k) sum = sum + a[i][k] * rtn[k];
                        rx[i] = sum;
                    }
                    const double dp1 = (fabs(aug[2]) < 1.e-5) ? rtn[2] : (aug[0] * rx[0] + aug[1] * rx[1]) / aug[2];
                    const double dph = aug[1] * rx[0] - aug[0] * rx[1];
                    const double dp2 = rx[2] - dp1 * a[2][2];
                    A.dsv[0 * n + g] = dp1;
                    A.dsv[1 * n + g] = dph;
                    A.dsv[2 * n + g] = dp2;
                    // dsv(3+i) = s1(i) - s0(i), s1 of the last PLASTC over every system
                    {
                        double th[AMAX];
                        double sumss = 0.0;
    #pragma unroll kUnrollA<AMAX>
                        for (int k = 0; k < AMAX; ++k)
                            if (k < nsa) sumss = sumss + fabs(gpre[k * n]);
                        const double rate = sumss / dt;
    #pragma unroll kUnrollA<AMAX>
                        for (int k = 0; k < AMAX; ++k) {
                            th[k] = 0.0;
                            if (k < nsa) {
                                const int j = M.uniform_par ? 0 : S.isa[k];
                                const double sst = M.taus[j] * pow_fast(rate, M.nsat[j]);
                                double arg = (sst != 0.0) ? 1.0 - (S.s0[k] / sst) : 0.0;
                                if (arg < 0.0) arg = 0.0;
                                th[k] = M.th4[j] + M.th0[j] * copysign(pow_fast(fabs(arg), M.alp[j]), arg);
                            }
                        }
    #pragma unroll
                        for (int i = 0; i < NS; ++i) {
                            const double s0i = A.sv[(3 + i) * n + g];
                            double s1 = s0i;
    #pragma unroll kUnrollA<AMAX>
                            for (int l = 0; l < AMAX; ++l)
                                if (l < nsa) s1 = s1 + (latent(M, i, S.isa[l]) * th[l]) * fabs(gpre[l * n]);
                            A.dsv[(3 + i) * n + g] = s1 - s0i;
                        }
                    }
                    // chi = djac^-1 (Cpstrs.f90:154-161), symmetrised (Sig_eps.f90:122-128)
                    double chi[6][6];
    #pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        double c[6];
    #pragma unroll
                        for (int k = 0; k < 6; ++k) c[k] = (k == i) ? 1.0 : 0.0;
                        solver<6>(djac, c);
    #pragma unroll
                        for (int j = 0; j < 6; ++j) chi[j][i] = c[j];
                    }
    #pragma unroll
                    for (int i = 0; i < 6; ++i)
    #pragma unroll
                        for (int j = i; j < 6; ++j) {
                            const double d = (chi[i][j] + chi[j][i]) / 2.0;
                            chi[i][j] = d;
                            chi[j][i] = d;
                        }
    #pragma unroll
                    for (int j = 0; j < 6; ++j)
    #pragma unroll
                        for (int i = 0; i < 6; ++i) A.chi[(i + 6 * j) * n + g] = chi[i][j];
    #pragma unroll
                    for (int i = 0; i < 6; ++i) {
                        A.t[i * n + g] = s[i];
                        A.ds[i * n + g] = s[i] - A.st[i * n + g];
                    }
                    A.dep[g] = fj2vom(false, epl);
                    A.erepmd[g] = erepmd;
                    flops += 6 * (int64_t)NS + 25 + 6 * solver_flops(6) + 82;
                    A.flops[g] = flops;
                    A.transc[g] = transc;
                    A.err[g] = err;
                    A.twinsys[g] = twinsys;
                    if (err) atomicMax(A.errmax, err);
                } else {
                    // checkpoint: stress at the next outer iteration and the work so far
    #pragma unroll
                    for (int i = 0; i < 6; ++i) A.ck_s[i * n + g] = s[i];
                    A.ck_iter[g] = iter;
                    A.ck_inner[g] = inner;
                    A.ck_kerr[g] = err;
                    A.flops[g] = flops;
                    A.transc[g] = transc;
                    A.twinsys[g] = twinsys;
                    want_ovf = ovf;
                    want_cont = !ovf;
                }
            }
        }
        list_append(want_cont, g, A.cont_list, A.cont_count);
        list_append(want_ovf, g, A.ovf_list, A.ovf_count);
        if (!A.queue) break;
        int next = 0;
        if (lane == 0) next = atomicAdd(A.queue, 64);
        base = __shfl(next, 0, 64);
        if (base >= (first ? A.n : *A.in_count)) break;
    }
}

template <int N>
void fill_material(DevMaterial<N>& d, const cpfem_material& m, double toler) {
    memset(&d, 0, sizeof(d));
    d.nss = m.nss;
    d.twinlim = m.twinlim;
    d.twinprob = m.twinprob;
    for (int i = 0; i < 9; ++i) d.e[i] = m.e[i];
    double amx = m.ampar[0][0];
    d.uniform_par = 1;
    d.two_valued = 1;
    d.hd = m.smatrx[0][0];
    const int ns = N == kNGeneric ? m.nss : N;
    d.hq = ns > 1 ? m.smatrx[0][1] : 0.0;
    for (int i = 0; i < ns; ++i) {
        d.ssang[i] = m.ssang[i];
        if (fabs(m.ssang[i]) > 1.5) d.any_twin = 1;
        if (m.ssang[i] != 0.0) d.any_unidir = 1;
        for (int k = 0; k < 3; ++k) {
            d.ab[i][k] = m.ab[i][k];
            d.an[i][k] = m.an[i][k];
        }
        d.am[i] = m.ampar[i][0];
        d.th0[i] = m.ampar[i][1];
        d.th4[i] = m.ampar[i][2];
        d.alp[i] = m.ampar[i][3];
        d.taus[i] = m.ampar[i][4];
        d.nsat[i] = m.ampar[i][5];
        for (int q = 0; q < 6; ++q)
            if (m.ampar[i][q] != m.ampar[0][q]) d.uniform_par = 0;
        for (int k = 0; k < ns; ++k) {
            d.smatrx[i][k] = m.smatrx[k][i];  // Fortran smatrx(i,k)
            const double want = (i == k) ? d.hd : d.hq;
            if (d.smatrx[i][k] != want) d.two_valued = 0;
        }
        if (m.ampar[i][0] > amx) amx = m.ampar[i][0];
    }
    d.conv = toler * amx;
}

DevControl make_control(const cpfem_control& c) {
    DevControl d;
    d.toler = c.toler;
    d.alpha = c.alpha;
    d.afactor = c.afactor;
    d.maxit = c.maxit;
    d.pad_ = 0;
    d.stall_tol = c.stall_tol;
    return d;
}

size_t material_bytes(int nss) {
    switch (nss) {
        case 12: return sizeof(DevMaterial<12>);
        case 18: return sizeof(DevMaterial<18>);
        default: return (nss >= 1 && nss <= kNGeneric) ? sizeof(DevMaterial<kNGeneric>) : 0;
    }
}

void build_material(int nss, void* host, const cpfem_material& m, double toler) {
    if (nss == 12)
        fill_material<12>(*reinterpret_cast<DevMaterial<12>*>(host), m, toler);
    else if (nss == 18)
        fill_material<18>(*reinterpret_cast<DevMaterial<18>*>(host), m, toler);
    else
        fill_material<kNGeneric>(*reinterpret_cast<DevMaterial<kNGeneric>*>(host), m, toler);
}

// Width of the first kernel (active systems held in registers) and the
// outer iterations every point gets before the unfinished ones are compacted.
#ifndef CPFEM_AMAX
#define CPFEM_AMAX 6
#endif
constexpr int kAmax = CPFEM_AMAX;
constexpr int kAmaxMid = 8;
constexpr int kAmaxWide = 12;
constexpr int kAmaxGeneric = kConstMaxWidth;   // fourth tier, generic system counts only

__global__ void k_zero_n(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero4(int* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

__global__ void k_set1(int* p, int v) {
    if (threadIdx.x == 0) *p = v;
}

// Ordering of a list of unfinished points before the pass that resumes them:
// a point's slip-Newton iteration count per outer iteration barely changes
// from one outer iteration to the next, so points are grouped by the mean so
// far (descending, in eighths of an iteration): the lanes of a wave then run
// inner loops of similar length, and the longest waves start first.  A
// counting sort over kSortBins bins; the order within a bin is arbitrary and
// changes no result (every point is updated independently).
constexpr int kSortBins = 1024;

// Sort key of a point: with flops (the work of the point's previous update,
// for the pass over every point) an eighth-octave bin of it, which is also
// saved in save_bin; with prev_bin that saved bin; otherwise the mean
// slip-Newton iterations per outer iteration so far in eighths.  Larger work
// sorts first.
struct SortKey {
    const int* ck_inner;
    const int* ck_iter;
    const long long* flops;
    const int* prev_bin;
    int* save_bin;
    __device__ __forceinline__ int bin(int g) const {
        if (prev_bin) return prev_bin[g];
        int m;
        if (flops) {
            const unsigned long long f = flops[g] > 0 ? (unsigned long long)flops[g] : 1ull;
            const int e = 63 - __clzll(f);
            const int frac = e >= 3 ? (int)((f >> (e - 3)) & 7ull) : (int)((f << (3 - e)) & 7ull);
            m = 8 * e + frac;
        } else {
            const int it = ck_iter[g] > 1 ? ck_iter[g] : 1;
            m = (8 * ck_inner[g]) / it;
        }
        return kSortBins - 1 - (m < kSortBins - 1 ? m : kSortBins - 1);
    }
};

// Bin counts: per block in LDS, then one device atomic per occupied bin
// (most points share a few bins, whose device counters would otherwise take
// one atomic per point).
__global__ __launch_bounds__(256) void k_sort_count(const int* __restrict__ list, int count, const int* count_dev,
                                                    const SortKey K, int* bins) {
    __shared__ int sh[kSortBins];
    for (int b = threadIdx.x; b < kSortBins; b += blockDim.x) sh[b] = 0;
    __syncthreads();
    if (count_dev) count = *count_dev;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) atomicAdd(&sh[K.bin(list ? list[i] : i)], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < kSortBins; b += blockDim.x)
        if (sh[b]) atomicAdd(&bins[b], sh[b]);
}

// exclusive prefix sum of the kSortBins counts, one block of kSortBins threads
__global__ __launch_bounds__(kSortBins) void k_sort_scan(int* bins) {
    __shared__ int sh[kSortBins];
    const int t = threadIdx.x;
    const int v = bins[t];
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kSortBins; off <<= 1) {
        const int add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    bins[t] = sh[t] - v;
}

// Scatter into bin order: each block reserves its run of every occupied bin
// with one device atomic, then ranks its points within the run in LDS.
__global__ __launch_bounds__(256) void k_sort_scatter(const int* __restrict__ list, int count, const int* count_dev,
                                                      const SortKey K, int* bins, int* __restrict__ out) {
    __shared__ int cnt[kSortBins];
    __shared__ int base[kSortBins];
    for (int b = threadIdx.x; b < kSortBins; b += blockDim.x) cnt[b] = 0;
    __syncthreads();
    if (count_dev) count = *count_dev;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < count;
    int g = 0, key = 0;
    if (live) {
        g = list ? list[i] : i;
        key = K.bin(g);
        if (K.save_bin) K.save_bin[g] = key;
        atomicAdd(&cnt[key], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSortBins; b += blockDim.x) {
        if (cnt[b]) base[b] = atomicAdd(&bins[b], cnt[b]);
        cnt[b] = 0;
    }
    __syncthreads();
    if (live) out[base[key] + atomicAdd(&cnt[key], 1)] = g;
}

// Sorted copy of the n-point list (null: points 0..n-1) into S.sorted, keyed
// by the previous updates' work (flops non-null; the bins are kept in
// S.prev_bin), by those kept bins (by_prev) or by the checkpoints.
// count_dev: the list's length read on the device (n is then the largest it can be).
static hipError_t sort_points(ConstSchedule& S, const int* list, int n, const long long* flops, hipStream_t stream,
                              bool by_prev = false, const int* count_dev = nullptr) {
    const int nb = (n + 255) / 256;
    const SortKey K{S.ck_inner, S.ck_iter, flops, by_prev ? S.prev_bin : nullptr, flops ? S.prev_bin : nullptr};
    hipLaunchKernelGGL(k_zero_n, dim3(kSortBins / 256), dim3(256), 0, stream, S.bins, kSortBins);
    hipLaunchKernelGGL(k_sort_count, dim3(nb), dim3(256), 0, stream, list, n, count_dev, K, S.bins);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(kSortBins), 0, stream, S.bins);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nb), dim3(256), 0, stream, list, n, count_dev, K, S.bins, S.sorted);
    return hipGetLastError();
}

// Workgroups of k_const<N, AMAX> that the device holds at once (0 if unknown).
template <int N, int AMAX>
static int resident_blocks() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_const<N, AMAX>, 64, 0) != hipSuccess)
            cus = per_cu = 0;
        cached = cus * per_cu;
    }
    return cached;
}

// One pass over nthreads list entries (an upper bound when the count is on
// the device).  When that needs more workgroups than the device holds at once,
// the launch is persistent: a grid of resident workgroups that take the list's
// chunks from the queue counter S.counts[4] (set here to the grid's own
// chunks), so that each wave starts its next chunk as soon as it finishes one.
template <int N, int AMAX>
static void launch_one(const void* dmat, const DevControl& C, double dt, ConstArgs A, int nthreads,
                       const ConstSchedule& S, hipStream_t stream) {
    if (nthreads <= 0) return;
    int grid = (nthreads + 63) / 64;
    A.queue = nullptr;
    const int cap = S.queue ? resident_blocks<N, AMAX>() : 0;
    if (cap > 0 && cap < grid) {
        grid = cap;
        A.queue = S.counts + 4;
        hipLaunchKernelGGL(k_set1, dim3(1), dim3(64), 0, stream, A.queue, grid * 64);
    }
    hipLaunchKernelGGL((k_const<N, AMAX>), dim3(grid), dim3(64), 0, stream,
                       reinterpret_cast<const DevMaterial<N>*>(dmat), C, dt, A);
}

template <int N>
static hipError_t read_counts(ConstSchedule& S, hipStream_t stream) {
    hipError_t e = hipMemcpyAsync(S.host_counts, S.counts, 4 * sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

// Tiers of active-set width: kAmax for every point, then kAmaxMid and
// kAmaxWide for the points whose active set outgrew the previous tier.
template <int N>
static hipError_t schedule(const void* dmat, const DevControl& C, double dt, ConstArgs A, ConstSchedule& S,
                           hipStream_t stream, int counts[4]) {
    // device counters: [0] list A, [1] list B, [2] overflow list, [3] overflow of the overflow
    hipLaunchKernelGGL(k_zero4, dim3(1), dim3(64), 0, stream, S.counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    A.ck_s = S.ck_s;
    A.ck_iter = S.ck_iter;
    A.ck_inner = S.ck_inner;
    A.ck_kerr = S.ck_kerr;
    A.el = S.el;
    A.gp = S.gp;
    // tier 1, pass 1: every point, at most S.first_iters outer iterations;
    // on a mesh, in the order of the work of each point's previous update
    A.first = 1;
    A.in_list = nullptr;
    A.in_count = nullptr;
    if (A.mesh) {
        if ((e = sort_points(S, nullptr, A.n, A.flops, stream)) != hipSuccess) return e;
        A.in_list = S.sorted;
    }
    A.cont_list = S.list_a;
    A.cont_count = S.counts + 0;
    A.ovf_list = S.list_ovf;
    A.ovf_count = S.counts + 2;
    const int fi = S.first_iters > 0 ? S.first_iters : 8;
    A.iter_limit = fi < C.maxit ? fi : C.maxit;
    const int wide_limit = S.wide_iters > 0 ? S.wide_iters : fi;
    for (int q = 0; q < 4; ++q) S.pass_ms[q] = 0.0;
    const bool timed = S.ev[0] != nullptr;
    auto mark = [&](int k) {
        if (timed) (void)hipEventRecord(S.ev[k], stream);
    };
    auto span = [&](int q, int a, int b) {
        float ms = 0.f;
        if (timed && hipEventElapsedTime(&ms, S.ev[a], S.ev[b]) == hipSuccess) S.pass_ms[q] += ms;
    };
    mark(0);
    launch_one<N, kAmax>(dmat, C, dt, A, A.n, S, stream);
    mark(1);
    if ((e = read_counts<N>(S, stream)) != hipSuccess) return e;
    span(0, 0, 1);
    counts[0] = S.host_counts[0];
    A.first = 0;
    // tier 1, pass 2: the unfinished points, compacted, up to maxit
    if (S.host_counts[0] > 0) {
        if ((e = sort_points(S, S.list_a, S.host_counts[0], nullptr, stream)) != hipSuccess) return e;
        A.in_list = S.sorted;
        A.in_count = S.counts + 0;
        A.cont_list = S.list_b;
        A.cont_count = S.counts + 1;
        A.iter_limit = C.maxit;
        mark(1);
        launch_one<N, kAmax>(dmat, C, dt, A, S.host_counts[0], S, stream);
        mark(2);
        if ((e = read_counts<N>(S, stream)) != hipSuccess) return e;
        span(1, 1, 2);
    }
    counts[1] = S.host_counts[2];
    // tier 2: overflow of tier 1 (list_ovf) -> overflow into list_a
    if (S.host_counts[2] > 0) {
        hipLaunchKernelGGL(k_zero4, dim3(1), dim3(64), 0, stream, S.counts);  // list_a reused below
        // the overflow count was read on the host; keep it on the device in slot 3
        hipLaunchKernelGGL(k_set1, dim3(1), dim3(64), 0, stream, S.counts + 3, S.host_counts[2]);
        // the points of the wider tier mostly outgrow the first kernel's width at
        // their first outer iteration and keep the wider set; what they will cost
        // is best told by their previous update's work, so on a mesh they run in
        // the order of the bins the first pass sorted by
        if ((e = sort_points(S, S.list_ovf, S.host_counts[2], nullptr, stream, A.mesh != 0)) != hipSuccess) return e;
        A.in_list = S.sorted;
        A.in_count = S.counts + 3;
        A.cont_list = S.list_b;
        A.cont_count = S.counts + 1;
        A.ovf_list = S.list_a;
        A.ovf_count = S.counts + 0;
        // as in tier 1: at most first_iters outer iterations, then the
        // unfinished points resume sorted by their inner iterations per outer
        // one (the split is exact: a point resumes from its checkpoint)
        A.iter_limit = wide_limit < C.maxit ? wide_limit : C.maxit;
        mark(2);
        launch_one<N, kAmaxMid>(dmat, C, dt, A, S.host_counts[2], S, stream);
        if ((e = read_counts<N>(S, stream)) != hipSuccess) return e;
        if (S.host_counts[1] > 0) {
            if ((e = sort_points(S, S.list_b, S.host_counts[1], nullptr, stream)) != hipSuccess) return e;
            A.in_list = S.sorted;
            A.in_count = S.counts + 1;
            A.cont_count = S.counts + 3;   // no point continues at iter_limit = maxit
            A.iter_limit = C.maxit;
            launch_one<N, kAmaxMid>(dmat, C, dt, A, S.host_counts[1], S, stream);
        }
        mark(3);
        if ((e = read_counts<N>(S, stream)) != hipSuccess) return e;
        span(2, 2, 3);
        counts[2] = S.host_counts[0];
        // tier 3: overflow of tier 2 (list_a)
        if (S.host_counts[0] > 0) {
            A.in_list = S.list_a;
            A.in_count = S.counts + 0;
            A.cont_list = S.list_b;
            A.cont_count = S.counts + 1;
            A.ovf_list = S.list_ovf;
            A.ovf_count = S.counts + 2;
            hipLaunchKernelGGL(k_set1, dim3(1), dim3(64), 0, stream, S.counts + 2, 0);
            mark(3);
            launch_one<N, kAmaxWide>(dmat, C, dt, A, S.host_counts[0], S, stream);
            mark(4);
            if ((e = read_counts<N>(S, stream)) != hipSuccess) return e;
            span(3, 3, 4);
            counts[3] = S.host_counts[2];
            if constexpr (N == kNGeneric) {
                // tier 4: overflow of tier 3 (list_ovf); materials with more than
                // 18 systems can activate up to kAmaxGeneric
                if (S.host_counts[2] > 0) {
                    hipLaunchKernelGGL(k_set1, dim3(1), dim3(64), 0, stream, S.counts + 3, S.host_counts[2]);
                    hipLaunchKernelGGL(k_set1, dim3(1), dim3(64), 0, stream, S.counts + 0, 0);
                    A.in_list = S.list_ovf;
                    A.in_count = S.counts + 3;
                    A.cont_list = S.list_b;
                    A.cont_count = S.counts + 1;
                    A.ovf_list = S.list_a;
                    A.ovf_count = S.counts + 0;
                    mark(3);
                    launch_one<N, kAmaxGeneric>(dmat, C, dt, A, S.host_counts[2], S, stream);
                    mark(4);
                    if ((e = read_counts<N>(S, stream)) != hipSuccess) return e;
                    span(3, 3, 4);
                    if (S.host_counts[0] > 0) return hipErrorNotSupported;
                }
            } else {
                // more than kAmaxWide active systems at one point: not compiled in
                if (S.host_counts[2] > 0) return hipErrorNotSupported;
            }
        }
    }
    return hipGetLastError();
}

// The schedule above with every decision on the device, for capture in a
// graph: no count is read back; each pass and tier is launched over the
// largest list it can get (its threads past the device count return at
// once) and the sorts read their list lengths on the device.  report[0] and
// report[1] receive the points continued after the first pass and those
// needing a wider tier, report[2] the points wider than the widest tier.
__global__ void k_tier_prep(int* counts, int* report) {
    if (threadIdx.x == 0) {
        report[0] = counts[0];
        report[1] = counts[2];
        counts[3] = counts[2];
        counts[0] = 0;
        counts[1] = 0;
        counts[2] = 0;
    }
}
__global__ void k_tier4_prep(int* counts) {
    if (threadIdx.x == 0) {
        counts[3] = counts[2];
        counts[0] = 0;
    }
}
__global__ void k_width_report(const int* counts, int slot, int* report) {
    if (threadIdx.x == 0) report[2] = counts[slot];
}

template <int N>
static hipError_t schedule_dev(const void* dmat, const DevControl& C, double dt, ConstArgs A, ConstSchedule& S,
                               hipStream_t stream, int* report) {
    const int n = A.n;
    hipLaunchKernelGGL(k_zero4, dim3(1), dim3(64), 0, stream, S.counts);
    A.ck_s = S.ck_s;
    A.ck_iter = S.ck_iter;
    A.ck_inner = S.ck_inner;
    A.ck_kerr = S.ck_kerr;
    A.el = S.el;
    A.gp = S.gp;
    A.first = 1;
    A.in_list = nullptr;
    A.in_count = nullptr;
    hipError_t e;
    if (A.mesh) {
        if ((e = sort_points(S, nullptr, n, A.flops, stream)) != hipSuccess) return e;
        A.in_list = S.sorted;
    }
    A.cont_list = S.list_a;
    A.cont_count = S.counts + 0;
    A.ovf_list = S.list_ovf;
    A.ovf_count = S.counts + 2;
    const int fi = S.first_iters > 0 ? S.first_iters : 8;
    A.iter_limit = fi < C.maxit ? fi : C.maxit;
    launch_one<N, kAmax>(dmat, C, dt, A, n, S, stream);
    // pass 2: the unfinished points
    A.first = 0;
    if ((e = sort_points(S, S.list_a, n, nullptr, stream, false, S.counts + 0)) != hipSuccess) return e;
    A.in_list = S.sorted;
    A.in_count = S.counts + 0;
    A.cont_list = S.list_b;
    A.cont_count = S.counts + 1;
    A.iter_limit = C.maxit;
    launch_one<N, kAmax>(dmat, C, dt, A, n, S, stream);
    // tier 2: the overflow of tier 1, by the previous update's work on a mesh
    hipLaunchKernelGGL(k_tier_prep, dim3(1), dim3(64), 0, stream, S.counts, report);
    if ((e = sort_points(S, S.list_ovf, n, nullptr, stream, A.mesh != 0, S.counts + 3)) != hipSuccess) return e;
    A.in_list = S.sorted;
    A.in_count = S.counts + 3;
    A.cont_list = S.list_b;
    A.cont_count = S.counts + 1;
    A.ovf_list = S.list_a;
    A.ovf_count = S.counts + 0;
    const int wide_limit = S.wide_iters > 0 ? S.wide_iters : fi;
    A.iter_limit = wide_limit < C.maxit ? wide_limit : C.maxit;
    launch_one<N, kAmaxMid>(dmat, C, dt, A, n, S, stream);
    if ((e = sort_points(S, S.list_b, n, nullptr, stream, false, S.counts + 1)) != hipSuccess) return e;
    A.in_list = S.sorted;
    A.in_count = S.counts + 1;
    A.cont_count = S.counts + 3;
    A.iter_limit = C.maxit;
    launch_one<N, kAmaxMid>(dmat, C, dt, A, n, S, stream);
    // tier 3: the overflow of tier 2
    A.in_list = S.list_a;
    A.in_count = S.counts + 0;
    A.ovf_list = S.list_ovf;
    A.ovf_count = S.counts + 2;
    launch_one<N, kAmaxWide>(dmat, C, dt, A, n, S, stream);
    if constexpr (N == kNGeneric) {
        hipLaunchKernelGGL(k_tier4_prep, dim3(1), dim3(64), 0, stream, S.counts);
        A.in_list = S.list_ovf;
        A.in_count = S.counts + 3;
        A.ovf_list = S.list_a;
        A.ovf_count = S.counts + 0;
        launch_one<N, kAmaxGeneric>(dmat, C, dt, A, n, S, stream);
        hipLaunchKernelGGL(k_width_report, dim3(1), dim3(64), 0, stream, (const int*)S.counts, 0, report);
    } else {
        hipLaunchKernelGGL(k_width_report, dim3(1), dim3(64), 0, stream, (const int*)S.counts, 2, report);
    }
    return hipGetLastError();
}

hipError_t run_constitutive_dev(int nss, const void* dmat, const DevControl& C, double dt, const ConstArgs& A,
                                ConstSchedule& S, hipStream_t stream, int* report) {
    switch (nss) {
        case 12: return schedule_dev<12>(dmat, C, dt, A, S, stream, report);
        case 18: return schedule_dev<18>(dmat, C, dt, A, S, stream, report);
        default:
            if (nss >= 1 && nss <= kNGeneric) return schedule_dev<kNGeneric>(dmat, C, dt, A, S, stream, report);
            return hipErrorInvalidValue;
    }
}

hipError_t run_constitutive(int nss, const void* dmat, const DevControl& C, double dt, const ConstArgs& A,
                            ConstSchedule& S, hipStream_t stream, int counts[4]) {
    switch (nss) {
        case 12: return schedule<12>(dmat, C, dt, A, S, stream, counts);
        case 18: return schedule<18>(dmat, C, dt, A, S, stream, counts);
        default:
            if (nss >= 1 && nss <= kNGeneric) return schedule<kNGeneric>(dmat, C, dt, A, S, stream, counts);
            return hipErrorInvalidValue;
    }
}

// Resident workgroups of every compiled kernel, looked up once outside any
// stream capture.
static void prime_resident() {
    (void)resident_blocks<12, kAmax>();
    (void)resident_blocks<12, kAmaxMid>();
    (void)resident_blocks<12, kAmaxWide>();
    (void)resident_blocks<18, kAmax>();
    (void)resident_blocks<18, kAmaxMid>();
    (void)resident_blocks<18, kAmaxWide>();
    (void)resident_blocks<kNGeneric, kAmax>();
    (void)resident_blocks<kNGeneric, kAmaxMid>();
    (void)resident_blocks<kNGeneric, kAmaxWide>();
    (void)resident_blocks<kNGeneric, kAmaxGeneric>();
}

static bool queue_enabled() {
    const char* v = getenv("CPFEM_CONST_QUEUE");
    return !(v && atoi(v) == 0);
}

hipError_t alloc_schedule(ConstSchedule& S, int n) {
    hipError_t e;
    S.queue = queue_enabled();
    prime_resident();
    if ((e = hipMalloc((void**)&S.ck_s, sizeof(double) * 6 * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.ck_iter, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.ck_kerr, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.ck_inner, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.sorted, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.prev_bin, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.bins, sizeof(int) * kSortBins)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.list_a, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.list_b, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.list_ovf, sizeof(int) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.el, sizeof(double) * 21 * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.gp, sizeof(double) * kConstMaxWidth * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&S.counts, sizeof(int) * 8)) != hipSuccess) return e;
    if ((e = hipHostMalloc((void**)&S.host_counts, sizeof(int) * 4, 0)) != hipSuccess) return e;
    for (int k = 0; k < 5; ++k)
        if ((e = hipEventCreate(&S.ev[k])) != hipSuccess) return e;
    return hipSuccess;
}

void free_schedule(ConstSchedule& S) {
    void* p[] = {S.ck_s, S.ck_iter, S.ck_kerr, S.ck_inner, S.sorted, S.prev_bin, S.bins, S.list_a,
                 S.list_b, S.list_ovf, S.el, S.gp, S.counts};
    for (void* q : p)
        if (q) (void)hipFree(q);
    if (S.host_counts) (void)hipHostFree(S.host_counts);
    for (int k = 0; k < 5; ++k)
        if (S.ev[k]) (void)hipEventDestroy(S.ev[k]);
    S = ConstSchedule();
}

}  // namespace cpfem

using namespace cpfem;

// Stream epoch of the twin draws of the next cpfem_dstrss_batch call: every
// call takes its own, so repeated calls draw afresh as the reference's
// RANDOM_NUMBER stream does (crys-cons.f90:169).
static std::atomic<uint64_t> g_batch_epoch{1};

void cpfem::set_batch_epoch(uint64_t next) { g_batch_epoch.store(next); }

extern "C" int cpfem_dstrss_batch(const cpfem_material* mat, const cpfem_control* ctl, double dt, int32_t n,
                                  const int32_t* iel, const double* de, const double* r, const double* st,
                                  double* ds, const double* sv, double* dsv, int32_t ldsv, double* t, double* chi,
                                  double* deff, double* srate, int32_t* twinsys, const int32_t* timestwinned,
                                  int32_t* tflag, double* sumcpiter, int64_t* flops, int32_t* err) {
    if (!mat || !ctl || n < 0 || ldsv < 3 + mat->nss) return CPFEM_E_ARG;
    const int nss = mat->nss;
    const size_t mbytes = material_bytes(nss);
    if (mbytes == 0) return CPFEM_E_NSS;
    if (n == 0) return 0;
    for (int g = 0; g < n; ++g)
        if (iel[g] != 1 && iel[g] != -1) return CPFEM_E_ARG;  // one material set per batch
    const size_t N = (size_t)n, nsv = 3 + (size_t)nss;
    // field-major staging of the point arrays
    std::vector<double> h((6 + 3 + 6 + 6 + 6 + 36 + 2 + 2 * nsv + nss) * N);
    double* hde = h.data();
    double* hr = hde + 6 * N;
    double* hst = hr + 3 * N;
    double* hds = hst + 6 * N;
    double* ht = hds + 6 * N;
    double* hchi = ht + 6 * N;
    double* hdeff = hchi + 36 * N;
    double* hrep = hdeff + N;
    double* hsv = hrep + N;
    double* hdsv = hsv + nsv * N;
    double* hsr = hdsv + nsv * N;
    for (size_t g = 0; g < N; ++g) {
        for (int i = 0; i < 6; ++i) {
            hde[i * N + g] = de[6 * g + i];
            hst[i * N + g] = st[6 * g + i];
            hds[i * N + g] = ds[6 * g + i];
        }
        for (int i = 0; i < 3; ++i) hr[i * N + g] = r[3 * g + i];
        for (size_t i = 0; i < nsv; ++i) {
            hsv[i * N + g] = sv[(size_t)ldsv * g + i];
            hdsv[i * N + g] = dsv[(size_t)ldsv * g + i];
        }
        for (int i = 0; i < nss; ++i) hsr[i * N + g] = srate[CPFEM_MAXSS * g + i];
    }
    std::vector<char> hm(mbytes);
    build_material(nss, hm.data(), *mat, ctl->toler);
    DevBuf<double> dd;
    DevBuf<int32_t> di;
    DevBuf<int64_t> dl;
    DevBuf<char> dm;
    CPFEM_HIP(dd.alloc(h.size()));
    CPFEM_HIP(di.alloc(4 * N + 2));
    CPFEM_HIP(dl.alloc(2 * N));
    CPFEM_HIP(dm.alloc(mbytes));
    CPFEM_HIP(hipMemcpy(dd.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(dm.p, hm.data(), mbytes, hipMemcpyHostToDevice));
    std::vector<int32_t> hi(4 * N + 2, 0);
    for (size_t g = 0; g < N; ++g) {
        hi[g] = iel[g];
        hi[N + g] = twinsys[g];
        hi[2 * N + g] = timestwinned[g];
    }
    hi[4 * N] = *tflag;
    CPFEM_HIP(hipMemcpy(di.p, hi.data(), hi.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    ConstArgs A;
    memset(&A, 0, sizeof(A));
    A.n = n;
    A.mesh = 0;
    A.iel = di.p;
    A.twinsys = di.p + N;
    A.timestwinned = di.p + 2 * N;
    A.err = di.p + 3 * N;
    A.tflag = di.p + 4 * N;
    A.errmax = di.p + 4 * N + 1;
    A.de = dd.p + (hde - h.data());
    A.r = dd.p + (hr - h.data());
    A.st = dd.p + (hst - h.data());
    A.ds = dd.p + (hds - h.data());
    A.t = dd.p + (ht - h.data());
    A.chi = dd.p + (hchi - h.data());
    A.dep = dd.p + (hdeff - h.data());
    A.erepmd = dd.p + (hrep - h.data());
    A.sv = dd.p + (hsv - h.data());
    A.dsv = dd.p + (hdsv - h.data());
    A.srate = dd.p + (hsr - h.data());
    A.flops = reinterpret_cast<long long*>(dl.p);
    A.transc = reinterpret_cast<long long*>(dl.p + N);
    A.rng_epoch = g_batch_epoch.fetch_add(1);
    ConstSchedule S;
    memset(&S, 0, sizeof(S));
    hipError_t he = alloc_schedule(S, n);
    if (he != hipSuccess) {
        free_schedule(S);
        return CPFEM_E_HIP;
    }
    DevControl C = make_control(*ctl);
    int counts[4];
    he = run_constitutive(nss, dm.p, C, dt, A, S, 0, counts);
    if (he == hipSuccess) he = hipDeviceSynchronize();
    free_schedule(S);
    if (he == hipErrorNotSupported) return CPFEM_E_WIDTH;
    if (he != hipSuccess) return CPFEM_E_HIP;
    CPFEM_HIP(hipMemcpy(h.data(), dd.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    CPFEM_HIP(hipMemcpy(hi.data(), di.p, hi.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> hl(2 * N);
    CPFEM_HIP(hipMemcpy(hl.data(), dl.p, hl.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    int worst = 0;
    double sum = sumcpiter ? *sumcpiter : 0.0;
    for (size_t g = 0; g < N; ++g) {
        for (int i = 0; i < 6; ++i) {
            ds[6 * g + i] = hds[i * N + g];
            t[6 * g + i] = ht[i * N + g];
        }
        for (int k = 0; k < 36; ++k) chi[36 * g + k] = hchi[k * N + g];
        deff[g] = hdeff[g];
        for (size_t i = 0; i < nsv; ++i) dsv[(size_t)ldsv * g + i] = hdsv[i * N + g];
        for (int i = 0; i < nss; ++i) srate[CPFEM_MAXSS * g + i] = hsr[i * N + g];
        twinsys[g] = hi[N + g];
        if (err) err[g] = hi[3 * N + g];
        if (flops) flops[g] = hl[g];
        if (hi[3 * N + g] > worst) worst = hi[3 * N + g];
        sum += hrep[g];
    }
    if (sumcpiter) *sumcpiter = sum;
    if (hi[4 * N]) *tflag = 1;
    return worst;
}

extern "C" int cpfem_device_info(char* name, int32_t len, int32_t* cus) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess) return CPFEM_E_NOGPU;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return CPFEM_E_NOGPU;
    if (name && len > 0) {
        strncpy(name, p.gcnArchName, (size_t)len - 1);
        name[len - 1] = 0;
    }
    if (cus) *cus = p.multiProcessorCount;
    return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free variant of cpfem_dstrss_batch.  Every device buffer it
// needs (the staging of the caller's host arrays, the field-major point
// arrays, the material record and the pass schedule) is carved from one
// file-scope device array, `scratch`; each GPU holds its own copy, whose
// address is looked up once per device.  Calls on one GPU share it, so calls
// that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {
constexpr int32_t kNoAllocMaxPoints = 8192;
constexpr size_t kScratchBytes = size_t(40) << 20;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static unsigned char* g_scratch_ptr[kMaxDevices];      // device address of `scratch` on each GPU
static int g_noalloc_counts[kMaxDevices][4];            // host side of the pass counters
static std::mutex g_scratch_mu;

// soa[i*n + g] = aos[g*ld + i] for i < k (field-major copy of a per-point array)
__global__ void k_gather_fields(const double* __restrict__ aos, int ld, int k, int n, double* __restrict__ soa) {
    const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (idx < (size_t)k * n) {
        const size_t g = idx % (size_t)n, i = idx / (size_t)n;
        soa[idx] = aos[g * (size_t)ld + i];
    }
}

// aos[g*ld + i] = soa[i*n + g] for i < k (entries i >= k of aos are left as they are)
__global__ void k_scatter_fields(const double* __restrict__ soa, int k, int n, double* __restrict__ aos, int ld) {
    const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (idx < (size_t)k * n) {
        const size_t g = idx % (size_t)n, i = idx / (size_t)n;
        aos[g * (size_t)ld + i] = soa[idx];
    }
}

// sumcpiter + sum of erepmd in point order (the order of the host sum of
// cpfem_dstrss_batch, so the two entry points give the same bits)
__global__ void k_ordered_sum(const double* __restrict__ v, int n, double start, double* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = start;
        for (int g = 0; g < n; ++g) s += v[g];
        *out = s;
    }
}

// Carves 256-byte-aligned pieces out of the scratch array.
struct Carver {
    unsigned char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += count * sizeof(T);
        return p;
    }
};

static int scratch_for_device(unsigned char** out, int* dev_out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return CPFEM_E_NOGPU;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "cpfem: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return CPFEM_E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    *dev_out = dev;
    return 0;
}

extern "C" int cpfem_dstrss_batch_noalloc(const cpfem_material* mat, const cpfem_control* ctl, double dt, int32_t n,
                                          const int32_t* iel, const double* de, const double* r, const double* st,
                                          double* ds, const double* sv, double* dsv, int32_t ldsv, double* t,
                                          double* chi, double* deff, double* srate, int32_t* twinsys,
                                          const int32_t* timestwinned, int32_t* tflag, double* sumcpiter,
                                          int64_t* flops, int32_t* err) {
    if (!mat || !ctl || n < 0 || ldsv < 3 + mat->nss || ldsv > CPFEM_MAXSV) return CPFEM_E_ARG;
    if (n > kNoAllocMaxPoints) {
        fprintf(stderr, "cpfem_dstrss_batch_noalloc: %d points exceed the scratch limit of %d\n", n,
                kNoAllocMaxPoints);
        return CPFEM_E_ARG;
    }
    const int nss = mat->nss;
    const size_t mbytes = material_bytes(nss);
    if (mbytes == 0) return CPFEM_E_NSS;
    if (n == 0) return 0;
    for (int g = 0; g < n; ++g)
        if (iel[g] != 1 && iel[g] != -1) return CPFEM_E_ARG;  // one material set per batch
    unsigned char* base = nullptr;
    int dev = 0;
    if (const int rc = scratch_for_device(&base, &dev)) return rc;
    const size_t N = (size_t)n;
    const int nsv = 3 + nss;
    Carver cv{base};
    // staging of the caller's per-point arrays in their own layout
    double* s_de = cv.take<double>(6 * N);
    double* s_r = cv.take<double>(3 * N);
    double* s_st = cv.take<double>(6 * N);
    double* s_ds = cv.take<double>(6 * N);
    double* s_sv = cv.take<double>((size_t)ldsv * N);
    double* s_dsv = cv.take<double>((size_t)ldsv * N);
    double* s_srate = cv.take<double>((size_t)CPFEM_MAXSS * N);
    double* s_t = cv.take<double>(6 * N);
    double* s_chi = cv.take<double>(36 * N);
    // field-major point arrays the kernels read and write
    double* f_de = cv.take<double>(6 * N);
    double* f_r = cv.take<double>(3 * N);
    double* f_st = cv.take<double>(6 * N);
    double* f_ds = cv.take<double>(6 * N);
    double* f_t = cv.take<double>(6 * N);
    double* f_chi = cv.take<double>(36 * N);
    double* f_deff = cv.take<double>(N);
    double* f_rep = cv.take<double>(N);
    double* f_sv = cv.take<double>((size_t)nsv * N);
    double* f_dsv = cv.take<double>((size_t)nsv * N);
    double* f_srate = cv.take<double>((size_t)nss * N);
    int32_t* f_int = cv.take<int32_t>(4 * N + 2);          // iel, twinsys, timestwinned, err, tflag, errmax
    long long* f_l = cv.take<long long>(2 * N);            // flops, transc
    double* f_sum = cv.take<double>(1);
    char* f_mat = cv.take<char>(mbytes);
    ConstSchedule S;
    S.ck_s = cv.take<double>(6 * N);
    S.ck_iter = cv.take<int>(N);
    S.ck_kerr = cv.take<int>(N);
    S.ck_inner = cv.take<int>(N);
    S.sorted = cv.take<int>(N);
    S.bins = cv.take<int>(1024);
    S.list_a = cv.take<int>(N);
    S.list_b = cv.take<int>(N);
    S.list_ovf = cv.take<int>(N);
    S.el = cv.take<double>(21 * N);
    S.gp = cv.take<double>((size_t)kConstMaxWidth * N);
    S.counts = cv.take<int>(8);
    S.host_counts = g_noalloc_counts[dev];
    S.queue = queue_enabled();
    prime_resident();
    if (cv.off > kScratchBytes) return CPFEM_E_ARG;   // cannot happen within kNoAllocMaxPoints
    hipStream_t s0 = 0;
    // material record, built in host stack memory
    alignas(16) unsigned char hm[sizeof(DevMaterial<kNGeneric>)];
    build_material(nss, hm, *mat, ctl->toler);
    CPFEM_HIP(hipMemcpy(f_mat, hm, mbytes, hipMemcpyHostToDevice));
    // inputs
    CPFEM_HIP(hipMemcpy(s_de, de, 6 * N * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(s_r, r, 3 * N * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(s_st, st, 6 * N * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(s_ds, ds, 6 * N * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(s_sv, sv, (size_t)ldsv * N * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(s_dsv, dsv, (size_t)ldsv * N * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(s_srate, srate, (size_t)CPFEM_MAXSS * N * sizeof(double), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(f_int, iel, N * sizeof(int32_t), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(f_int + N, twinsys, N * sizeof(int32_t), hipMemcpyHostToDevice));
    CPFEM_HIP(hipMemcpy(f_int + 2 * N, timestwinned, N * sizeof(int32_t), hipMemcpyHostToDevice));
    const int32_t flags0[2] = {*tflag, 0};
    CPFEM_HIP(hipMemcpy(f_int + 4 * N, flags0, 2 * sizeof(int32_t), hipMemcpyHostToDevice));
    auto gather = [&](const double* aos, int ld, int k, double* soa) {
        const size_t tot = (size_t)k * N;
        hipLaunchKernelGGL(k_gather_fields, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s0, aos, ld, k, n, soa);
    };
    auto scatter = [&](const double* soa, int k, double* aos, int ld) {
        const size_t tot = (size_t)k * N;
        hipLaunchKernelGGL(k_scatter_fields, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s0, soa, k, n, aos,
                           ld);
    };
    gather(s_de, 6, 6, f_de);
    gather(s_r, 3, 3, f_r);
    gather(s_st, 6, 6, f_st);
    gather(s_ds, 6, 6, f_ds);
    gather(s_sv, ldsv, nsv, f_sv);
    gather(s_dsv, ldsv, nsv, f_dsv);
    gather(s_srate, CPFEM_MAXSS, nss, f_srate);
    // the err array and the counters start at zero (the scratch keeps the last call's contents)
    hipLaunchKernelGGL(k_zero_n, dim3((n + 255) / 256), dim3(256), 0, s0, f_int + 3 * N, n);
    CPFEM_HIP(hipGetLastError());
    ConstArgs A;
    memset(&A, 0, sizeof(A));
    A.n = n;
    A.mesh = 0;
    A.iel = f_int;
    A.twinsys = f_int + N;
    A.timestwinned = f_int + 2 * N;
    A.err = f_int + 3 * N;
    A.tflag = f_int + 4 * N;
    A.errmax = f_int + 4 * N + 1;
    A.de = f_de;
    A.r = f_r;
    A.st = f_st;
    A.ds = f_ds;
    A.t = f_t;
    A.chi = f_chi;
    A.dep = f_deff;
    A.erepmd = f_rep;
    A.sv = f_sv;
    A.dsv = f_dsv;
    A.srate = f_srate;
    A.flops = f_l;
    A.transc = f_l + N;
    A.rng_epoch = g_batch_epoch.fetch_add(1);
    DevControl C = make_control(*ctl);
    int counts[4];
    hipError_t he = run_constitutive(nss, f_mat, C, dt, A, S, s0, counts);
    if (he == hipErrorNotSupported) return CPFEM_E_WIDTH;
    if (he != hipSuccess) return CPFEM_E_HIP;
    // outputs back into the staging (caller's layout), then to the host
    scatter(f_ds, 6, s_ds, 6);
    scatter(f_t, 6, s_t, 6);
    scatter(f_chi, 36, s_chi, 36);
    scatter(f_dsv, nsv, s_dsv, ldsv);
    scatter(f_srate, nss, s_srate, CPFEM_MAXSS);
    hipLaunchKernelGGL(k_ordered_sum, dim3(1), dim3(64), 0, s0, (const double*)f_rep, n,
                       sumcpiter ? *sumcpiter : 0.0, f_sum);
    CPFEM_HIP(hipGetLastError());
    CPFEM_HIP(hipDeviceSynchronize());
    CPFEM_HIP(hipMemcpy(ds, s_ds, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    CPFEM_HIP(hipMemcpy(t, s_t, 6 * N * sizeof(double), hipMemcpyDeviceToHost));
    CPFEM_HIP(hipMemcpy(chi, s_chi, 36 * N * sizeof(double), hipMemcpyDeviceToHost));
    CPFEM_HIP(hipMemcpy(deff, f_deff, N * sizeof(double), hipMemcpyDeviceToHost));
    CPFEM_HIP(hipMemcpy(dsv, s_dsv, (size_t)ldsv * N * sizeof(double), hipMemcpyDeviceToHost));
    CPFEM_HIP(hipMemcpy(srate, s_srate, (size_t)CPFEM_MAXSS * N * sizeof(double), hipMemcpyDeviceToHost));
    CPFEM_HIP(hipMemcpy(twinsys, f_int + N, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (err) CPFEM_HIP(hipMemcpy(err, f_int + 3 * N, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (flops) CPFEM_HIP(hipMemcpy(flops, f_l, N * sizeof(int64_t), hipMemcpyDeviceToHost));
    int32_t flags[2];
    CPFEM_HIP(hipMemcpy(flags, f_int + 4 * N, 2 * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (sumcpiter) CPFEM_HIP(hipMemcpy(sumcpiter, f_sum, sizeof(double), hipMemcpyDeviceToHost));
    if (flags[0]) *tflag = 1;
    return flags[1];
}
