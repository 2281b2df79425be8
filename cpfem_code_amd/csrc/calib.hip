// Calibration support: replay of a recorded load path through the
// constitutive update, and the adjoint of that replay.
//
// Recording.  After NRITER of each increment (Stif_sol.f90:10-88) the inputs
// of every point's last DSTRSS call (Sig_eps.f90:84) are kept: strain and
// rotation increments de, r, the angle increments dsv(1:3) and the stress
// increment ds the call started from, and the point's integration weight.
// Replay.  With these increments fixed, the load path reduces to independent
// points: for each increment, DSTRSS on the point's own state (the product's
// constitutive kernels, run with the replay's material and tolerance) and the
// commit of UPDATE_GPTV (Fasolt3.f90:254-284).  The macroscopic curve is the
// weight-averaged committed stress component per increment.
// Adjoint.  The loss sum_n (curve_n - target_n)^2 is differentiated with
// respect to tau0 (the initial slip resistance of every system), taus, h0,
// the latent ratio q and the rate sensitivity m.  The reverse sweep keeps the
// state at the start of every ckpt_every-th increment; each segment is run
// forward again from its checkpoint (states and outputs kept per increment)
// and then reversed one increment at a time by k_adjoint, one point per lane.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "adjoint.hpp"
#include "common.hpp"
#include "constitutive.hpp"
#include "mesh.hpp"

struct cpfem_ctx : cpfem::Ctx {};

namespace cpfem {

namespace {

constexpr int kCalBlock = 256;
// fields of one increment's record: de 6, r 3, dg 3, dsg 6, curve weight 1
// (zero for excluded points), integration weight 1
constexpr int kRec = 20;
constexpr int kRecDe = 0, kRecR = 6, kRecDg = 9, kRecDs = 12, kRecW = 18, kRecWi = 19;
constexpr int kAdjMaxSS = 24;   // slip systems the reverse kernel is compiled for

struct Calib {
    int ninc = 0, every = 0, nck = 0, n = 0, nss = 0, nfs = 0;   // nfs = 3 + nss state fields of sv
    double toler = 0.0;
    int maxit = 0;
    double* hist = nullptr;      // [ninc][kRec][n]
    std::vector<double> wsum;    // [ninc] total weight per increment
    std::vector<int> have;       // [ninc] 1 once recorded
    std::vector<int> have_eq;    // [ninc] 1 once recorded with coordinates and matrix
    double *st0 = nullptr, *ph0 = nullptr;   // initial stress [6][n], angles [3][n]
    // replay point state, in the layout of the constitutive kernels
    double *st = nullptr, *ds = nullptr, *sv = nullptr, *dsv = nullptr, *t = nullptr, *chi = nullptr;
    double *erepmd = nullptr, *srate = nullptr, *de = nullptr, *dr = nullptr, *dep = nullptr;
    int *twinsys = nullptr, *timestwinned = nullptr, *gperr = nullptr, *flags = nullptr;
    long long *flops = nullptr, *transc = nullptr;
    void* dmat = nullptr;        // replay material (parameters applied)
    size_t mat_bytes = 0;
    double* ck = nullptr;        // [nck][6 + nfs][n] states at segment starts
    double* seg_state = nullptr; // [every][6 + nfs][n]
    double* seg_out = nullptr;   // [every][6 + nss][n] new stress, slip rates
    double *stb = nullptr, *phb = nullptr, *s0b = nullptr, *pb = nullptr;  // cotangents
    int* counters = nullptr;     // [4] device: unconverged, singular, elastic
    int* bad = nullptr;          // [n] 1: the replay stopped at maxit in some increment
    std::vector<int> excluded;   // [n] host: points given zero weight so far
    // equilibrium correction: coordinates and Newton matrix of every increment
    int equil = 0;
    double* xhist = nullptr;     // [ninc][3*nnod]
    double* vhist = nullptr;     // [ninc][nvals]
    double* kcot = nullptr;      // [9][n] cotangents of de, r
    double* elam = nullptr;      // [6][n] strain increments of the multiplier field
    double* rlam = nullptr;      // [3][n]
    double* avec = nullptr;      // [3*nnod]
    long long solve_sweeps = 0;
    int solves = 0;
    double* partial = nullptr;   // [nblocks][4]
    std::vector<double> hpartial;
    int nblocks = 0;
    std::vector<void*> allocs;
};

template <typename T>
hipError_t calloc_dev(Calib* K, T** p, size_t n, hipStream_t s) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e == hipSuccess) {
        K->allocs.push_back((void*)*p);
        e = hipMemsetAsync(*p, 0, n * sizeof(T), s);
    }
    return e;
}

inline int nb_of(size_t n) { return (int)((n + kCalBlock - 1) / kCalBlock); }

// timing events of one call, released on every return path
struct Events {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    hipError_t create() {
        for (auto& x : e) {
            const hipError_t r = hipEventCreate(&x);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    ~Events() {
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// record slot: copies of the inputs of the last constitutive call
__global__ void k_record(int n, const double* __restrict__ de, const double* __restrict__ dr,
                         const double* __restrict__ dsv, const double* __restrict__ ds, double* __restrict__ rec) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = (size_t)n;
    for (int i = 0; i < 6; ++i) rec[(kRecDe + i) * N + g] = de[i * N + g];
    for (int i = 0; i < 3; ++i) rec[(kRecR + i) * N + g] = dr[i * N + g];
    for (int i = 0; i < 3; ++i) rec[(kRecDg + i) * N + g] = dsv[i * N + g];
    for (int i = 0; i < 6; ++i) rec[(kRecDs + i) * N + g] = ds[i * N + g];
}

// sum of a field over blocks (fixed order within a block)
__device__ void block_sum4(double v[4], double* partial) {
    __shared__ double sh[kCalBlock / 64][4];
    for (int q = 0; q < 4; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[q] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 4; ++q) sh[wave][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) {
            double s = 0.0;
            for (int w = 0; w < kCalBlock / 64; ++w) s += sh[w][q];
            partial[(size_t)blockIdx.x * 4 + q] = s;
        }
}

__global__ __launch_bounds__(kCalBlock) void k_wsum(int n, const double* __restrict__ w, double* partial) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {g < n ? w[g] : 0.0, 0.0, 0.0, 0.0};
    block_sum4(v, partial);
}

// initial replay state: stress and angles of the recorded start, every slip
// resistance tau0
__global__ void k_replay_init(int n, int nss, const double* __restrict__ st0, const double* __restrict__ ph0,
                              double tau0, double* __restrict__ st, double* __restrict__ sv) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = (size_t)n;
    for (int i = 0; i < 6; ++i) st[i * N + g] = st0[i * N + g];
    for (int i = 0; i < 3; ++i) sv[i * N + g] = ph0[i * N + g];
    for (int i = 0; i < nss; ++i) sv[(3 + i) * N + g] = tau0;
}

// state <-> buffer of 6 + nfs fields (stress, then sv)
__global__ void k_state_copy(int n, int nfs, const double* __restrict__ st, const double* __restrict__ sv,
                             double* __restrict__ out) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = (size_t)n;
    for (int i = 0; i < 6; ++i) out[i * N + g] = st[i * N + g];
    for (int i = 0; i < nfs; ++i) out[(6 + i) * N + g] = sv[i * N + g];
}
__global__ void k_state_restore(int n, int nfs, const double* __restrict__ in, double* __restrict__ st,
                                double* __restrict__ sv) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = (size_t)n;
    for (int i = 0; i < 6; ++i) st[i * N + g] = in[i * N + g];
    for (int i = 0; i < nfs; ++i) sv[i * N + g] = in[(6 + i) * N + g];
}

// inputs of one replayed DSTRSS call from the record
__global__ void k_replay_load(int n, int nfs, int nss, const double* __restrict__ rec, double* __restrict__ de,
                              double* __restrict__ dr, double* __restrict__ ds, double* __restrict__ dsv,
                              double* __restrict__ srate, int* __restrict__ twinsys, int* __restrict__ kerr) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = (size_t)n;
    for (int i = 0; i < 6; ++i) de[i * N + g] = rec[(kRecDe + i) * N + g];
    for (int i = 0; i < 3; ++i) dr[i * N + g] = rec[(kRecR + i) * N + g];
    for (int i = 0; i < 6; ++i) ds[i * N + g] = rec[(kRecDs + i) * N + g];
    for (int i = 0; i < 3; ++i) dsv[i * N + g] = rec[(kRecDg + i) * N + g];
    for (int i = 3; i < nfs; ++i) dsv[i * N + g] = 0.0;
    for (int i = 0; i < nss; ++i) srate[i * N + g] = 0.0;
    twinsys[g] = 0;
    kerr[g] = 0;
}

// UPDATE_GPTV for the replay (st += ds, sv += dsv), the weighted sum of the
// curve component and the count of points stopped at the iteration limit;
// with out != null the new stress and the slip rates are kept for the reverse
// sweep.
__global__ __launch_bounds__(kCalBlock) void k_replay_commit(int n, int nfs, int nss, int comp, double toler,
                                                             const double* __restrict__ rec, double* __restrict__ st,
                                                             const double* __restrict__ ds, double* __restrict__ sv,
                                                             const double* __restrict__ dsv,
                                                             const double* __restrict__ t,
                                                             const double* __restrict__ srate,
                                                             const double* __restrict__ erepmd, double* out,
                                                             int* bad, double* partial) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (g < n) {
        const size_t N = (size_t)n;
        for (int i = 0; i < 6; ++i) st[i * N + g] = st[i * N + g] + ds[i * N + g];
        for (int i = 0; i < nfs; ++i) sv[i * N + g] = sv[i * N + g] + dsv[i * N + g];
        if (out) {
            for (int i = 0; i < 6; ++i) out[i * N + g] = t[i * N + g];
            for (int i = 0; i < nss; ++i) out[(6 + i) * N + g] = srate[i * N + g];
        }
        const double w = rec[kRecW * N + g];
        v[0] = w * st[comp * N + g];
        v[1] = erepmd[g] > toler ? 1.0 : 0.0;
        if (bad && erepmd[g] > toler) bad[g] = 1;
    }
    block_sum4(v, partial);
}

template <int N, int NM>
__global__ __launch_bounds__(64) void k_adjoint(int n, int nss, double dt, int comp, double coef,
                                                const DevMaterial<N>* __restrict__ Mp,
                                                const int* __restrict__ gp_elem, const int* __restrict__ elem_le,
                                                const double* __restrict__ state, const double* __restrict__ rec,
                                                const double* __restrict__ out, double* __restrict__ stb,
                                                double* __restrict__ phb, double* __restrict__ s0b,
                                                double* __restrict__ pb, int* __restrict__ counters, int dry,
                                                const double* __restrict__ corr, double* __restrict__ kcot) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const DevMaterial<N>& M = Mp[0];
    const size_t Nn = (size_t)n;
    adj::PointIn<NM> P;
    for (int i = 0; i < 6; ++i) P.st[i] = state[i * Nn + g];
    for (int i = 0; i < 3; ++i) P.php[i] = state[(6 + i) * Nn + g];
    for (int i = 0; i < nss; ++i) P.s0[i] = state[(9 + i) * Nn + g];
    for (int i = 0; i < 6; ++i) P.de[i] = rec[(kRecDe + i) * Nn + g];
    for (int i = 0; i < 3; ++i) P.r[i] = rec[(kRecR + i) * Nn + g];
    for (int i = 0; i < 3; ++i) P.dg[i] = rec[(kRecDg + i) * Nn + g];
    for (int i = 0; i < 6; ++i) P.dsg[i] = rec[(kRecDs + i) * Nn + g];
    for (int i = 0; i < 6; ++i) P.so[i] = out[i * Nn + g];
    for (int i = 0; i < nss; ++i) P.srate[i] = out[(6 + i) * Nn + g];
    P.iel = elem_le[gp_elem[g]];
    double sb[6], ph[3], s0[NM], stn[6], p[adj::kParams], deb[6], rb[3];
    for (int i = 0; i < 6; ++i) sb[i] = stb[i * Nn + g];
    sb[comp] += coef * rec[kRecW * Nn + g];
    if (corr) {
        const double wi = rec[kRecWi * Nn + g];
        for (int i = 0; i < 6; ++i) sb[i] -= wi * corr[i * Nn + g];
    }
    for (int i = 0; i < 3; ++i) ph[i] = phb[i * Nn + g];
    for (int i = 0; i < nss; ++i) s0[i] = s0b[i * Nn + g];
    for (int i = 0; i < adj::kParams; ++i) p[i] = 0.0;
    const int rc = adj::point_reverse<N, NM>(M, nss, dt, P, sb, ph, s0, stn, p, deb, rb);
    if (kcot) {
        for (int i = 0; i < 6; ++i) kcot[i * Nn + g] = deb[i];
        for (int i = 0; i < 3; ++i) kcot[(6 + i) * Nn + g] = rb[i];
    }
    if (dry) return;
    if (rc < 0) atomicAdd(&counters[1], 1);
    if (rc == 0) atomicAdd(&counters[2], 1);
    for (int i = 0; i < 6; ++i) stb[i * Nn + g] = stn[i];
    for (int i = 0; i < 3; ++i) phb[i * Nn + g] = ph[i];
    for (int i = 0; i < nss; ++i) s0b[i * Nn + g] = s0[i];
    for (int i = 1; i < adj::kParams; ++i) pb[i * Nn + g] += p[i];
}

// zero weight for the points in bad, every increment
__global__ void k_exclude(int n, int ninc, const int* __restrict__ bad, double* __restrict__ hist) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n || !bad[g]) return;
    for (int k = 0; k < ninc; ++k) hist[((size_t)k * kRec + kRecW) * n + g] = 0.0;
}

// final sums: [0] tau0 (all slip-resistance cotangents at the start), [1..4]
__global__ __launch_bounds__(kCalBlock) void k_grad_sums(int n, int nss, const double* __restrict__ s0b,
                                                         const double* __restrict__ pb, int group, double* partial) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (g < n) {
        const size_t N = (size_t)n;
        if (group == 0) {
            double s = 0.0;
            for (int i = 0; i < nss; ++i) s += s0b[i * N + g];
            v[0] = s;
            v[1] = pb[1 * N + g];
            v[2] = pb[2 * N + g];
            v[3] = pb[3 * N + g];
        } else {
            v[0] = pb[4 * N + g];
        }
    }
    block_sum4(v, partial);
}

// Checkpoints and segment buffers for ckpt_every increments per segment
// (<= 0: one segment, the whole trajectory kept by the forward pass).
hipError_t alloc_segments(Calib* K, int ckpt_every, hipStream_t s) {
    for (double* p : {K->ck, K->seg_state, K->seg_out}) {
        if (!p) continue;
        auto it = std::find(K->allocs.begin(), K->allocs.end(), (void*)p);
        if (it != K->allocs.end()) K->allocs.erase(it);
        (void)hipFree(p);
    }
    K->ck = K->seg_state = K->seg_out = nullptr;
    K->every = ckpt_every <= 0 ? K->ninc : std::min(ckpt_every, K->ninc);
    K->nck = (K->ninc + K->every - 1) / K->every;
    const size_t n = (size_t)K->n;
    hipError_t e = calloc_dev(K, &K->ck, (size_t)K->nck * (6 + K->nfs) * n, s);
    if (e == hipSuccess) e = calloc_dev(K, &K->seg_state, (size_t)K->every * (6 + K->nfs) * n, s);
    if (e == hipSuccess) e = calloc_dev(K, &K->seg_out, (size_t)K->every * (6 + K->nss) * n, s);
    return e;
}

int sum_partials(Ctx* c, Calib* K, int nb, double out[4]) {
    CPFEM_HIP(hipMemcpyAsync(K->hpartial.data(), K->partial, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost,
                             c->stream));
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 4; ++q) out[q] = 0.0;
    for (int b = 0; b < nb; ++b)
        for (int q = 0; q < 4; ++q) out[q] += K->hpartial[(size_t)b * 4 + q];
    return 0;
}

// Material of the replay: material 1 with the five parameters applied.
int set_params(Ctx* c, Calib* K, const double p[5]) {
    cpfem_material m = c->mat0;
    for (int i = 0; i < m.nss; ++i) {
        m.ampar[i][0] = p[4];
        m.ampar[i][1] = p[2];
        m.ampar[i][4] = p[1];
        for (int k = 0; k < m.nss; ++k)
            if (k != i) m.smatrx[i][k] = p[3];
    }
    std::vector<unsigned char> host(K->mat_bytes);
    build_material(c->knss, host.data(), m, K->toler);
    CPFEM_HIP(hipMemcpyAsync(K->dmat, host.data(), K->mat_bytes, hipMemcpyHostToDevice, c->stream));
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// One replayed increment n (1-based): load, constitutive update, commit.
int replay_step(Ctx* c, Calib* K, int n, int comp, double* out, double sums[4], int* bad = nullptr) {
    hipStream_t s = c->stream;
    const int nb = nb_of(K->n);
    const double* rec = K->hist + (size_t)(n - 1) * kRec * K->n;
    hipLaunchKernelGGL(k_replay_load, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nfs, K->nss, rec, K->de, K->dr, K->ds,
                       K->dsv, K->srate, K->twinsys, c->sched.ck_kerr);
    CPFEM_HIP(hipGetLastError());
    ConstArgs A;
    memset(&A, 0, sizeof(A));
    A.n = K->n;
    A.mesh = 1;
    A.gp_elem = c->gp_elem;
    A.elem_le = c->elem_le;
    A.de = K->de;
    A.r = K->dr;
    A.dep = K->dep;
    A.st = K->st;
    A.ds = K->ds;
    A.sv = K->sv;
    A.dsv = K->dsv;
    A.t = K->t;
    A.chi = K->chi;
    A.erepmd = K->erepmd;
    A.srate = K->srate;
    A.twinsys = K->twinsys;
    A.timestwinned = K->timestwinned;
    A.flops = K->flops;
    A.transc = K->transc;
    A.err = K->gperr;
    A.tflag = K->flags;
    A.errmax = K->flags + 1;
    A.rng_epoch = 0;
    cpfem_control ctl = c->ctl;
    ctl.toler = K->toler;
    ctl.maxit = K->maxit;
    // points held by SCALE_IT's cap stop once their stress is stationary; they
    // still end above toler and are found by cpfem_calib_exclude
    ctl.stall_tol = 1e-14;
    const DevControl C = make_control(ctl);
    int counts[4] = {0, 0, 0, 0};
    const hipError_t he = run_constitutive(c->knss, K->dmat, C, c->tiem[n - 1], A, c->sched, s, counts);
    if (he == hipErrorNotSupported) return CPFEM_E_WIDTH;
    CPFEM_HIP(he);
    hipLaunchKernelGGL(k_replay_commit, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nfs, K->nss, comp, K->toler, rec,
                       K->st, (const double*)K->ds, K->sv, (const double*)K->dsv, (const double*)K->t,
                       (const double*)K->srate, (const double*)K->erepmd, out, bad, K->partial);
    CPFEM_HIP(hipGetLastError());
    return sum_partials(c, K, nb, sums);
}

template <int N, int NM>
void launch_adjoint(Ctx* c, Calib* K, int n, int comp, double coef, const double* state, const double* out, int dry,
                    const double* corr, double* kcot) {
    const double* rec = K->hist + (size_t)(n - 1) * kRec * K->n;
    hipLaunchKernelGGL((k_adjoint<N, NM>), dim3((K->n + 63) / 64), dim3(64), 0, c->stream, K->n, K->nss,
                       c->tiem[n - 1], comp, coef, (const DevMaterial<N>*)K->dmat, (const int*)c->gp_elem,
                       (const int*)c->elem_le, state, rec, out, K->stb, K->phb, K->s0b, K->pb, K->counters, dry,
                       corr, kcot);
}

void launch_adjoint_any(Ctx* c, Calib* K, int n, int comp, double coef, const double* state, const double* out,
                        int dry, const double* corr, double* kcot) {
    if (c->knss == 12)
        launch_adjoint<12, 12>(c, K, n, comp, coef, state, out, dry, corr, kcot);
    else if (c->knss == 18)
        launch_adjoint<18, 18>(c, K, n, comp, coef, state, out, dry, corr, kcot);
    else
        launch_adjoint<kNGeneric, kAdjMaxSS>(c, K, n, comp, coef, state, out, dry, corr, kcot);
}

// Reverse step of increment n.  With the equilibrium correction: a dry pass
// gives the cotangents of every point's de and r, their transpose through the
// kinematics is the right-hand side a of K lambda = a with the increment's
// recorded Newton matrix K (symmetric; the stiffness NRITER solves with,
// Stif_sol.f90:42-53), and the stress cotangents get -w B lambda: the
// response of the converged displacement increment to the parameters and to
// the state it started from.
int adjoint_step(Ctx* c, Calib* K, int n, int comp, double coef, const double* state, const double* out) {
    if (!K->equil) {
        launch_adjoint_any(c, K, n, comp, coef, state, out, 0, nullptr, nullptr);
        return 0;
    }
    const size_t D = 3 * (size_t)c->nnod;
    const double* xn = K->xhist + (size_t)(n - 1) * D;
    launch_adjoint_any(c, K, n, comp, coef, state, out, 1, nullptr, K->kcot);
    CPFEM_HIP(kin_transpose(c, xn, K->kcot, K->avec));
    int sweeps = 0;
    double err = 0.0;
    // consecutive increments have nearby multipliers: start from the last one
    const int rc = solve_with(c, K->vhist + (size_t)(n - 1) * c->A.nvals(), K->avec, &sweeps, &err, K->solves > 0);
    if (rc < 0 || (rc > 0 && rc != 904)) return rc;
    K->solve_sweeps += sweeps;
    K->solves += 1;
    CPFEM_HIP(kin_apply(c, xn, c->ddx, K->elam, K->rlam));
    launch_adjoint_any(c, K, n, comp, coef, state, out, 0, K->elam, nullptr);
    return 0;
}

}  // namespace

void calib_free(Ctx* c) {
    Calib* K = (Calib*)c->calib;
    if (!K) return;
    for (void* p : K->allocs) (void)hipFree(p);
    delete K;
    c->calib = nullptr;
}

}  // namespace cpfem

using namespace cpfem;

extern "C" int cpfem_calib_begin(cpfem_ctx* c, int32_t ninc, int32_t ckpt_every, double toler, int32_t maxit) {
    if (!c || ninc < 1 || ninc > c->ninc1 || !(toler > 0.0) || maxit < 1) return CPFEM_E_ARG;
    // one material, no twin systems, uniform parameters over the systems and
    // a two-valued latent matrix: the five parameters then have one meaning
    if (c->nmat != 1) return CPFEM_E_ARG;
    const cpfem_material& m = c->mat0;
    if (m.nss < 1 || m.nss > kAdjMaxSS) return CPFEM_E_NSS;
    for (int i = 0; i < m.nss; ++i) {
        if (fabs(m.ssang[i]) > 1.5) return CPFEM_E_ARG;
        for (int q = 0; q < 6; ++q)
            if (m.ampar[i][q] != m.ampar[0][q]) return CPFEM_E_ARG;
        for (int k = 0; k < m.nss; ++k)
            if (k != i && m.nss > 1 && m.smatrx[i][k] != m.smatrx[0][1]) return CPFEM_E_ARG;
    }
    calib_free(c);
    Calib* K = new Calib();
    c->calib = K;
    K->ninc = ninc;
    K->n = c->ngps;
    K->nss = m.nss;
    K->nfs = 3 + c->nss;
    K->toler = toler;
    K->maxit = maxit;
    K->wsum.assign(ninc, 0.0);
    K->have.assign(ninc, 0);
    K->have_eq.assign(ninc, 0);
    K->mat_bytes = material_bytes(c->knss);
    K->nblocks = nb_of(K->n);
    K->hpartial.assign((size_t)4 * K->nblocks, 0.0);
    hipStream_t s = c->stream;
    const size_t n = (size_t)K->n;
    hipError_t e = hipSuccess;
#define CAL_ALLOC(ptr, count) \
    if (e == hipSuccess) e = calloc_dev(K, &(ptr), (count), s)
    CAL_ALLOC(K->hist, (size_t)ninc * kRec * n);
    CAL_ALLOC(K->st0, 6 * n);
    CAL_ALLOC(K->ph0, 3 * n);
    CAL_ALLOC(K->st, 6 * n);
    CAL_ALLOC(K->ds, 6 * n);
    CAL_ALLOC(K->sv, (size_t)K->nfs * n);
    CAL_ALLOC(K->dsv, (size_t)K->nfs * n);
    CAL_ALLOC(K->t, 6 * n);
    CAL_ALLOC(K->chi, 36 * n);
    CAL_ALLOC(K->erepmd, n);
    CAL_ALLOC(K->srate, (size_t)c->nss * n);
    CAL_ALLOC(K->de, 6 * n);
    CAL_ALLOC(K->dr, 3 * n);
    CAL_ALLOC(K->dep, n);
    CAL_ALLOC(K->twinsys, n);
    CAL_ALLOC(K->timestwinned, n);
    CAL_ALLOC(K->gperr, n);
    CAL_ALLOC(K->flags, 4);
    CAL_ALLOC(K->flops, n);
    CAL_ALLOC(K->transc, n);
    CAL_ALLOC(K->stb, 6 * n);
    CAL_ALLOC(K->phb, 3 * n);
    CAL_ALLOC(K->s0b, (size_t)K->nss * n);
    CAL_ALLOC(K->pb, (size_t)adj::kParams * n);
    CAL_ALLOC(K->counters, 4);
    CAL_ALLOC(K->bad, n);
    CAL_ALLOC(K->partial, (size_t)4 * K->nblocks);
#undef CAL_ALLOC
    if (e == hipSuccess) {
        unsigned char* dm = nullptr;
        e = hipMalloc((void**)&dm, K->mat_bytes);
        if (e == hipSuccess) {
            K->allocs.push_back(dm);
            K->dmat = dm;
        }
    }
    if (e == hipSuccess) e = alloc_segments(K, ckpt_every, s);
    if (e != hipSuccess) {
        calib_free(c);
        fprintf(stderr, "cpfem: calibration buffers: %s\n", hipGetErrorString(e));
        return CPFEM_E_HIP;
    }
    // the start of the load path: current stress and angles
    CPFEM_HIP(hipMemcpyAsync(K->st0, c->st, sizeof(double) * 6 * n, hipMemcpyDeviceToDevice, s));
    CPFEM_HIP(hipMemcpyAsync(K->ph0, c->sv, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
    CPFEM_HIP(hipStreamSynchronize(s));
    return 0;
}

extern "C" int cpfem_calib_record(cpfem_ctx* c, int32_t inc) {
    if (!c || !c->calib) return CPFEM_E_ARG;
    Calib* K = (Calib*)c->calib;
    if (inc < 1 || inc > K->ninc) return CPFEM_E_ARG;
    const size_t n = (size_t)K->n;
    double* rec = K->hist + (size_t)(inc - 1) * kRec * n;
    const int nb = nb_of(n);
    hipLaunchKernelGGL(k_record, dim3(nb), dim3(kCalBlock), 0, c->stream, K->n, (const double*)c->de,
                       (const double*)c->dr, (const double*)c->dsv, (const double*)c->ds, rec);
    CPFEM_HIP(hipGetLastError());
    CPFEM_HIP(point_weights(c, rec + kRecW * n));
    CPFEM_HIP(hipMemcpyAsync(rec + kRecWi * n, rec + kRecW * n, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    K->have_eq[inc - 1] = K->equil;
    if (K->equil) {
        const size_t D = 3 * (size_t)c->nnod;
        CPFEM_HIP(hipMemcpyAsync(K->xhist + (size_t)(inc - 1) * D, c->x, sizeof(double) * D, hipMemcpyDeviceToDevice,
                                 c->stream));
        CPFEM_HIP(hipMemcpyAsync(K->vhist + (size_t)(inc - 1) * c->A.nvals(), c->A.vals,
                                 sizeof(double) * c->A.nvals(), hipMemcpyDeviceToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_wsum, dim3(nb), dim3(kCalBlock), 0, c->stream, K->n, (const double*)(rec + kRecW * n),
                       K->partial);
    CPFEM_HIP(hipGetLastError());
    double sums[4];
    const int rc = sum_partials(c, K, nb, sums);
    if (rc) return rc;
    K->wsum[inc - 1] = sums[0];
    K->have[inc - 1] = 1;
    return 0;
}

extern "C" int cpfem_calib_history(cpfem_ctx* c, int32_t inc, double* rec_out) {
    if (!c || !c->calib || !rec_out) return CPFEM_E_ARG;
    Calib* K = (Calib*)c->calib;
    if (inc < 1 || inc > K->ninc || !K->have[inc - 1]) return CPFEM_E_ARG;
    const size_t n = (size_t)K->n;
    std::vector<double> f((size_t)kRec * n);
    CPFEM_HIP(hipMemcpy(f.data(), K->hist + (size_t)(inc - 1) * kRec * n, sizeof(double) * f.size(),
                        hipMemcpyDeviceToHost));
    for (size_t g = 0; g < n; ++g)
        for (int q = 0; q < 19; ++q) rec_out[g * 19 + q] = f[q * n + g];
    return 0;
}

extern "C" int cpfem_calib_run(cpfem_ctx* c, const double* params, const double* target, int32_t comp,
                               double* curve, double* loss, double* grad, cpfem_calib_stats* stats) {
    if (!c || !c->calib || !params || !target || comp < 0 || comp > 5) return CPFEM_E_ARG;
    Calib* K = (Calib*)c->calib;
    for (int i = 0; i < K->ninc; ++i)
        if (!K->have[i] || (K->equil && grad && !K->have_eq[i])) return CPFEM_E_ARG;
    if (!(params[0] > 0.0) || !(params[4] > 0.0)) return CPFEM_E_ARG;
    int rc = set_params(c, K, params);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const int nb = nb_of(K->n);
    const size_t n = (size_t)K->n;
    const size_t fs = (size_t)(6 + K->nfs) * n, fo = (size_t)(6 + K->nss) * n;
    CPFEM_HIP(hipMemsetAsync(K->counters, 0, 4 * sizeof(int), s));
    K->solves = 0;
    K->solve_sweeps = 0;
    Events ev;
    CPFEM_HIP(ev.create());
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    CPFEM_HIP(hipEventRecord(e0, s));
    // forward replay with a checkpoint at the start of every segment
    hipLaunchKernelGGL(k_replay_init, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nss, (const double*)K->st0,
                       (const double*)K->ph0, params[0], K->st, K->sv);
    CPFEM_HIP(hipGetLastError());
    std::vector<double> cur(K->ninc), coef(K->ninc);
    double L = 0.0;
    long long unconv = 0;
    // the last segment's states and outputs are kept by this pass, so the
    // reverse sweep starts without recomputing it
    const int last0 = 1 + (K->nck - 1) * K->every;
    for (int inc = 1; inc <= K->ninc; ++inc) {
        if (grad && (inc - 1) % K->every == 0) {
            const int k = (inc - 1) / K->every;
            hipLaunchKernelGGL(k_state_copy, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nfs, (const double*)K->st,
                               (const double*)K->sv, K->ck + (size_t)k * fs);
        }
        double* keep = nullptr;
        if (grad && inc >= last0) {
            hipLaunchKernelGGL(k_state_copy, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nfs, (const double*)K->st,
                               (const double*)K->sv, K->seg_state + (size_t)(inc - last0) * fs);
            keep = K->seg_out + (size_t)(inc - last0) * fo;
        }
        double sums[4];
        if ((rc = replay_step(c, K, inc, comp, keep, sums))) return rc;
        const double W = K->wsum[inc - 1];
        cur[inc - 1] = W > 0.0 ? sums[0] / W : 0.0;
        unconv += (long long)llround(sums[1]);
        const double d = cur[inc - 1] - target[inc - 1];
        L += d * d;
        coef[inc - 1] = W > 0.0 ? 2.0 * d / W : 0.0;
    }
    CPFEM_HIP(hipEventRecord(e1, s));
    if (curve)
        for (int i = 0; i < K->ninc; ++i) curve[i] = cur[i];
    if (loss) *loss = L;
    if (grad) {
        CPFEM_HIP(hipMemsetAsync(K->stb, 0, sizeof(double) * 6 * n, s));
        CPFEM_HIP(hipMemsetAsync(K->phb, 0, sizeof(double) * 3 * n, s));
        CPFEM_HIP(hipMemsetAsync(K->s0b, 0, sizeof(double) * K->nss * n, s));
        CPFEM_HIP(hipMemsetAsync(K->pb, 0, sizeof(double) * adj::kParams * n, s));
        for (int k = K->nck - 1; k >= 0; --k) {
            const int n0 = 1 + k * K->every, n1 = std::min(K->ninc, n0 + K->every - 1);
            hipLaunchKernelGGL(k_state_restore, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nfs,
                               (const double*)(K->ck + (size_t)k * fs), K->st, K->sv);
            for (int inc = n0; inc <= n1 && k < K->nck - 1; ++inc) {
                double* state = K->seg_state + (size_t)(inc - n0) * fs;
                hipLaunchKernelGGL(k_state_copy, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nfs,
                                   (const double*)K->st, (const double*)K->sv, state);
                double sums[4];
                if ((rc = replay_step(c, K, inc, comp, K->seg_out + (size_t)(inc - n0) * fo, sums))) return rc;
            }
            for (int inc = n1; inc >= n0; --inc)
                if ((rc = adjoint_step(c, K, inc, comp, coef[inc - 1], K->seg_state + (size_t)(inc - n0) * fs,
                                       K->seg_out + (size_t)(inc - n0) * fo)))
                    return rc;
            CPFEM_HIP(hipGetLastError());
        }
        double g0[4], g1[4];
        hipLaunchKernelGGL(k_grad_sums, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nss, (const double*)K->s0b,
                           (const double*)K->pb, 0, K->partial);
        if ((rc = sum_partials(c, K, nb, g0))) return rc;
        hipLaunchKernelGGL(k_grad_sums, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nss, (const double*)K->s0b,
                           (const double*)K->pb, 1, K->partial);
        if ((rc = sum_partials(c, K, nb, g1))) return rc;
        grad[0] = g0[0];
        grad[1] = g0[1];
        grad[2] = g0[2];
        grad[3] = g0[3];
        grad[4] = g1[0];
    }
    CPFEM_HIP(hipEventRecord(e2, s));
    CPFEM_HIP(hipEventSynchronize(e2));
    int cnt[4];
    CPFEM_HIP(hipMemcpy(cnt, K->counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->points = K->n;
        stats->increments = K->ninc;
        stats->unconverged = (int32_t)unconv;
        stats->singular = cnt[1];
        stats->elastic = cnt[2];
        stats->equilibrium_solves = K->solves;
        stats->solve_sweeps = (double)K->solve_sweeps;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        stats->t_forward_ms = ms;
        ms = 0.f;
        if (grad) (void)hipEventElapsedTime(&ms, e1, e2);
        stats->t_reverse_ms = ms;
    }
    return cnt[1] > 0 ? CPFEM_E_SINGULAR : 0;
}

extern "C" int cpfem_calib_exclude(cpfem_ctx* c, const double* params, int32_t* excluded) {
    if (!c || !c->calib || !params) return CPFEM_E_ARG;
    Calib* K = (Calib*)c->calib;
    for (int i = 0; i < K->ninc; ++i)
        if (!K->have[i]) return CPFEM_E_ARG;
    if (!(params[0] > 0.0) || !(params[4] > 0.0)) return CPFEM_E_ARG;
    int rc = set_params(c, K, params);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const int nb = nb_of(K->n);
    CPFEM_HIP(hipMemsetAsync(K->bad, 0, sizeof(int) * (size_t)K->n, s));
    hipLaunchKernelGGL(k_replay_init, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->nss, (const double*)K->st0,
                       (const double*)K->ph0, params[0], K->st, K->sv);
    CPFEM_HIP(hipGetLastError());
    for (int inc = 1; inc <= K->ninc; ++inc) {
        double sums[4];
        if ((rc = replay_step(c, K, inc, 2, nullptr, sums, K->bad))) return rc;
    }
    hipLaunchKernelGGL(k_exclude, dim3(nb), dim3(kCalBlock), 0, s, K->n, K->ninc, (const int*)K->bad, K->hist);
    CPFEM_HIP(hipGetLastError());
    for (int inc = 1; inc <= K->ninc; ++inc) {
        const double* rec = K->hist + (size_t)(inc - 1) * kRec * K->n;
        hipLaunchKernelGGL(k_wsum, dim3(nb), dim3(kCalBlock), 0, s, K->n, rec + (size_t)kRecW * K->n, K->partial);
        CPFEM_HIP(hipGetLastError());
        double sums[4];
        if ((rc = sum_partials(c, K, nb, sums))) return rc;
        K->wsum[inc - 1] = sums[0];
    }
    std::vector<int> b(K->n);
    CPFEM_HIP(hipMemcpy(b.data(), K->bad, sizeof(int) * b.size(), hipMemcpyDeviceToHost));
    if (K->excluded.size() != b.size()) K->excluded.assign(b.size(), 0);
    int cnt = 0;
    for (size_t g = 0; g < b.size(); ++g) {
        K->excluded[g] |= b[g] != 0;
        cnt += K->excluded[g];
    }
    if (excluded) *excluded = cnt;
    return 0;
}

extern "C" int cpfem_calib_options(cpfem_ctx* c, int32_t flags) {
    if (!c || !c->calib) return CPFEM_E_ARG;
    Calib* K = (Calib*)c->calib;
    if (!(flags & CPFEM_CALIB_EQUILIBRIUM)) {
        K->equil = 0;
        return 0;
    }
    if (c->precision != 0 || c->npts8 != c->ngps) return CPFEM_E_ARG;
    // switching the correction on: before the first record, or on a record
    // whose increments all kept their coordinates and matrices
    for (int i = 0; i < K->ninc; ++i)
        if (K->have[i] && !K->have_eq[i]) return CPFEM_E_ARG;
    const size_t D = 3 * (size_t)c->nnod, n = (size_t)K->n;
    hipError_t e = hipSuccess;
    if (!K->xhist) {
        e = calloc_dev(K, &K->xhist, (size_t)K->ninc * D, c->stream);
        if (e == hipSuccess) e = calloc_dev(K, &K->vhist, (size_t)K->ninc * c->A.nvals(), c->stream);
        if (e == hipSuccess) e = calloc_dev(K, &K->kcot, 9 * n, c->stream);
        if (e == hipSuccess) e = calloc_dev(K, &K->elam, 6 * n, c->stream);
        if (e == hipSuccess) e = calloc_dev(K, &K->rlam, 3 * n, c->stream);
        if (e == hipSuccess) e = calloc_dev(K, &K->avec, D, c->stream);
    }
    CPFEM_HIP(e);
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    K->equil = 1;
    return 0;
}

extern "C" int cpfem_calib_checkpoints(cpfem_ctx* c, int32_t ckpt_every) {
    if (!c || !c->calib) return CPFEM_E_ARG;
    Calib* K = (Calib*)c->calib;
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    CPFEM_HIP(alloc_segments(K, ckpt_every, c->stream));
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int cpfem_calib_end(cpfem_ctx* c) {
    if (!c) return CPFEM_E_ARG;
    calib_free(c);
    return 0;
}
