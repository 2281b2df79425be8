// Constitutive kernels and the batch entry point of the C ABI.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "common.hpp"
#include "constitutive.hpp"
#include "constitutive_launch.hpp"

namespace cpfem {

// One lane per point; SoA arrays with stride n (field-major).
template <int N>
__global__ __launch_bounds__(64) void k_dstrss_points(const DevMaterial<N>* __restrict__ Mp, const DevControl C,
                                                      const double dt, const PointArrays A) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.n) return;
    const DevMaterial<N>& M = *Mp;
    const size_t n = (size_t)A.n;
    PointIO io;
    io.iel = A.iel[g];
    io.err = 0;
    io.tflag = 0;
    io.deff = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        io.de[i] = A.de[i * n + g];
        io.st[i] = A.st[i * n + g];
        io.ds[i] = A.ds[i * n + g];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        io.r[i] = A.r[i * n + g];
        io.sv0[i] = A.sv[i * n + g];
        io.dsv0[i] = A.dsv[i * n + g];
    }
    io.twinsys = A.twinsys[g];
    io.timestwinned = A.timestwinned[g];
    io.rng_key = (A.rng_epoch << 32) ^ (uint64_t)(g + A.gp_offset);
    double s0[N], dsv[3 + N], srate[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        s0[i] = A.sv[(3 + i) * n + g];
        srate[i] = A.srate[i * n + g];
    }
#pragma unroll
    for (int i = 0; i < 3 + N; ++i) dsv[i] = A.dsv[i * n + g];
    Count cnt{0, 0, 0, 0};
    dstrss<N>(M, C, dt, io, s0, dsv, srate, cnt);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        A.t[i * n + g] = io.t[i];
        A.ds[i * n + g] = io.ds[i];
    }
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int i = 0; i < 6; ++i) A.chi[(i + 6 * j) * n + g] = io.chi[i][j];
    A.deff[g] = io.deff;
    A.erepmd[g] = io.erepmd;
#pragma unroll
    for (int i = 0; i < 3 + N; ++i) A.dsv[i * n + g] = dsv[i];
#pragma unroll
    for (int i = 0; i < N; ++i) A.srate[i * n + g] = srate[i];
    A.twinsys[g] = io.twinsys;
    if (A.flops) A.flops[g] = cnt.flops;
    if (A.transc) A.transc[g] = cnt.transc;
    if (A.err) A.err[g] = io.err;
    if (io.err) atomicMax(A.errmax, io.err);
    if (io.tflag) atomicOr(A.tflag, 1);
}

template <int N>
void fill_material(DevMaterial<N>& d, const cpfem_material& m, double toler) {
    memset(&d, 0, sizeof(d));
    d.nss = m.nss;
    d.twinlim = m.twinlim;
    d.twinprob = m.twinprob;
    for (int i = 0; i < 9; ++i) d.e[i] = m.e[i];
    double amx = m.ampar[0][0];
    for (int i = 0; i < N; ++i) {
        d.ssang[i] = m.ssang[i];
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
        for (int k = 0; k < N; ++k) d.smatrx[i][k] = m.smatrx[k][i];  // Fortran smatrx(i,k)
        if (m.ampar[i][0] > amx) amx = m.ampar[i][0];
    }
    d.conv = toler * amx;
}

template void fill_material<12>(DevMaterial<12>&, const cpfem_material&, double);
template void fill_material<18>(DevMaterial<18>&, const cpfem_material&, double);

DevControl make_control(const cpfem_control& c) {
    DevControl d;
    d.toler = c.toler;
    d.alpha = c.alpha;
    d.afactor = c.afactor;
    d.maxit = c.maxit;
    d.pad_ = 0;
    return d;
}

template <int N>
hipError_t launch_dstrss_points(const void* dmat, const DevControl& C, double dt, const PointArrays& A,
                                hipStream_t stream) {
    if (A.n <= 0) return hipSuccess;
    const int block = 64;
    const int grid = (A.n + block - 1) / block;
    hipLaunchKernelGGL(k_dstrss_points<N>, dim3(grid), dim3(block), 0, stream,
                       reinterpret_cast<const DevMaterial<N>*>(dmat), C, dt, A);
    return hipGetLastError();
}

hipError_t launch_dstrss(int nss, const void* dmat, const DevControl& C, double dt, const PointArrays& A,
                         hipStream_t stream) {
    switch (nss) {
        case 12: return launch_dstrss_points<12>(dmat, C, dt, A, stream);
        case 18: return launch_dstrss_points<18>(dmat, C, dt, A, stream);
        default: return hipErrorInvalidValue;
    }
}

size_t material_bytes(int nss) {
    switch (nss) {
        case 12: return sizeof(DevMaterial<12>);
        case 18: return sizeof(DevMaterial<18>);
        default: return 0;
    }
}

void build_material(int nss, void* host, const cpfem_material& m, double toler) {
    if (nss == 12) fill_material<12>(*reinterpret_cast<DevMaterial<12>*>(host), m, toler);
    if (nss == 18) fill_material<18>(*reinterpret_cast<DevMaterial<18>*>(host), m, toler);
}

}  // namespace cpfem

using namespace cpfem;

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
    PointArrays A;
    memset(&A, 0, sizeof(A));
    const ptrdiff_t off = 0;
    A.n = n;
    A.iel = di.p;
    A.twinsys = di.p + N;
    A.timestwinned = di.p + 2 * N;
    A.err = di.p + 3 * N;
    A.tflag = di.p + 4 * N;
    A.errmax = di.p + 4 * N + 1;
    A.de = dd.p + (hde - h.data()) + off;
    A.r = dd.p + (hr - h.data());
    A.st = dd.p + (hst - h.data());
    A.ds = dd.p + (hds - h.data());
    A.t = dd.p + (ht - h.data());
    A.chi = dd.p + (hchi - h.data());
    A.deff = dd.p + (hdeff - h.data());
    A.erepmd = dd.p + (hrep - h.data());
    A.sv = dd.p + (hsv - h.data());
    A.dsv = dd.p + (hdsv - h.data());
    A.srate = dd.p + (hsr - h.data());
    A.flops = reinterpret_cast<long long*>(dl.p);
    A.transc = reinterpret_cast<long long*>(dl.p + N);
    A.rng_epoch = 0;
    A.gp_offset = 0;
    DevControl C = make_control(*ctl);
    CPFEM_HIP(launch_dstrss(nss, dm.p, C, dt, A, 0));
    CPFEM_HIP(hipDeviceSynchronize());
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
