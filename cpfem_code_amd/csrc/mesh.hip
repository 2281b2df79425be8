// Device-resident increment of FASOLT3: point kernels, colour-ordered element
// assembly, penalty boundary conditions, multicolour block SOR and the damped
// Newton update, behind the C ABI of include/cpfem.h.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "constitutive.hpp"
#include "kernels.hpp"
#include "mesh.hpp"

struct cpfem_ctx : cpfem::Ctx {};

namespace cpfem {

constexpr int kBlock = 256;
constexpr int kXcd = 8;   // MI355X: 8 XCDs, blocks dispatched to them round-robin

// Logical block of a launch of nb_pad = kXcd * per_xcd blocks: consecutive
// logical blocks (neighbouring rows, i.e. neighbouring mesh nodes) stay on one
// XCD, so each XCD's L2 holds the vector slice its rows gather from.
__device__ __forceinline__ int xcd_block(int per_xcd) {
    return (int)(blockIdx.x % kXcd) * per_xcd + (int)(blockIdx.x / kXcd);
}

// ---------------------------------------------------------------------------
// deterministic reductions: per-block partials, then one block in fixed order
// ---------------------------------------------------------------------------
template <int Q>
__device__ __forceinline__ void block_partials(double v[Q], double* partial, int slot) {
    __shared__ double sh[kBlock / 64][Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[q] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < Q; ++q) sh[wave][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double s = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w][q];
            partial[(size_t)slot * 4 + q] = s;
        }
    }
}

// Sums of the nblocks partial slots in a fixed order, by one block of kBlock
// threads; thread 0 receives them in out.
__device__ __forceinline__ void reduce_partials(const double* __restrict__ partial, int nblocks, double out[4]) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] += partial[(size_t)b * 4 + q];
    __shared__ double sh[kBlock][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) sh[threadIdx.x][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 4; ++q) {
            double s = 0.0;
            for (int t = 0; t < (int)blockDim.x; ++t) s += sh[t][q];
            out[q] = s;
        }
    }
}

__global__ void k_reduce_final(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
    double r[4];
    reduce_partials(partial, nblocks, r);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) out[q] = r[q];
}

// ---------------------------------------------------------------------------
// simple vector kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* __restrict__ p, size_t n, T v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_axpy_nodes(double* __restrict__ y, const double* __restrict__ x, double a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

// Sums of erepmd, flops and transcendentals over the points and the count of
// points whose slip Newton ended with FAFERR 987 (report values).
__global__ void k_point_sums(const double* __restrict__ erepmd, const long long* __restrict__ flops,
                             const long long* __restrict__ transc, const int* __restrict__ err, int n,
                             double* partial) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += erepmd[i];
        v[1] += (double)flops[i];
        v[2] += (double)transc[i];
        v[3] += err[i] == 987 ? 1.0 : 0.0;
    }
    block_partials<4>(v, partial, blockIdx.x);
}

// ---------------------------------------------------------------------------
// element assembly: one 64-lane wave per element, lane = node pair (i, j)
// ---------------------------------------------------------------------------
struct AsmArgs {
    int nel_color;
    int per_xcd;                 // logical blocks per XCD of the 8-node launch
    const int* elems;            // elements of this colour (element-block mode: element k is wave k)
    const int* elem_nodes;
    const int* elem_gps;
    const int* elem_lt;
    const int* elem_np;
    const int* scatter;          // entry 0 of each node pair's block (entry e at +64*e) ...
    const int* scat_start;       // ... from scat_start[element], pair (a, b) at a*nd + b
    const double* x;
    const double* chi;           // [36][ngps]
    const double* t;             // [6][ngps]
    int ngps;
    double* fc;                  // [3*nnod]
    double* efc;                 // element-block mode: [nel][8][3] forces of each element node
};

// BLOCKS = false: the element's blocks are added to the matrix through the
// scatter map and its node forces to fc (one launch per element colour).
// BLOCKS = true: one launch over all elements; the block of node pair
// (a, b) of element e goes to vals[e*576 + a*72 + q*8 + b] (q: entry; the
// 72 values of node a's pairs are contiguous, read by node a's matrix row
// only) and the forces of node a to efc[(e*8 + a)*3 + c], with plain
// stores; the gathers below then sum them into the matrix and fc.
template <typename T, bool BLOCKS>
__global__ __launch_bounds__(256) void k_assemble(const AsmArgs A, T* __restrict__ vals) {
    constexpr int EPB = 4;  // elements per block at most (blockDim.x / 64 of the launch: 1, 2 or 4)
    __shared__ T sh_dn[EPB][8][8][3];
    __shared__ T sh_w[EPB][8];
    __shared__ T sh_d[EPB][8][36];
    __shared__ T sh_s[EPB][8][6];
    __shared__ double sh_x[EPB][8][3];
    // internal forces: weight of each point and its stress times the node's
    // shape-function gradient, sum_k s(i,k) dN_a/dx_k, for every node a
    __shared__ double sh_wf[EPB][8];
    __shared__ double sh_sf[EPB][8][8][3];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // consecutive elements of the colour (neighbours in the mesh, whose
    // blocks share cache lines of the matrix) on one XCD, so that the partial
    // lines they write meet in that XCD's L2 before they are written back
    const int k = xcd_block(A.per_xcd) * (int)(blockDim.x >> 6) + wave;
    const bool live = k < A.nel_color;
    const int e = live ? (BLOCKS ? k : A.elems[k]) : 0;
    const int ng = live ? A.elem_np[e] : 0;
    const int lt = live ? A.elem_lt[e] : 13;
    if (live && lane < 24) {
        const int a = lane / 3, c = lane % 3;
        sh_x[wave][a][c] = A.x[3 * A.elem_nodes[kMaxNodes * e + a] + c];
    }
    __syncthreads();
    const size_t n = (size_t)A.ngps;
    if (live && lane < ng) {
        double xl[8][3], sfd[8][3];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) xl[a][c] = sh_x[wave][a][c];
        const double w = brick_shape(lt, lane, xl, sfd);
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) sh_dn[wave][lane][a][c] = (T)sfd[a][c];
        sh_w[wave][lane] = (T)w;
        const int gp = A.elem_gps[8 * e + lane];
#pragma unroll
        for (int q = 0; q < 36; ++q) sh_d[wave][lane][q] = (T)A.chi[q * n + gp];
        double st[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            st[q] = A.t[q * n + gp];
            sh_s[wave][lane][q] = (T)st[q];
        }
        sh_wf[wave][lane] = w;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            sh_sf[wave][lane][a][0] = st[0] * sfd[a][0] + st[5] * sfd[a][1] + st[4] * sfd[a][2];
            sh_sf[wave][lane][a][1] = st[1] * sfd[a][1] + st[3] * sfd[a][2] + st[5] * sfd[a][0];
            sh_sf[wave][lane][a][2] = st[2] * sfd[a][2] + st[4] * sfd[a][0] + st[3] * sfd[a][1];
        }
    }
    __syncthreads();
    if (!live) return;
    const int i = lane >> 3, j = lane & 7;
    T acc[3][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[p][q] = (T)0;
    for (int g = 0; g < ng; ++g) {
        T a[3], b[3], s[6], gm[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a[c] = sh_dn[wave][g][i][c];
            b[c] = sh_dn[wave][g][j][c];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) s[q] = sh_s[wave][g][q];
        pair_block<T>(a, b, &sh_d[wave][g][0], s, gm);
        const T w = sh_w[wave][g];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[p][q] = acc[p][q] + w * gm[p][q];
    }
    if (BLOCKS) {
        T* dst = vals + (size_t)e * 576 + i * 72 + j;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) dst[(3 * p + q) * 8] = acc[p][q];
    } else {
        const int off = A.scatter[A.scat_start[e] + lane];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                T* dst = vals + off + (3 * p + q) * 64;
                *dst = *dst + acc[p][q];
            }
    }
    // internal forces of node i (Elvals.f90:122-126), lanes with j == 0, from
    // the points' staged products, summed over the points in order
    if (j == 0) {
        double fl[3] = {0.0, 0.0, 0.0};
        for (int g = 0; g < ng; ++g) {
            const double w = sh_wf[wave][g];
#pragma unroll
            for (int c = 0; c < 3; ++c) fl[c] = fl[c] + w * sh_sf[wave][g][i][c];
        }
        if (BLOCKS) {
#pragma unroll
            for (int c = 0; c < 3; ++c) A.efc[((size_t)e * 8 + i) * 3 + c] = fl[c];
        } else {
            const int nd = A.elem_nodes[kMaxNodes * e + i];
#pragma unroll
            for (int c = 0; c < 3; ++c) A.fc[3 * nd + c] = A.fc[3 * nd + c] + fl[c];
        }
    }
}

// Gather assembly: the lists of the contributions of each matrix block and
// each node force, built at context creation.  A segment is one lane slot p
// of one row group (64 lanes, 9 entries each: the layout of the matrix
// values); lane l of it finds its contributions at list[start + k*64 + l],
// k < cnt (the largest count of the segment's lanes; -1 pads), in the order
// of the element colours.
struct GatherArgs {
    int nseg;            // segments = waves of the launch
    int per_xcd;         // logical blocks per XCD
    const int* order;    // [nseg] segment of each wave: the colours' groups of one mesh region in turn
    const int* start;    // [nseg]
    const int* cnt;      // [nseg]
    const int* list;     // element-block indices e*576 + a*72 + b
};

// Matrix values from the element blocks.  Each block starts from zero and adds
// its contributions in element-colour order, the order in which the
// colour-by-colour scatter adds them, so both give the same values bit for bit.
// The lanes of a row read the node pairs (a, b) of one element node a: 576
// contiguous bytes per element.
template <typename T>
__global__ __launch_bounds__(256) void k_gather_blocks(const GatherArgs G, const T* __restrict__ eblk,
                                                       T* __restrict__ vals) {
    const int lane = threadIdx.x & 63;
    const int w = xcd_block(G.per_xcd) * (int)(blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= G.nseg) return;   // wave-uniform
    const int sg = G.order[w];
    const int* li = G.list + G.start[sg] + lane;
    const int n = G.cnt[sg];
    T acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = (T)0;
    for (int k0 = 0; k0 < n; k0 += 4) {
        int id[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) id[u] = (k0 + u < n) ? li[(k0 + u) * 64] : -1;
        T v[4][9];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 9; ++q) v[u][q] = id[u] >= 0 ? eblk[(size_t)id[u] + q * 8] : (T)0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (id[u] >= 0)
#pragma unroll
                for (int q = 0; q < 9; ++q) acc[q] = acc[q] + v[u][q];
    }
    T* dst = vals + (size_t)sg * 576 + lane;
#pragma unroll
    for (int q = 0; q < 9; ++q) dst[q * 64] = acc[q];
}

// Node forces from the element node forces, in element-colour order; nodes
// in chunks of 64 (one wave each), entry k of node nd at
// list[start[nd/64] + k*64 + nd%64] (efc index, -1 pads).
__global__ void k_gather_forces(int nnod, const int* __restrict__ start, const int* __restrict__ cnt,
                                const int* __restrict__ list, const double* __restrict__ efc, double* __restrict__ fc) {
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= nnod) return;
    const int ch = nd >> 6, lane = nd & 63;
    const int n = cnt[ch];
    const int* li = list + start[ch] + lane;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < n; ++k) {
        const int id = li[k * 64];
        if (id >= 0)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] = acc[c] + efc[(size_t)id + c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) fc[3 * (size_t)nd + c] = acc[c];
}

// ---------------------------------------------------------------------------
// 20-node brick (SHAPE_22, Shapef.f90:210-342)
// ---------------------------------------------------------------------------
// Tabulated shape functions and local derivatives at the 2x2x2 points, as the
// reference holds them (6 decimals): the quadratic serendipity functions at
// +-g, g = 1/sqrt(3) in single precision, each rounded to 6 decimals.  Filled
// from the host at context creation.
__constant__ double kTab22N[8][20];
__constant__ double kTab22D[8][3][20];

// Local coordinates of the 20 nodes: bottom corners 1,3,5,7 and mid-sides
// 2,4,6,8 anticlockwise seen from the top, 9-12 halfway up the vertical edges,
// top face 13-20 like the bottom (Shapef.f90:191-200).
static const int kBrick20Loc[20][3] = {
    {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}, {-1, 1, -1}, {-1, 0, -1},
    {-1, -1, 0},  {1, -1, 0},  {1, 1, 0},   {-1, 1, 0},
    {-1, -1, 1},  {0, -1, 1},  {1, -1, 1},  {1, 0, 1},  {1, 1, 1},  {0, 1, 1},  {-1, 1, 1},  {-1, 0, 1}};
static const int kBrickSign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

static hipError_t upload_brick20_tables() {
    double tn[8][20], td[8][3][20];
    const double g = (double)(float)(1.0 / sqrt(3.0));
    auto r6 = [](double v) { return round(v * 1.0e6) / 1.0e6; };
    for (int ig = 0; ig < 8; ++ig) {
        double xi[3];
        for (int k = 0; k < 3; ++k) xi[k] = kBrickSign[ig][k] * g;
        for (int a = 0; a < 20; ++a) {
            const int* c = kBrick20Loc[a];
            double n, d[3];
            if (c[0] != 0 && c[1] != 0 && c[2] != 0) {
                double f[3], sum = -2.0;
                for (int k = 0; k < 3; ++k) {
                    f[k] = 1.0 + xi[k] * c[k];
                    sum += xi[k] * c[k];
                }
                n = f[0] * f[1] * f[2] * sum / 8.0;
                for (int k = 0; k < 3; ++k)
                    d[k] = c[k] * f[(k + 1) % 3] * f[(k + 2) % 3] * sum / 8.0 + f[0] * f[1] * f[2] * c[k] / 8.0;
            } else {
                int z = 0;
                while (c[z] != 0) ++z;
                const int p = (z + 1) % 3, q = (z + 2) % 3;
                const double qz = 1.0 - xi[z] * xi[z], fp = 1.0 + xi[p] * c[p], fq = 1.0 + xi[q] * c[q];
                n = qz * fp * fq / 4.0;
                d[z] = -2.0 * xi[z] * fp * fq / 4.0;
                d[p] = qz * c[p] * fq / 4.0;
                d[q] = qz * fp * c[q] / 4.0;
            }
            tn[ig][a] = r6(n);
            for (int k = 0; k < 3; ++k) td[ig][k][a] = r6(d[k]);
        }
    }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(kTab22N), tn, sizeof(tn));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(kTab22D), td, sizeof(td));
    return e;
}

// SHAPE_21 (Shapef.f90:104-189) and SHAPE_23 (:360-485) tabulate their local
// derivatives at the points to 6 decimals as well.  Generated here from the
// quadratic functions:
//  - 10-node tetrahedron: volume coordinates of corners 1,3,5,10 with local
//    axes xi1 = L10, xi2 = L3, xi3 = L1, L5 = 1 - xi1 - xi2 - xi3; corner
//    N = L(2L - 1), mid-edge N = 4 La Lb (2: 1-3, 4: 3-5, 6: 1-5, 7: 1-10,
//    8: 3-10, 9: 5-10); point k has L = (5 + 3 sqrt 5)/20 at the k-th corner
//    and (5 - sqrt 5)/20 at the others (this reproduces the whole table);
//  - 15-node pentahedron: area coordinates L1 = xi1, L3 = xi2, L5 = 1 - xi1 -
//    xi2 of base corners 1,3,5 (top corners 10,12,14 above them), zeta = xi3;
//    corner N = L((2L - 1)(1 + zeta zc) - (1 - zeta^2))/2, triangle-edge N =
//    2 La Lb (1 + zeta zc), vertical-edge N = L(1 - zeta^2); points 1-3 at
//    zeta = -1/sqrt 3 nearest corners 1,3,5 (area coordinates 2/3, 1/6, 1/6),
//    4-6 at +1/sqrt 3 (60 of the 270 tabulated values differ from this by
//    1e-6 or 2e-6, the table's own rounding).
__constant__ double kTab21D[4][3][10];
__constant__ double kTab23D[6][3][15];

static hipError_t upload_tet_wedge_tables() {
    auto r6 = [](double v) { return round(v * 1.0e6) / 1.0e6; };
    double t21[4][3][10], t23[6][3][15];
    {
        static const int corner_pos[4] = {0, 2, 4, 9};
        static const int mids[6][3] = {{1, 0, 1}, {3, 1, 2}, {5, 0, 2}, {6, 0, 3}, {7, 1, 3}, {8, 2, 3}};
        static const double dl[4][3] = {{0, 0, 1}, {0, 1, 0}, {-1, -1, -1}, {1, 0, 0}};
        const double a = (5.0 + 3.0 * sqrt(5.0)) / 20.0, b = (5.0 - sqrt(5.0)) / 20.0;
        for (int ig = 0; ig < 4; ++ig) {
            double L[4];
            for (int c = 0; c < 4; ++c) L[c] = (c == ig) ? a : b;
            for (int c = 0; c < 4; ++c)
                for (int j = 0; j < 3; ++j) t21[ig][j][corner_pos[c]] = r6((4.0 * L[c] - 1.0) * dl[c][j]);
            for (int m = 0; m < 6; ++m) {
                const int p = mids[m][1], q = mids[m][2];
                for (int j = 0; j < 3; ++j) t21[ig][j][mids[m][0]] = r6(4.0 * (L[q] * dl[p][j] + L[p] * dl[q][j]));
            }
        }
    }
    {
        // node -> (kind: 0 corner, 1 triangle edge, 2 vertical edge; corners p, q of the base; zc)
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
                double d[3];
                if (kind == 0) {
                    const double l = L[p];
                    const double dn = 0.5 * ((4.0 * l - 1.0) * (1.0 + z * zc) - (1.0 - z * z));
                    d[0] = dn * dl[p][0];
                    d[1] = dn * dl[p][1];
                    d[2] = 0.5 * l * ((2.0 * l - 1.0) * zc + 2.0 * z);
                } else if (kind == 1) {
                    const double f = 1.0 + z * zc;
                    for (int j = 0; j < 2; ++j) d[j] = 2.0 * f * (L[q] * dl[p][j] + L[p] * dl[q][j]);
                    d[2] = 2.0 * L[p] * L[q] * zc;
                } else {
                    for (int j = 0; j < 2; ++j) d[j] = (1.0 - z * z) * dl[p][j];
                    d[2] = -2.0 * z * L[p];
                }
                for (int j = 0; j < 3; ++j) t23[ig][j][a] = r6(d[j]);
            }
        }
    }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(kTab21D), t21, sizeof(t21));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(kTab23D), t23, sizeof(t23));
    return e;
}

// Local derivative table entry d N_a / d xi_j at point ig of the NN-node
// tabulated element.
template <int NN>
__device__ __forceinline__ double tab_dn(int ig, int j, int a) {
    if constexpr (NN == 20)
        return kTab22D[ig][j][a];
    else if constexpr (NN == 10)
        return kTab21D[ig][j][a];
    else
        return kTab23D[ig][j][a];
}

// Global derivatives and weight of a tabulated element at point ig: detJ
// (SHAPE_22), -0.0416667 detJ (SHAPE_21, Shapef.f90:187) or 0.16666667 detJ
// (SHAPE_23, Shapef.f90:483).
template <int NN>
__device__ __forceinline__ double tab_shape(int ig, const double xl[NN][3], double sfd[NN][3]) {
    double jac[3][3];  // jac[j][k] = sum_a dndn(j,a) x(k,a)
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NN; ++a) s = s + tab_dn<NN>(ig, j, a) * xl[a][k];
            jac[j][k] = s;
        }
    const double det = invert_3(jac);
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) s = s + jac[k][j] * tab_dn<NN>(ig, j, a);
            sfd[a][k] = s;
        }
    if constexpr (NN == 20)
        return 1.0 * det;
    else if constexpr (NN == 10)
        return -(0.0416667 * det);
    else
        return 0.16666667 * det;
}

// ---------------------------------------------------------------------------
// point kinematics (Elvals.f90:87-112): shape functions, weight check (920)
// and DSTRNS (905, 906) for the points of NN-node elements, one per thread
// ---------------------------------------------------------------------------
struct KinArgs {
    int n;
    const int* gp_elem;
    const int* gp_local;
    const int* elem_nodes;
    const int* elem_lt;
    const int* elem_nd;
    const double* x;
    const double* dx;
    double smale, smalr;
    double* de;    // [6][n]
    double* r;     // [3][n]
    int* kerr;     // [n]
};

template <int NN>
__global__ __launch_bounds__(kBlock) void k_kin(const KinArgs K) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= K.n) return;
    const int e = K.gp_elem[g];
    if (K.elem_nd[e] != NN) return;
    double xl[NN][3], dxl[NN][3], sfd[NN][3];
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        const int nd = K.elem_nodes[kMaxNodes * e + a];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xl[a][k] = K.x[3 * nd + k];
            dxl[a][k] = K.dx[3 * nd + k];
        }
    }
    double w;
    if constexpr (NN == 8)
        w = brick_shape(K.elem_lt[e], K.gp_local[g], xl, sfd);
    else
        w = tab_shape<NN>(K.gp_local[g], xl, sfd);
    int kerr = (w < 1.e-7) ? 920 : 0;
    double de[6], r[3], deff;
    const int e2 = dstrns<NN>(sfd, dxl, K.smale, K.smalr, de, r, deff);
    if (e2 > kerr) kerr = e2;
    const size_t n = (size_t)K.n;
#pragma unroll
    for (int i = 0; i < 6; ++i) K.de[i * n + g] = de[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) K.r[i * n + g] = r[i];
    K.kerr[g] = kerr;
}

// ---------------------------------------------------------------------------
// volume averages over the points (the macroscopic state): point weights are
// the integration weights at the current coordinates; four sums per launch,
// group 0: (weight, weight*eps, weight*st11, weight*st22), 1: (st33, st23,
// st31, st12), 2: (ep11, ep22, ep33, ep23), 3: (ep31, ep12, 0, 0)
// ---------------------------------------------------------------------------
// Integration weight of point g (element e, NN nodes) at the current coordinates.
template <int NN>
__device__ __noinline__ double point_weight(const KinArgs& K, int g, int e) {
    double xl[NN][3], sfd[NN][3];
#pragma unroll
    for (int a = 0; a < NN; ++a) {
        const int nd = K.elem_nodes[kMaxNodes * e + a];
#pragma unroll
        for (int k = 0; k < 3; ++k) xl[a][k] = K.x[3 * nd + k];
    }
    if constexpr (NN == 8)
        return brick_shape(K.elem_lt[e], K.gp_local[g], xl, sfd);
    else
        return tab_shape<NN>(K.gp_local[g], xl, sfd);
}

__global__ __launch_bounds__(kBlock) void k_vol_sums(const KinArgs K, const double* __restrict__ st,
                                                     const double* __restrict__ ep, const double* __restrict__ eps,
                                                     int group, double* partial, int slot0) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < K.n) {
        const int e = K.gp_elem[g];
        const int nn = K.elem_nd[e];
        if (nn == 8 || nn == 10 || nn == 15 || nn == 20) {
            const double w = nn == 8    ? point_weight<8>(K, g, e)
                             : nn == 10 ? point_weight<10>(K, g, e)
                             : nn == 15 ? point_weight<15>(K, g, e)
                                        : point_weight<20>(K, g, e);
            const size_t n = (size_t)K.n;
            if (group == 0) {
                v[0] = w;
                v[1] = w * eps[g];
                v[2] = w * st[0 * n + g];
                v[3] = w * st[1 * n + g];
            } else if (group == 1) {
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] = w * st[(2 + q) * n + g];
            } else if (group == 2) {
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] = w * ep[q * n + g];
            } else {
                v[0] = w * ep[4 * n + g];
                v[1] = w * ep[5 * n + g];
            }
        }
    }
    block_partials<4>(v, partial, slot0 + blockIdx.x);
}

// Transpose of the point kinematics of 8-node bricks (DSTRNS, Sig_eps.f90:40-50):
// for cotangents deb (engineering shears) and rb of every point's strain and
// rotation increments, out[3*node+k] += sum over the element's points of
// sum_i dN_a/dx_i Gbar(i,k), where Gbar is the cotangent of the displacement
// gradient d(i,k) = sum_a dN_a/dx_i du_k.  One thread per element of one
// colour (no two share a node).
__global__ __launch_bounds__(kBlock) void k_kin_transpose8(const KinArgs K, int nel_color, const int* __restrict__ elems,
                                                           const int* __restrict__ elem_gps,
                                                           const double* __restrict__ cot, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel_color) return;
    const int e = elems[t];
    double xl[8][3], sfd[8][3], ub[8][3];
    for (int a = 0; a < 8; ++a) {
        const int nd = K.elem_nodes[kMaxNodes * e + a];
        for (int k = 0; k < 3; ++k) {
            xl[a][k] = K.x[3 * nd + k];
            ub[a][k] = 0.0;
        }
    }
    const size_t n = (size_t)K.n;
    for (int q = 0; q < 8; ++q) {
        const int g = elem_gps[8 * e + q];
        if (g < 0) continue;
        (void)brick_shape(K.elem_lt[e], K.gp_local[g], xl, sfd);
        double gb[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1) % 3, k = (j + 1) % 3;
            const double deb = cot[i * n + g], dsh = cot[(i + 3) * n + g], rb = cot[(6 + i) * n + g];
            gb[i][i] += deb;
            gb[k][j] += dsh + 0.5 * rb;   // del(i+3) = d(k,j) + d(j,k), rl(i) = (d(k,j) - d(j,k))/2
            gb[j][k] += dsh - 0.5 * rb;
        }
        for (int a = 0; a < 8; ++a)
            for (int k = 0; k < 3; ++k) ub[a][k] += sfd[a][0] * gb[0][k] + sfd[a][1] * gb[1][k] + sfd[a][2] * gb[2][k];
    }
    for (int a = 0; a < 8; ++a) {
        const int nd = K.elem_nodes[kMaxNodes * e + a];
        for (int k = 0; k < 3; ++k) out[3 * nd + k] += ub[a][k];
    }
}

__global__ __launch_bounds__(kBlock) void k_point_weights(const KinArgs K, double* __restrict__ w) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= K.n) return;
    const int e = K.gp_elem[g];
    const int nn = K.elem_nd[e];
    w[g] = nn == 8    ? point_weight<8>(K, g, e)
           : nn == 10 ? point_weight<10>(K, g, e)
           : nn == 15 ? point_weight<15>(K, g, e)
           : nn == 20 ? point_weight<20>(K, g, e)
                      : 0.0;
}

// ---------------------------------------------------------------------------
// Assembly of the tabulated quadratic elements (20-node bricks, 10-node
// tetrahedra, 15-node pentahedra): one 256-thread block per element, thread
// = node pair (i, j) of the NN^2 (several per thread above 256)
// ---------------------------------------------------------------------------
template <typename T, int NN>
__global__ __launch_bounds__(256) void k_assemble_tab(const AsmArgs A, T* __restrict__ vals) {
    __shared__ double sh_x[NN][3];
    __shared__ double sh_dn64[8][NN][3];
    __shared__ double sh_w64[8];
    __shared__ double sh_s64[8][6];
    __shared__ T sh_dn[8][NN][3];
    __shared__ T sh_w[8];
    __shared__ T sh_d[8][36];
    __shared__ T sh_s[8][6];
    const int e = A.elems[blockIdx.x];
    const int tid = threadIdx.x;
    const int ng = A.elem_np[e];
    if (tid < 3 * NN) sh_x[tid / 3][tid % 3] = A.x[3 * A.elem_nodes[kMaxNodes * e + tid / 3] + tid % 3];
    __syncthreads();
    const size_t n = (size_t)A.ngps;
    if (tid < ng) {
        double xl[NN][3], sfd[NN][3];
#pragma unroll
        for (int a = 0; a < NN; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) xl[a][c] = sh_x[a][c];
        const double w = tab_shape<NN>(tid, xl, sfd);
#pragma unroll
        for (int a = 0; a < NN; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                sh_dn64[tid][a][c] = sfd[a][c];
                sh_dn[tid][a][c] = (T)sfd[a][c];
            }
        sh_w64[tid] = w;
        sh_w[tid] = (T)w;
        const int gp = A.elem_gps[8 * e + tid];
#pragma unroll
        for (int q = 0; q < 36; ++q) sh_d[tid][q] = (T)A.chi[q * n + gp];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double v = A.t[q * n + gp];
            sh_s64[tid][q] = v;
            sh_s[tid][q] = (T)v;
        }
    }
    __syncthreads();
    const int base = A.scat_start[e];
    for (int p = tid; p < NN * NN; p += 256) {
        const int i = p / NN, j = p - NN * (p / NN);
        T acc[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) acc[a][b] = (T)0;
        for (int g = 0; g < ng; ++g) {
            T a[3], b[3], sg[6], gm[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a[c] = sh_dn[g][i][c];
                b[c] = sh_dn[g][j][c];
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) sg[q] = sh_s[g][q];
            pair_block<T>(a, b, &sh_d[g][0], sg, gm);
            const T w = sh_w[g];
#pragma unroll
            for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2) acc[a2][b2] = acc[a2][b2] + w * gm[a2][b2];
        }
        const int off = A.scatter[base + p];
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
            for (int b2 = 0; b2 < 3; ++b2) {
                T* dst = vals + off + (3 * a2 + b2) * 64;
                *dst = *dst + acc[a2][b2];
            }
    }
    // internal forces of node tid (Elvals.f90:122-126)
    if (tid < NN) {
        double fl[3] = {0.0, 0.0, 0.0};
        for (int g = 0; g < ng; ++g) {
            const double w = sh_w64[g];
            const double* s = sh_s64[g];
            const double* d = sh_dn64[g][tid];
            fl[0] = fl[0] + w * (s[0] * d[0] + s[5] * d[1] + s[4] * d[2]);
            fl[1] = fl[1] + w * (s[1] * d[1] + s[3] * d[2] + s[5] * d[0]);
            fl[2] = fl[2] + w * (s[2] * d[2] + s[4] * d[0] + s[3] * d[1]);
        }
        const int nd = A.elem_nodes[kMaxNodes * e + tid];
#pragma unroll
        for (int c = 0; c < 3; ++c) A.fc[3 * nd + c] = A.fc[3 * nd + c] + fl[c];
    }
}

// ---------------------------------------------------------------------------
// 'tie' elements (no points, Elvals.f90:62-77): penalty efix on DOF k of the
// first node and -efix coupling it to DOF k of the second; lt 1-3 tie DOF lt,
// lt 4-6 the two DOFs 1+mod(lt-3,3) and the next one.  One thread per element.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_ties(int nt, const int* __restrict__ elems, const int* __restrict__ elem_lt,
                       const int* __restrict__ scatter, const int* __restrict__ scat_start, double efix,
                       T* __restrict__ vals) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int e = elems[t];
    const int itp = elem_lt[e];
    const int b = scat_start[e];
    const int off11 = scatter[b + 0], off12 = scatter[b + 1];  // pairs (1,1) and (1,2) of the 2 nodes
    int dofs[2] = {-1, -1};
    if (itp >= 1 && itp <= 3) {
        dofs[0] = itp - 1;
    } else if (itp >= 4 && itp <= 6) {
        dofs[0] = (itp - 3) % 3;
        dofs[1] = (dofs[0] + 1) % 3;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int k = dofs[q];
        if (k < 0) continue;
        vals[off11 + 4 * k * 64] = vals[off11 + 4 * k * 64] + (T)efix;
        vals[off12 + 4 * k * 64] = vals[off12 + 4 * k * 64] + (T)(-efix);
    }
}

// ---------------------------------------------------------------------------
// SUVALS: rhs = fc + penalty terms; penalty stiffness on the diagonal blocks
// ---------------------------------------------------------------------------
struct BcArgs {
    int nnod;
    const int* ib;
    const int* diag_off;  // [nnod] entry 0 of the node's diagonal block; entry e at +64*e
    const double* x;
    const double* dx;
    const double* fc;
    const double* xs;   // x_spec of this increment [99][15]
    const double* fs;   // f_spec of this increment [99][3]
    double efix;
    double* rhs;
    const int* inc_dev;   // non-null: xs, fs are the whole tables and the increment is read here
};

template <typename T>
__global__ void k_suvals(BcArgs A, T* __restrict__ vals, double* partial) {
    const int nd = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (A.inc_dev) {
        const size_t k = (size_t)(*A.inc_dev - 1);
        A.xs += k * 99 * 15;
        A.fs += k * 99 * 3;
    }
    if (nd < A.nnod) {
        const int idof = 3 * nd;
        double r[3] = {A.fc[idof], A.fc[idof + 1], A.fc[idof + 2]};
        const int ibc = A.ib[nd];
        if (ibc >= 100) {
            const int i_kin = (ibc / 100) % 100;
            if (i_kin > 0) {
                const double* xs = A.xs + (size_t)(i_kin - 1) * 15;
                double xspec[3];
                if (i_kin < 50) {
                    xspec[0] = xs[6];
                    xspec[1] = xs[10];
                    xspec[2] = xs[14];
                } else {
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double s = 0.0;
#pragma unroll
                        for (int k = 0; k < 3; ++k) s = s + xs[6 + 3 * j + k] * (A.x[idof + k] + A.dx[idof + k]);
                        xspec[j] = s;
                    }
                }
                const int point[3][3] = {{0, 1, 3}, {1, 2, 4}, {3, 4, 5}};
                T* dg = vals + A.diag_off[nd];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const double add = A.efix * xs[point[i][j]];
                        T* dst = dg + (3 * i + j) * 64;
                        *dst = *dst + (T)add;
                    }
                    double sum = 0.0;
#pragma unroll
                    for (int j = 0; j < 3; ++j) sum = sum + A.efix * xs[point[i][j]] * (A.dx[idof + j] - xspec[j]);
                    r[i] = r[i] + sum;
                }
            }
            const int i_spf = (ibc / 10000) % 100;
            if (i_spf > 0) {
#pragma unroll
                for (int i = 0; i < 3; ++i) r[i] = r[i] + A.fs[(size_t)(i_spf - 1) * 3 + i];
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) A.rhs[idof + i] = r[i];
        if (ibc < 10) v[0] = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        v[1] = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    }
    block_partials<4>(v, partial, blockIdx.x);
}

// ---------------------------------------------------------------------------
// multicolour block SOR
// ---------------------------------------------------------------------------
struct SorArgs {
    int nrows;           // rows of the launch (one colour)
    int gbase;           // first row group of the launch
    int per_xcd;         // logical blocks per XCD of the launch
    int P;               // slots per lane
    const int* cols;     // [group][P][64]
    double omega;
    const int* stop;     // device flag: nonzero -> skip the sweep
    const double* omega_dev = nullptr;   // non-null: the relaxation factor read on the device
    const int* gate = nullptr;           // non-null: the residual runs only while *gate != 0
};

constexpr int kWavesPerBlock = kBlock / 64;

// Row sums of one group: lane (l, i) multiplies its slots p*kRowLanes + l of row
// i with the gathered y and the kRowLanes partial sums of a row are combined on
// its l == 0 lane.  Matrix values and columns are streamed with nontemporal
// loads (each is read once per sweep); y stays cached.  With DIAG the slot 0
// block is part of the sum (residual); without it lane l == 0 keeps the
// diagonal block in d and the row's node in nd (relaxation).
template <typename T, typename V, int PC, bool DIAG>
__device__ __forceinline__ void group_row_sums(const T* __restrict__ vg, const int* __restrict__ cg, int Prt, int l,
                                               const V* __restrict__ y, T sum[3], T d[9], int& nd) {
    const int P = PC > 0 ? PC : Prt;
    if (PC > 0) {
        int c[PC > 0 ? PC : 1];
        T m[PC > 0 ? PC : 1][9];
        T yv[PC > 0 ? PC : 1][3];
#pragma unroll
        for (int p = 0; p < PC; ++p) c[p] = __builtin_nontemporal_load(cg + p * 64);
#pragma unroll
        for (int p = 0; p < PC; ++p)
#pragma unroll
            for (int e = 0; e < 9; ++e) m[p][e] = __builtin_nontemporal_load(vg + (p * 9 + e) * 64);
#pragma unroll
        for (int p = 0; p < PC; ++p)
#pragma unroll
            for (int a = 0; a < 3; ++a) yv[p][a] = (T)y[3 * c[p] + a];
#pragma unroll
        for (int p = 0; p < PC; ++p) {
            if (p == 0 && l == 0) {
#pragma unroll
                for (int e = 0; e < 9; ++e) d[e] = m[0][e];
                nd = c[0];
                if (!DIAG) continue;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
                sum[a] = sum[a] + m[p][3 * a + 0] * yv[p][0] + m[p][3 * a + 1] * yv[p][1] + m[p][3 * a + 2] * yv[p][2];
        }
    } else {
        for (int p = 0; p < P; ++p) {
            const int c = __builtin_nontemporal_load(cg + p * 64);
            T m[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) m[e] = __builtin_nontemporal_load(vg + (p * 9 + e) * 64);
            const T y0 = (T)y[3 * c], y1 = (T)y[3 * c + 1], y2 = (T)y[3 * c + 2];
            if (p == 0 && l == 0) {
#pragma unroll
                for (int e = 0; e < 9; ++e) d[e] = m[e];
                nd = c;
                if (!DIAG) continue;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) sum[a] = sum[a] + m[3 * a + 0] * y0 + m[3 * a + 1] * y1 + m[3 * a + 2] * y2;
        }
    }
}

// The same for PC > 0 with the group's column indices already loaded (cpre):
// the colour kernel issues them before it reads its stop flag.
template <typename T, typename V, int PC>
__device__ __forceinline__ void group_row_sums_pre(const T* __restrict__ vg, const int (&cpre)[PC], int l,
                                                   const V* __restrict__ y, T sum[3], T d[9], int& nd) {
    T m[PC][9];
    T yv[PC][3];
#pragma unroll
    for (int p = 0; p < PC; ++p)
#pragma unroll
        for (int e = 0; e < 9; ++e) m[p][e] = __builtin_nontemporal_load(vg + (p * 9 + e) * 64);
#pragma unroll
    for (int p = 0; p < PC; ++p)
#pragma unroll
        for (int a = 0; a < 3; ++a) yv[p][a] = (T)y[3 * cpre[p] + a];
#pragma unroll
    for (int p = 0; p < PC; ++p) {
        if (p == 0 && l == 0) {
#pragma unroll
            for (int e = 0; e < 9; ++e) d[e] = m[0][e];
            nd = cpre[0];
            continue;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
            sum[a] = sum[a] + m[p][3 * a + 0] * yv[p][0] + m[p][3 * a + 1] * yv[p][1] + m[p][3 * a + 2] * yv[p][2];
    }
}

// FP64 row sums with three slots per lane in two phases: slots 0 and 1, then
// slot 2 once their products are summed (the compiler barrier keeps slot 2's
// loads behind them), so that at most two slots' values are live per lane.
// The products are summed in the same order as in group_row_sums_pre.
template <typename V>
__device__ __forceinline__ void group_row_sums_2ph(const double* __restrict__ vg, const int (&cpre)[3], int l,
                                                   const V* __restrict__ y, double sum[3], double d[9], int& nd) {
    {
        double m[2][9];
        double yv[2][3];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int e = 0; e < 9; ++e) m[p][e] = __builtin_nontemporal_load(vg + (p * 9 + e) * 64);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int a = 0; a < 3; ++a) yv[p][a] = (double)y[3 * cpre[p] + a];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            if (p == 0 && l == 0) {
#pragma unroll
                for (int e = 0; e < 9; ++e) d[e] = m[0][e];
                nd = cpre[0];
                continue;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
                sum[a] = sum[a] + m[p][3 * a + 0] * yv[p][0] + m[p][3 * a + 1] * yv[p][1] + m[p][3 * a + 2] * yv[p][2];
        }
    }
    asm volatile("" ::: "memory");
    {
        double m[9];
        double yv[3];
#pragma unroll
        for (int e = 0; e < 9; ++e) m[e] = __builtin_nontemporal_load(vg + (2 * 9 + e) * 64);
#pragma unroll
        for (int a = 0; a < 3; ++a) yv[a] = (double)y[3 * cpre[2] + a];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            sum[a] = sum[a] + m[3 * a + 0] * yv[0] + m[3 * a + 1] * yv[1] + m[3 * a + 2] * yv[2];
    }
}

// Sum of the kRowLanes partial sums of row i onto its lane l == 0 (lane i).
// In FP64 each component is summed before the next one's shuffles are issued
// (the empty asm takes the sum as an operand), which bounds the registers
// that hold shuffled values to one component's.
template <typename T>
__device__ __forceinline__ void row_combine(T sum[3], int lane) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T acc = sum[a];
#pragma unroll
        for (int q = 1; q < kRowLanes; ++q) acc = acc + __shfl(sum[a], (lane + q * kGroupRows) & 63, 64);
        if constexpr (sizeof(T) == 8) asm volatile("" : "+v"(acc));
        sum[a] = acc;
    }
}

// One colour of one sweep on A y = b (y, b nodal vectors in DOF order), one
// row group per wave.  After the row sums, lane l == 0 of each row relaxes the
// node's three DOFs in x, y, z order with the fresh values of the earlier ones,
// as the reference's natural-order sweep does within a node.  PC > 0 fixes the
// slots per lane at compile time (3 for the 27-block rows of a brick mesh) so
// that every load of a lane is issued before the first one is consumed.  With
// ACC the kernel also accumulates |y_new - y_old|^2 and |y_new|^2 for the
// relative-update test.
template <typename T, typename V, bool ACC, int PC>
__global__ __launch_bounds__(kBlock) void k_sor_color(const SorArgs A, const T* __restrict__ vals,
                                                      const V* __restrict__ b, V* __restrict__ y, double* partial,
                                                      int pslot0) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int lane = threadIdx.x & 63;
    // the wave's row group, uniform over the wave: kept in scalar registers so
    // that loads can address it as a scalar base plus the lane's offset
    const int w = __builtin_amdgcn_readfirstlane(xcd_block(A.per_xcd) * (int)(blockDim.x >> 6) + (threadIdx.x >> 6));
    const int l = lane / kGroupRows, i = lane - l * kGroupRows;
    const bool live = (l < kRowLanes) && (w * kGroupRows + i < A.nrows);
    const size_t g = (size_t)A.gbase + w;
    // With the slots per lane fixed at compile time the group's column indices
    // are loaded before the stop flag is read, so that the two memory round
    // trips of a launch's critical path (flag, then columns, then the gather of
    // y) overlap; a skipped launch reads only the columns.
    int cpre[PC > 0 ? PC : 1];
    if (PC > 0 && live) {
        const int* cg = A.cols + g * (size_t)A.P * 64 + lane;
#pragma unroll
        for (int p = 0; p < (PC > 0 ? PC : 1); ++p) cpre[p] = __builtin_nontemporal_load(cg + p * 64);
    }
    if (*A.stop) return;   // uniform: the flag is written by earlier kernels only
    T sum[3] = {(T)0, (T)0, (T)0};
    T d[9];
    int nd = 0;
    if (live) {
        if constexpr (PC == 3 && sizeof(T) == 8)
            group_row_sums_2ph<V>(vals + g * (size_t)A.P * 576 + lane, cpre, l, y, sum, d, nd);
        else if constexpr (PC > 0)
            group_row_sums_pre<T, V, PC>(vals + g * (size_t)A.P * 576 + lane, cpre, l, y, sum, d, nd);
        else
            group_row_sums<T, V, PC, false>(vals + g * (size_t)A.P * 576 + lane, A.cols + g * (size_t)A.P * 64 + lane,
                                            A.P, l, y, sum, d, nd);
    }
    row_combine<T>(sum, lane);
    if (live && l == 0) {
        T yn[3] = {(T)y[3 * nd], (T)y[3 * nd + 1], (T)y[3 * nd + 2]};
        const T om = (T)(A.omega_dev ? *A.omega_dev : A.omega);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T s = sum[a];
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (k != a) s = s + d[3 * a + k] * yn[k];
            const T old = yn[a];
            const T nw = old + om * (((T)b[3 * nd + a] - s) / d[3 * a + a] - old);
            yn[a] = nw;
            if (ACC) {
                const double dd = (double)nw - (double)old;
                v[0] += dd * dd;
                v[1] += (double)nw * (double)nw;
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) y[3 * nd + a] = (V)yn[a];
    }
    if (ACC) block_partials<4>(v, partial, pslot0 + blockIdx.x);
}

// Residual r = b - A y in FP64 over all row groups (A read in its stored
// precision, products accumulated in FP64).  Partial sums: |D^-1 r|^2 and
// |D^-1 b|^2 with D the scalar diagonal, so that penalty rows (efix ~ 1e20)
// and stiffness rows are measured in the same displacement units; the
// unscaled |r|^2 and |b|^2 are kept in slots 2, 3.  Padded rows carry column
// -1 and are skipped.
template <typename T, int PC>
__global__ __launch_bounds__(kBlock) void k_residual(const SorArgs A, const T* __restrict__ vals,
                                                     const double* __restrict__ b, const double* __restrict__ y,
                                                     double* __restrict__ res, double* partial) {
    if (A.gate && *A.gate == 0) return;   // uniform over the grid
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int lane = threadIdx.x & 63;
    const int w = xcd_block(A.per_xcd) * kWavesPerBlock + (threadIdx.x >> 6);
    const int l = lane / kGroupRows;
    const size_t g = (size_t)w;
    bool live = (l < kRowLanes) && (w < A.nrows);   // nrows = number of groups here
    const int* cg = A.cols + g * (size_t)A.P * 64 + lane;
    if (live) live = cg[0] >= 0;
    double sum[3] = {0.0, 0.0, 0.0};
    double d[9];
    int nd = 0;
    if (live) {
        const T* vg = vals + g * (size_t)A.P * 576 + lane;
        for (int p = 0; p < (PC > 0 ? PC : A.P); ++p) {
            const int c = __builtin_nontemporal_load(cg + p * 64);
            double m[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) m[e] = (double)__builtin_nontemporal_load(vg + (p * 9 + e) * 64);
            if (p == 0 && l == 0) {
#pragma unroll
                for (int e = 0; e < 9; ++e) d[e] = m[e];
                nd = c;
            }
            const double y0 = y[3 * c], y1 = y[3 * c + 1], y2 = y[3 * c + 2];
#pragma unroll
            for (int a = 0; a < 3; ++a) sum[a] += m[3 * a + 0] * y0 + m[3 * a + 1] * y1 + m[3 * a + 2] * y2;
        }
    }
    row_combine<double>(sum, lane);
    if (live && l == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ra = b[3 * nd + a] - sum[a];
            const double dd = d[4 * a];
            if (res) res[3 * nd + a] = ra;
            v[0] += (ra / dd) * (ra / dd);
            v[1] += (b[3 * nd + a] / dd) * (b[3 * nd + a] / dd);
            v[2] += ra * ra;
            v[3] += b[3 * nd + a] * b[3 * nd + a];
        }
    }
    block_partials<4>(v, partial, blockIdx.x);
}

__global__ void k_round_vals(const double* __restrict__ a, float* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)__builtin_nontemporal_load(a + i);
}

__global__ void k_to_float(const double* __restrict__ a, float* __restrict__ b, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_add_float(double* __restrict__ y, const float* __restrict__ d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + (double)d[i];
}

// Stop flag of the relative-update test (Stif_sol.f90:208-210) after a sweep.
__global__ void k_sor_check(const double* __restrict__ partial, int nblocks, double tol2, int* stop, int* sweeps,
                            double* err_out, int maxit) {
    __shared__ double sh[kBlock][2];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
        a += partial[(size_t)i * 4];
        b += partial[(size_t)i * 4 + 1];
    }
    sh[threadIdx.x][0] = a;
    sh[threadIdx.x][1] = b;
    __syncthreads();
    if (threadIdx.x == 0 && *stop == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            s0 += sh[t][0];
            s1 += sh[t][1];
        }
        // a non-finite |y|^2 (NaN or Inf in the iterate) makes the measure NaN
        const double err = !isfinite(s1) ? s1 - s1 : (s1 > 0.0 ? s0 / s1 : 0.0);
        *err_out = err;
        *sweeps += 1;
        // a non-finite measure (a matrix or right-hand side with NaN or Inf
        // entries, or sweeps that diverge) ends the solve at once
        if (err < tol2 || *sweeps >= maxit || !isfinite(err)) *stop = 1;
    }
}

// NR update of Stif_sol.f90:70-74.
__global__ void k_nr_update(double* __restrict__ x, double* __restrict__ dx, const double* __restrict__ ddx,
                            double alp2, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double xc = alp2 * ddx[i];
        x[i] = x[i] - xc;
        dx[i] = dx[i] - xc;
    }
}

__global__ void k_norms2(const double* __restrict__ a, const double* __restrict__ b, int n, double* partial) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += a[i] * a[i];
        v[1] += b[i] * b[i];
    }
    block_partials<4>(v, partial, blockIdx.x);
}

__global__ void k_dot(const double* __restrict__ a, const double* __restrict__ b, int n, double* partial) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[0] += a[i] * b[i];
    block_partials<4>(v, partial, blockIdx.x);
}

// UPDATE_GPTV (Fasolt3.f90:254-284) on field-major arrays.
// ---------------------------------------------------------------------------
// Newton iteration on the device (NrDev): the residual-rule solve's stopping
// decisions and the damped update, so that a whole iteration can run as one
// graph.  Every kernel below that takes the record returns at once when its
// condition does not hold (a uniform read), which is how one captured
// sequence serves solves of any length.
// ---------------------------------------------------------------------------
// Batches before the next residual check, from the residuals at the last two
// checks and the sweeps between them: at most 80 % of the remainder that the
// observed geometric rate predicts (so the check after them is passed by at
// most one batch), at most kMaxChecks; one when the residual did not fall.
__device__ __forceinline__ int check_batches(double rel, double prev_rel, int sweeps, int prev_sweeps, double tol,
                                             int batch, int max_batches) {
#pragma clang fp contract(off)
    if (!(prev_rel > 0.0 && rel < prev_rel)) return 1;
    const double rate = log(rel / prev_rel) / (double)(sweeps - prev_sweeps);
    const double sw = 0.8 * (log(tol / rel) / rate);
    if (!(sw < (double)(max_batches * batch))) return max_batches;
    return (int)sw / batch;
}

__global__ void k_set_omega(NrDev* nd, double omega) {
    if (threadIdx.x == 0) nd->omega = omega;
}

// Inputs of a Newton iteration run as a graph (written before its launch).
__global__ void k_nr_set(NrDev* nd, int inc, double dt, double omega, uint64_t epoch) {
    if (threadIdx.x == 0) {
        nd->inc = inc;
        nd->dt = dt;
        nd->omega = omega;
        nd->epoch = epoch;
    }
}

__global__ void k_epoch_inc(NrDev* nd) {
    if (threadIdx.x == 0) nd->epoch += 1;
}

__global__ void k_flags_to_nr(const int* flags, NrDev* nd) {
    if (threadIdx.x == 0) {
        nd->flags[0] = flags[0];
        nd->flags[1] = flags[1];
    }
}

// Start of a residual-rule solve (Stif_sol.f90:179 starts from zero; the
// caller clears ddx).  flags non-null: the ELVALS code in flags[1] decides
// whether the iteration goes on (codes above 500 other than 987 and 910 end
// NRITER before the solve).
__global__ void k_solve_init(NrDev* nd, const int* flags, int maxit) {
    if (threadIdx.x != 0) return;
    int fatal = 0;
    if (flags) {
        const int e = flags[1];
        fatal = e < 0 || (e > 500 && e != 987 && e != 910);
    }
    nd->fatal = fatal;
    nd->updated = fatal;
    nd->stop = fatal || maxit <= 0;
    nd->sweeps = 0;
    nd->left = 1;
    nd->check_now = 0;
    nd->cycle = 1;
    nd->prev_sweeps = 0;
    nd->rel = 1.0;
    nd->prev_rel = -1.0;
    nd->final_rate = 0.0;
}

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Mixed precision, first batch of a stretch: the FP32 correction equation
// A d = r from d = 0, r the FP64 residual of the previous check.
__global__ void k_cycle_start(const NrDev* nd, const double* __restrict__ r, float* __restrict__ r32,
                              float* __restrict__ d32, int n) {
    if (nd->stop || !nd->cycle) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        r32[i] = (float)r[i];
        d32[i] = 0.0f;
    }
}

__global__ void k_solve_after(NrDev* nd, int batch) {
    if (threadIdx.x != 0 || nd->stop) return;
    nd->sweeps += batch;
    nd->left -= 1;
    nd->cycle = 0;
    nd->check_now = nd->left <= 0;
}

// Mixed precision, end of a stretch: ddx += d.
__global__ void k_cycle_end(const NrDev* nd, double* __restrict__ y, const float* __restrict__ d, int n) {
    if (!nd->check_now) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + (double)d[i];
}

// Residual check after a stretch: relative Jacobi-scaled residual from the
// residual kernel's partials; stop below tol or at maxit sweeps, otherwise
// the next stretch's length.  The solve's rate (log residual reduction per
// sweep) is kept from the latest stretch of at least four batches and, once
// the solve meets its tolerance, replaced by its whole-solve average
// (log of the final relative residual over the sweeps; the solve starts at
// relative residual 1).  The average is what the relaxation factor's search
// compares: above the optimal factor the iteration's dominant eigenvalues are
// complex and the residual oscillates, so rates over single stretches scatter
// more than the sweep counts they stand for.
__global__ __launch_bounds__(kBlock) void k_solve_check(const double* __restrict__ partial, int nblocks, NrDev* nd,
                                                        double tol, int maxit, int batch, int max_batches) {
#pragma clang fp contract(off)
    if (!nd->check_now) return;
    double r[4];
    reduce_partials(partial, nblocks, r);
    if (threadIdx.x != 0) return;
    // a right-hand side with NaN or Inf entries makes the measure NaN, not 0
    const double rel = !isfinite(r[1]) ? r[1] - r[1] : (r[1] > 0.0 ? sqrt(r[0] / r[1]) : 0.0);
    const int sweeps = nd->sweeps;
    const double prev_rel = nd->prev_rel;
    const int prev_sweeps = nd->prev_sweeps;
    if (prev_rel > 0.0 && rel > 0.0 && rel < prev_rel && sweeps - prev_sweeps >= 4 * batch)
        nd->final_rate = log(rel / prev_rel) / (double)(sweeps - prev_sweeps);
    nd->rel = rel;
    nd->check_now = 0;
    if (rel < tol) {
        if (rel > 0.0 && sweeps > 0) nd->final_rate = log(rel) / (double)sweeps;
        nd->stop = 1;
        return;
    }
    if (!isfinite(rel)) {
        // NaN or Inf in the system or diverging sweeps: the solve ends at once
        // and the iteration's gated Newton update is skipped, so that the state
        // stays as it was (the host reports 904 and ends NRITER)
        nd->stop = 1;
        nd->updated = 1;
        return;
    }
    int b = check_batches(rel, prev_rel, sweeps, prev_sweeps, tol, batch, max_batches);
    nd->prev_rel = rel;
    nd->prev_sweeps = sweeps;
    if (sweeps >= maxit) {
        nd->stop = 1;
        return;
    }
    const int cap = (maxit - sweeps) / batch;
    b = b < cap ? b : cap;
    nd->left = b > 1 ? b : 1;
    nd->cycle = 1;
}

// The Newton update of Stif_sol.f90:66-74 on the device.  gate non-null: run
// only once the solve has stopped and before the update has run.
__device__ __forceinline__ bool update_gate(const NrDev* g) { return !g || (g->stop && !g->updated); }

__global__ void k_norms2_g(const double* __restrict__ a, const double* __restrict__ b, int n, double* partial,
                           const NrDev* gate) {
    if (!update_gate(gate)) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += a[i] * a[i];
        v[1] += b[i] * b[i];
    }
    block_partials<4>(v, partial, blockIdx.x);
}

// sdx = |ddx| / |dx| (toler when dx = 0) and the damping alp2 = 1 / (1 + alpha sdx).
__global__ __launch_bounds__(kBlock) void k_nr_alpha(const double* __restrict__ partial, int nblocks, NrDev* nd,
                                                     double toler, double alpha, const NrDev* gate) {
    // the host's arithmetic (no fused multiply-add), as the reference evaluates it
#pragma clang fp contract(off)
    if (!update_gate(gate)) return;
    double r[4];
    reduce_partials(partial, nblocks, r);
    if (threadIdx.x != 0) return;
    for (int q = 0; q < 4; ++q) nd->nrm[q] = r[q];
    const double errdx = sqrt(r[0]), erddx = sqrt(r[1]);
    const double sdx = errdx == 0.0 ? toler : erddx / errdx;
    nd->sdx = sdx;
    nd->alp2 = 1.0 / (1.0 + alpha * sdx);
}

__global__ void k_nr_update_g(double* __restrict__ x, double* __restrict__ dx, const double* __restrict__ ddx,
                              const NrDev* nd, int n, const NrDev* gate) {
    if (!update_gate(gate)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double xc = nd->alp2 * ddx[i];
        x[i] = x[i] - xc;
        dx[i] = dx[i] - xc;
    }
}

__global__ void k_nr_mark(NrDev* nd) {
    if (threadIdx.x == 0 && nd->stop && !nd->updated) nd->updated = 1;
}

__global__ void k_commit(int n, int nsv, double* eps, const double* dep, double* rt, const double* dr, double* ep,
                         const double* de, double* st, const double* ds, double* sv, const double* dsv) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = (size_t)n;
    eps[g] = eps[g] + dep[g];
    for (int j = 0; j < 3; ++j) rt[j * N + g] = rt[j * N + g] + dr[j * N + g];
    for (int j = 0; j < 6; ++j) {
        ep[j * N + g] = ep[j * N + g] + de[j * N + g];
        st[j * N + g] = st[j * N + g] + ds[j * N + g];
    }
    for (int j = 0; j < nsv; ++j) sv[j * N + g] = sv[j * N + g] + dsv[j * N + g];
}

// Cumulative slip of Stif_sol.f90:80-84.
__global__ void k_cumslip(int n, int nss, double* cumslip, const double* srate, double dt) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t N = (size_t)n;
    for (int k = 0; k < nss; ++k) cumslip[k * N + g] = cumslip[k * N + g] + (srate[k * N + g] * dt);
}

// DoReorient (Fasolt3.f90:79-178) with the material-1 twin tables.
struct TwinTables {
    double ssang[60];
    double ab[60][3], an[60][3];
};

__global__ void k_reorient(int n, double* sv, int* twinsys, int* timestwinned, const TwinTables* __restrict__ T,
                           int* count) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || twinsys[j] == 0) return;
    const size_t N = (size_t)n;
    const int i = twinsys[j] - 1;
    const double e1 = sv[j], e2 = sv[N + j], e3 = sv[2 * N + j];
    const double ang = T->ssang[i] * 3.14159 / 180.;
    const double c1 = cos(e1), c2 = cos(e2), c3 = cos(e3), ss1 = sin(e1), s2 = sin(e2), s3 = sin(e3);
    double rm[3][3];
    rm[0][0] = c1 * c3 - c2 * ss1 * s3;
    rm[0][1] = -c1 * s3 - c2 * c3 * ss1;
    rm[0][2] = ss1 * s2;
    rm[1][0] = c3 * ss1 + c1 * c2 * s3;
    rm[1][1] = c1 * c2 * c3 - ss1 * s3;
    rm[1][2] = -c1 * s2;
    rm[2][0] = s2 * s3;
    rm[2][1] = c3 * s2;
    rm[2][2] = c2;
    double nv[3], dv[3];
    for (int p = 0; p < 3; ++p) {
        double sn = 0.0, sd = 0.0;
        for (int q = 0; q < 3; ++q) {
            sn = sn + rm[p][q] * T->an[i][q];
            sd = sd + rm[p][q] * T->ab[i][q];
        }
        nv[p] = sn;
        dv[p] = sd;
    }
    const double u1 = nv[1] * dv[2] - nv[2] * dv[1];
    const double u2 = nv[2] * dv[0] - nv[0] * dv[2];
    const double u3 = nv[0] * dv[1] - nv[1] * dv[0];
    const double c = cos(ang), s = sin(ang);
    double erm[3][3];
    erm[0][0] = c + u1 * u1 * (1. - c);
    erm[0][1] = u1 * u2 * (1. - c) - u3 * s;
    erm[0][2] = u1 * u3 * (1. - c) + u2 * s;
    erm[1][0] = u1 * u2 * (1. - c) + u3 * s;
    erm[1][1] = c + u2 * u2 * (1. - c);
    erm[1][2] = u2 * u3 * (1. - c) - u1 * s;
    erm[2][0] = u1 * u3 * (1. - c) - u2 * s;
    erm[2][1] = u2 * u3 * (1. - c) + u1 * s;
    erm[2][2] = c + u3 * u3 * (1. - c);
    double frm[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
            double sum = 0.0;
            for (int k = 0; k < 3; ++k) sum = sum + erm[p][k] * rm[k][q];
            frm[p][q] = sum;
        }
    const double e1n = atan2(frm[0][2], -frm[1][2]);
    if (fabs(frm[2][2]) > 1.) frm[2][2] = frm[2][2] / fabs(frm[2][2]);
    const double e2n = acos(frm[2][2]);
    const double e3n = atan2(frm[2][0], frm[2][1]);
    sv[j] = e1n;
    sv[N + j] = e2n;
    sv[2 * N + j] = e3n;
    twinsys[j] = 0;
    timestwinned[j] = timestwinned[j] + 1;
    atomicAdd(count, 1);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline int nblk(size_t n, int b = kBlock) { return (int)((n + b - 1) / b); }
// Blocks per XCD of a launch over ng row groups, wpb groups (waves) per block.
static int group_per_xcd(int ng, int wpb = kWavesPerBlock) { return ((ng + wpb - 1) / wpb + kXcd - 1) / kXcd; }

template <typename T>
static hipError_t dalloc(Ctx* c, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e == hipSuccess) {
        c->allocs.push_back((void*)*p);
        e = hipMemsetAsync(*p, 0, n * sizeof(T), c->stream);
    }
    return e;
}

static double reduce_now(Ctx* c, int nblocks, double out[4]) {
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, c->stream, c->red.partial, nblocks, c->red.result);
    (void)hipMemcpyAsync(c->red.host, c->red.result, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    (void)hipStreamSynchronize(c->stream);
    for (int q = 0; q < 4; ++q) out[q] = c->red.host[q];
    return out[0];
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

// Gather assembly on meshes of 8-node bricks (CPFEM_ASSEMBLY=scatter keeps
// the colour-by-colour scatter): for every matrix block and every node, the
// list of its element contributions in the order of the element colours (ebc
// holds the elements colour by colour), padded per segment of 64 lanes.
static int build_gather(Ctx* c, int nel, int nnod, const std::vector<int>& en, const std::vector<int>& kind,
                        const std::vector<int>& scat, const std::vector<int>& scat_start, const std::vector<int>& ebc) {
    c->A.gather = false;
    const char* am = getenv("CPFEM_ASSEMBLY");
    if (am && strcmp(am, "scatter") == 0) return 0;
    if (nel == 0 || (size_t)nel * 576 > 0x7fffffffull) return 0;
    for (int e = 0; e < nel; ++e)
        if (kind[e] != kElem8) return 0;
    const int nseg = c->A.ngroups * c->A.P;
    // matrix blocks: block entry 0 at offset off = segment*576 + lane
    std::vector<int> cnt((size_t)nseg * 64, 0);
    for (int e = 0; e < nel; ++e)
        for (int pr = 0; pr < 64; ++pr) {
            const int off = scat[scat_start[e] + pr];
            cnt[(size_t)(off / 576) * 64 + off % 64]++;
        }
    std::vector<int> seg_cnt(nseg, 0), seg_start(nseg, 0);
    size_t total = 0;
    for (int sg = 0; sg < nseg; ++sg) {
        int m = 0;
        for (int l = 0; l < 64; ++l) m = std::max(m, cnt[(size_t)sg * 64 + l]);
        seg_cnt[sg] = m;
        if (total > 0x7fffffffull) return 0;
        seg_start[sg] = (int)total;
        total += (size_t)m * 64;
    }
    if (total > 0x7fffffffull) return 0;
    std::vector<int> glist(std::max<size_t>(total, 1), -1);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int t = 0; t < nel; ++t) {
        const int e = ebc[t];
        for (int pr = 0; pr < 64; ++pr) {
            const int off = scat[scat_start[e] + pr];
            const int sg = off / 576, l = off % 64;
            const int k = cnt[(size_t)sg * 64 + l]++;
            glist[(size_t)seg_start[sg] + (size_t)k * 64 + l] = e * 576 + (pr >> 3) * 72 + (pr & 7);
        }
    }
    // wave order of the block gather: group j of every colour in turn (the
    // colours' j-th groups cover neighbouring nodes, which share elements)
    std::vector<int> order;
    order.reserve(nseg);
    const int ncol = c->A.ncolor, P = c->A.P;
    int maxg = 0;
    for (int k = 0; k < ncol; ++k) maxg = std::max(maxg, c->A.gstart[k + 1] - c->A.gstart[k]);
    for (int j = 0; j < maxg; ++j)
        for (int k = 0; k < ncol; ++k)
            if (j < c->A.gstart[k + 1] - c->A.gstart[k])
                for (int p = 0; p < P; ++p) order.push_back((c->A.gstart[k] + j) * P + p);
    if ((int)order.size() != nseg) return CPFEM_E_ARG;
    // node forces: chunks of 64 nodes
    const int nch = (nnod + 63) / 64;
    std::vector<int> fcnt(nnod, 0), fch_cnt(nch, 0), fch_start(nch, 0);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < 8; ++a) fcnt[en[kMaxNodes * e + a]]++;
    size_t ftotal = 0;
    for (int ch = 0; ch < nch; ++ch) {
        int m = 0;
        for (int nd = ch * 64; nd < std::min(nnod, ch * 64 + 64); ++nd) m = std::max(m, fcnt[nd]);
        fch_cnt[ch] = m;
        fch_start[ch] = (int)ftotal;
        ftotal += (size_t)m * 64;
    }
    if (ftotal > 0x7fffffffull) return 0;
    std::vector<int> flist(std::max<size_t>(ftotal, 1), -1);
    std::fill(fcnt.begin(), fcnt.end(), 0);
    for (int t = 0; t < nel; ++t) {
        const int e = ebc[t];
        for (int a = 0; a < 8; ++a) {
            const int nd = en[kMaxNodes * e + a];
            const int k = fcnt[nd]++;
            flist[(size_t)fch_start[nd / 64] + (size_t)k * 64 + nd % 64] = (e * 8 + a) * 3;
        }
    }
    const size_t vbytes = (c->precision == 1) ? sizeof(float) : sizeof(double);
    char* eb = nullptr;
    if (dalloc(c, &eb, (size_t)nel * 576 * vbytes) || dalloc(c, &c->A.efc, (size_t)nel * 24) ||
        dalloc(c, &c->A.g_order, nseg) || dalloc(c, &c->A.g_start, nseg) || dalloc(c, &c->A.g_cnt, nseg) ||
        dalloc(c, &c->A.g_list, glist.size()) || dalloc(c, &c->A.f_start, nch) || dalloc(c, &c->A.f_cnt, nch) ||
        dalloc(c, &c->A.f_list, flist.size()))
        return CPFEM_E_HIP;
    c->A.eblk = eb;
    hipStream_t s = c->stream;
    CPFEM_HIP(hipMemcpyAsync(c->A.g_order, order.data(), nseg * sizeof(int), hipMemcpyHostToDevice, s));
    CPFEM_HIP(hipMemcpyAsync(c->A.g_start, seg_start.data(), nseg * sizeof(int), hipMemcpyHostToDevice, s));
    CPFEM_HIP(hipMemcpyAsync(c->A.g_cnt, seg_cnt.data(), nseg * sizeof(int), hipMemcpyHostToDevice, s));
    CPFEM_HIP(hipMemcpyAsync(c->A.g_list, glist.data(), glist.size() * sizeof(int), hipMemcpyHostToDevice, s));
    CPFEM_HIP(hipMemcpyAsync(c->A.f_start, fch_start.data(), nch * sizeof(int), hipMemcpyHostToDevice, s));
    CPFEM_HIP(hipMemcpyAsync(c->A.f_cnt, fch_cnt.data(), nch * sizeof(int), hipMemcpyHostToDevice, s));
    CPFEM_HIP(hipMemcpyAsync(c->A.f_list, flist.data(), flist.size() * sizeof(int), hipMemcpyHostToDevice, s));
    CPFEM_HIP(hipStreamSynchronize(s));
    c->A.gather = true;
    return 0;
}

// Sparsity (the node blocks SETPTA's row lists cover, Stif_sol.f90:98-149),
// node and element colourings, the row-group layout and the scatter map.
// en: [nel][kMaxNodes] 0-based nodes, nde: nodes per element, kind: ElemKind.
static int build_structure(Ctx* c, const cpfem_mesh* m, const std::vector<int>& en, const std::vector<int>& nde,
                           const std::vector<int>& kind) {
    const int nel = m->nlmnt, nnod = m->nnod;
    // node neighbour lists (nodes sharing an element), self first
    std::vector<std::vector<int>> nb(nnod);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < nde[e]; ++a)
            for (int b = 0; b < nde[e]; ++b) nb[en[kMaxNodes * e + a]].push_back(en[kMaxNodes * e + b]);
    int S = 0;
    size_t nnzb = 0;
    for (int i = 0; i < nnod; ++i) {
        auto& v = nb[i];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        v.erase(std::remove(v.begin(), v.end(), i), v.end());
        v.insert(v.begin(), i);
        S = std::max<int>(S, (int)v.size());
        nnzb += v.size();
    }
    // greedy node colouring (nodes sharing an element differ)
    std::vector<int> col(nnod, -1);
    int ncol = 0;
    std::vector<char> used;
    for (int i = 0; i < nnod; ++i) {
        used.assign(64, 0);
        for (int j : nb[i])
            if (j != i && col[j] >= 0 && col[j] < 64) used[col[j]] = 1;
        int k = 0;
        while (k < 64 && used[k]) ++k;
        if (k == 64) return CPFEM_E_ARG;
        col[i] = k;
        ncol = std::max(ncol, k + 1);
    }
    std::vector<int> row_node(nnod), node_row(nnod);
    c->A.cstart.assign(ncol + 1, 0);
    for (int i = 0; i < nnod; ++i) c->A.cstart[col[i] + 1]++;
    for (int k = 0; k < ncol; ++k) c->A.cstart[k + 1] += c->A.cstart[k];
    {
        std::vector<int> pos(c->A.cstart.begin(), c->A.cstart.end() - 1);
        for (int i = 0; i < nnod; ++i) {
            const int r = pos[col[i]]++;
            row_node[r] = i;
            node_row[i] = r;
        }
    }
    const int R = nnod;
    // row groups: each colour's rows cut into groups of kGroupRows
    const int P = (S + kRowLanes - 1) / kRowLanes;
    c->A.gstart.assign(ncol + 1, 0);
    for (int k = 0; k < ncol; ++k)
        c->A.gstart[k + 1] = c->A.gstart[k] + (c->A.cstart[k + 1] - c->A.cstart[k] + kGroupRows - 1) / kGroupRows;
    const int ngroups = c->A.gstart[ncol];
    if ((size_t)ngroups * P * 576 > 0x7fffffffull) return CPFEM_E_ARG;
    std::vector<int> gpos(R);   // group position q of each row
    for (int k = 0; k < ncol; ++k)
        for (int r = c->A.cstart[k]; r < c->A.cstart[k + 1]; ++r)
            gpos[r] = c->A.gstart[k] * kGroupRows + (r - c->A.cstart[k]);
    std::vector<int> cols((size_t)ngroups * P * 64, -1), diag_off(nnod);
    for (int r = 0; r < R; ++r) {
        const int nd = row_node[r];
        for (int s = 0; s < P * kRowLanes; ++s)
            cols[col_index(gpos[r], P, s)] = s < (int)nb[nd].size() ? nb[nd][s] : nd;
        diag_off[nd] = (int)val_index(gpos[r], P, 0, 0);
    }
    // element scatter map: entry 0 of the block of each node pair (a, b) at
    // scat_start[e] + a*nd + b
    std::vector<int> scat_start(nel);
    size_t nscat = 0;
    for (int e = 0; e < nel; ++e) {
        scat_start[e] = (int)nscat;
        nscat += (size_t)nde[e] * nde[e];
    }
    if (nscat > 0x7fffffffull) return CPFEM_E_ARG;
    std::vector<int> scat(std::max<size_t>(nscat, 1));
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < nde[e]; ++a) {
            const int ni = en[kMaxNodes * e + a];
            const auto& v = nb[ni];
            for (int b = 0; b < nde[e]; ++b) {
                const int nj = en[kMaxNodes * e + b];
                int slot = 0;
                if (nj != ni) slot = (int)(std::lower_bound(v.begin() + 1, v.end(), nj) - v.begin());
                scat[scat_start[e] + a * nde[e] + b] = (int)val_index(gpos[node_row[ni]], P, slot, 0);
            }
        }
    // element colouring (elements sharing a node differ)
    std::vector<std::vector<int>> node_elems(nnod);
    for (int e = 0; e < nel; ++e)
        for (int a = 0; a < nde[e]; ++a) node_elems[en[kMaxNodes * e + a]].push_back(e);
    std::vector<int> ecol(nel, -1);
    int necol = 0;
    for (int e = 0; e < nel; ++e) {
        used.assign(256, 0);
        for (int a = 0; a < nde[e]; ++a)
            for (int f : node_elems[en[kMaxNodes * e + a]])
                if (ecol[f] >= 0) used[ecol[f]] = 1;
        int k = 0;
        while (k < 256 && used[k]) ++k;
        if (k == 256) return CPFEM_E_ARG;
        ecol[e] = k;
        necol = std::max(necol, k + 1);
    }
    // elements ordered by colour, then kind
    c->ekstart.assign(kKinds * necol + 1, 0);
    for (int e = 0; e < nel; ++e) c->ekstart[kKinds * ecol[e] + kind[e] + 1]++;
    for (int k = 0; k < kKinds * necol; ++k) c->ekstart[k + 1] += c->ekstart[k];
    std::vector<int> ebc(nel);
    {
        std::vector<int> pos(c->ekstart.begin(), c->ekstart.end() - 1);
        for (int e = 0; e < nel; ++e) ebc[pos[kKinds * ecol[e] + kind[e]]++] = e;
    }
    c->A.R = R;
    c->A.S = S;
    c->A.P = P;
    c->A.ncolor = ncol;
    c->A.ngroups = ngroups;
    // colour order of the sweeps: the colour numbers, except with 8 colours
    // the reflected Gray code with bit roles (1, 4, 2): 0 1 5 4 6 7 3 2.  On a
    // brick mesh the greedy colours are node parity classes, and a sweep that
    // visits them one axis flip at a time needs 20-30 % fewer sweeps than
    // 0..7 (64^3 plane strain, 10^3-16^3 plane strain and tension, tuned omega)
    c->corder.resize(ncol);
    for (int k = 0; k < ncol; ++k) c->corder[k] = k;
    if (ncol == 8) c->corder = {0, 1, 5, 4, 6, 7, 3, 2};
    c->A.nnzb = nnzb;
    const size_t vbytes = (c->precision == 1) ? sizeof(float) : sizeof(double);
    if (dalloc(c, &c->A.diag_off, nnod) || dalloc(c, &c->A.cols, cols.size()) ||
        dalloc(c, &c->A.scatter, scat.size()) || dalloc(c, &c->A.scat_start, nel) ||
        dalloc(c, &c->elem_by_color, nel))
        return CPFEM_E_HIP;
    char* vp = nullptr;
    if (dalloc(c, &vp, c->A.nvals() * vbytes)) return CPFEM_E_HIP;
    c->A.vals = vp;
    if (c->precision == 2 && dalloc(c, &c->A.vals32, c->A.nvals())) return CPFEM_E_HIP;
    CPFEM_HIP(hipMemcpyAsync(c->A.diag_off, diag_off.data(), nnod * sizeof(int), hipMemcpyHostToDevice, c->stream));
    CPFEM_HIP(hipMemcpyAsync(c->A.cols, cols.data(), cols.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    CPFEM_HIP(hipMemcpyAsync(c->A.scatter, scat.data(), scat.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    CPFEM_HIP(hipMemcpyAsync(c->A.scat_start, scat_start.data(), nel * sizeof(int), hipMemcpyHostToDevice,
                             c->stream));
    CPFEM_HIP(hipMemcpyAsync(c->elem_by_color, ebc.data(), nel * sizeof(int), hipMemcpyHostToDevice, c->stream));
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    return build_gather(c, nel, nnod, en, kind, scat, scat_start, ebc);
}


}  // namespace cpfem

using namespace cpfem;

extern "C" int cpfem_create(const cpfem_mesh* m, const cpfem_material* mats, int32_t nmat, const cpfem_control* ctl,
                            cpfem_ctx** out) {
    if (!m || !mats || nmat < 1 || !ctl || !out) return CPFEM_E_ARG;
    *out = nullptr;
    // one material set of 12 or 18 systems runs in the kernels of its own count,
    // anything else in the generic ones (record and count read per point)
    int nss = 0;
    for (int k = 0; k < nmat; ++k) {
        if (material_bytes(mats[k].nss) == 0) return CPFEM_E_NSS;
        nss = std::max(nss, (int)mats[k].nss);
    }
    int knss = mats[0].nss;
    if (nmat > 1 || (knss != 12 && knss != 18)) knss = kNGeneric;
    // sv(1:3) Euler angles, sv(4:3+nss) slip resistances; FS3PRE may declare
    // more state variables (fs3p-gps.f90:30-31), which are carried unchanged
    if (m->nsv < 3 + nss || m->nsv > 63) return CPFEM_E_ARG;
    if (ctl->precision < 0 || ctl->precision > 2) return CPFEM_E_ARG;
    if (ctl->precision != 0 && ctl->sor_kind != 1) return CPFEM_E_ARG;
    if (!(ctl->sor_omega < 2.0)) return CPFEM_E_ARG;   // SOR on an SPD system converges for 0 < omega < 2
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CPFEM_E_NOGPU;
    // element tables.  Types: 12 (8-node brick, 1 point, SHAPE_12), 13 (8-node
    // brick, 2x2x2 points), 22 (20-node brick, 2x2x2 points, SHAPE_22) and the
    // two-node 'tie' elements 1-6 without points (Elvals.f90:62-77).
    const int nel = m->nlmnt;
    std::vector<int> en((size_t)nel * kMaxNodes, 0), eg((size_t)nel * 8, -1), lt(nel), le(nel), np(nel), nde(nel),
        kind(nel);
    std::vector<int> gpe(m->ngps, -1), gpl(m->ngps, -1);
    int npts8 = 0, npts20 = 0, npts10 = 0, npts15 = 0;
    for (int e = 0; e < nel; ++e) {
        const int t = m->lt[e], nd = m->nd[e], ng = m->np[e];
        if (t == 12 && nd == 8 && ng == 1) {
            kind[e] = kElem8;
        } else if (t == 13 && nd == 8 && ng == 8) {
            kind[e] = kElem8;
        } else if (t == 22 && nd == 20 && ng == 8) {
            kind[e] = kElem20;
        } else if (t == 21 && nd == 10 && ng == 4) {
            kind[e] = kElem10;
        } else if (t == 23 && nd == 15 && ng == 6) {
            kind[e] = kElem15;
        } else if (t >= 1 && t <= 6 && nd == 2 && ng == 0) {
            kind[e] = kTie;
        } else {
            return CPFEM_E_ARG;
        }
        if (kind[e] != kTie && (m->le[e] == 0 || abs(m->le[e]) > nmat)) return CPFEM_E_ARG;
        lt[e] = t;
        le[e] = m->le[e];
        np[e] = ng;
        nde[e] = nd;
        for (int a = 0; a < nd; ++a) {
            const int node = m->ln[20 * e + a] - 1;
            if (node < 0 || node >= m->nnod) return CPFEM_E_ARG;
            en[kMaxNodes * e + a] = node;
        }
        for (int g = 0; g < ng; ++g) {
            const int gp = m->lg[8 * e + g] - 1;
            if (gp < 0 || gp >= m->ngps || gpe[gp] >= 0) return CPFEM_E_ARG;
            eg[8 * e + g] = gp;
            gpe[gp] = e;
            gpl[gp] = g;
        }
        if (kind[e] == kElem8) npts8 += ng;
        if (kind[e] == kElem20) npts20 += ng;
        if (kind[e] == kElem10) npts10 += ng;
        if (kind[e] == kElem15) npts15 += ng;
    }
    for (int g = 0; g < m->ngps; ++g)
        if (gpe[g] < 0) return CPFEM_E_ARG;
    cpfem_ctx* c = new cpfem_ctx();
    c->nlmnt = nel;
    c->nnod = m->nnod;
    c->ngps = m->ngps;
    c->nsv = m->nsv;
    c->nss = nss;
    c->knss = knss;
    c->ninc1 = m->ninc1;
    c->nmat = nmat;
    c->ctl = *ctl;
    c->mat0 = mats[0];
    c->precision = ctl->precision;
    c->omega_auto = ctl->sor_omega <= 0.0;
    if (c->omega_auto) c->ctl.sor_omega = 1.0;   // until the first solve tunes it
    c->tiem.assign(m->tiem, m->tiem + m->ninc1);
    c->npts8 = npts8;
    c->npts20 = npts20;
    c->npts10 = npts10;
    c->npts15 = npts15;
    memset(&c->stats, 0, sizeof(c->stats));
    int rc = 0;
#define TRY(x)                      \
    do {                            \
        if ((x) != hipSuccess) {    \
            rc = CPFEM_E_HIP;       \
            goto fail;              \
        }                           \
    } while (0)
    {
        TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        {
            const char* g = getenv("CPFEM_SOR_GRAPHS");
            c->use_graphs = !(g && g[0] == '0');
        }
        for (int i = 0; i < Ctx::kEvents; ++i) TRY(hipEventCreate(&c->ev[i]));
        const size_t G = c->ngps, D = 3 * (size_t)c->nnod, nsv = c->nsv;
        TRY(dalloc(c, &c->elem_nodes, en.size()));
        TRY(dalloc(c, &c->elem_gps, eg.size()));
        TRY(dalloc(c, &c->elem_lt, nel));
        TRY(dalloc(c, &c->elem_le, nel));
        TRY(dalloc(c, &c->elem_np, nel));
        TRY(dalloc(c, &c->elem_nd, nel));
        TRY(dalloc(c, &c->gp_elem, G));
        TRY(dalloc(c, &c->gp_local, G));
        TRY(dalloc(c, &c->ib, c->nnod));
        TRY(dalloc(c, &c->x_spec, (size_t)m->ninc1 * 99 * 15));
        TRY(dalloc(c, &c->f_spec, (size_t)m->ninc1 * 99 * 3));
        TRY(hipMemcpyAsync(c->elem_nodes, en.data(), en.size() * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->elem_gps, eg.data(), eg.size() * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->elem_lt, lt.data(), nel * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->elem_le, le.data(), nel * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->elem_np, np.data(), nel * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->elem_nd, nde.data(), nel * 4, hipMemcpyHostToDevice, c->stream));
        if (npts20 > 0) TRY(upload_brick20_tables());
        if (npts10 + npts15 > 0) TRY(upload_tet_wedge_tables());
        TRY(hipMemcpyAsync(c->gp_elem, gpe.data(), G * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->gp_local, gpl.data(), G * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->ib, m->ib, c->nnod * 4, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->x_spec, m->x_spec, (size_t)m->ninc1 * 99 * 15 * 8, hipMemcpyHostToDevice, c->stream));
        TRY(hipMemcpyAsync(c->f_spec, m->f_spec, (size_t)m->ninc1 * 99 * 3 * 8, hipMemcpyHostToDevice, c->stream));
        TRY(dalloc(c, &c->x, D));
        TRY(dalloc(c, &c->dx, D));
        TRY(dalloc(c, &c->fc, D));
        TRY(dalloc(c, &c->rhs, D));
        TRY(dalloc(c, &c->ddx, D));
        TRY(dalloc(c, &c->xold, D));
        if (c->precision != 0) {
            TRY(dalloc(c, &c->rhs32, D));
            TRY(dalloc(c, &c->del32, D));
        }
        TRY(dalloc(c, &c->eps, G));
        TRY(dalloc(c, &c->dep, G));
        TRY(dalloc(c, &c->st, 6 * G));
        TRY(dalloc(c, &c->ep, 6 * G));
        TRY(dalloc(c, &c->ds, 6 * G));
        TRY(dalloc(c, &c->de, 6 * G));
        TRY(dalloc(c, &c->t, 6 * G));
        TRY(dalloc(c, &c->rt, 3 * G));
        TRY(dalloc(c, &c->dr, 3 * G));
        TRY(dalloc(c, &c->sv, nsv * G));
        TRY(dalloc(c, &c->dsv, nsv * G));
        TRY(dalloc(c, &c->srate, (size_t)nss * G));
        TRY(dalloc(c, &c->cumslip, (size_t)nss * G));
        TRY(dalloc(c, &c->chi, 36 * G));
        TRY(dalloc(c, &c->erepmd, G));
        TRY(dalloc(c, &c->twinsys, G));
        TRY(dalloc(c, &c->timestwinned, G));
        TRY(dalloc(c, &c->gperr, G));
        TRY(dalloc(c, &c->flops, G));
        TRY(dalloc(c, &c->transc, G));
        TRY(dalloc(c, &c->flags, 8));
        // reduction scratch: enough partial slots for the largest launch
        // (row-group launches use at most nnod / (4 * kGroupRows) + 8 blocks per colour)
        // (the volume averages use two point-sized launches)
        const int nbr = std::max(nblk(std::max<size_t>(std::max<size_t>(2 * G, D), 8 * (size_t)c->nnod)) + 64,
                                 (int)((size_t)c->nnod / kGroupRows) + 16 * 64 + 64);
        c->red.nblocks = nbr;
        TRY(dalloc(c, &c->red.partial, (size_t)nbr * 4));
        TRY(dalloc(c, &c->red.result, 8));
        TRY(hipHostMalloc((void**)&c->red.host, 8 * sizeof(double), 0));
        TRY(dalloc(c, &c->nrdev, 1));
        TRY(hipHostMalloc((void**)&c->nrhost, sizeof(NrDev), 0));
        // material
        {
            // one device record per material, indexed by |le| - 1
            const size_t mb = material_bytes(knss);
            std::vector<char> hm(mb * nmat);
            for (int k = 0; k < nmat; ++k) build_material(knss, hm.data() + k * mb, mats[k], ctl->toler);
            char* dm = nullptr;
            TRY(dalloc(c, &dm, hm.size()));
            c->dmat = dm;
            TRY(hipMemcpyAsync(dm, hm.data(), hm.size(), hipMemcpyHostToDevice, c->stream));
            TRY(hipStreamSynchronize(c->stream));
        }
        TRY(alloc_schedule(c->sched, c->ngps));
        rc = build_structure(c, m, en, nde, kind);
        if (rc) goto fail;
    }
    *out = c;
    return 0;
fail:
    cpfem_destroy(c);
    return rc ? rc : CPFEM_E_HIP;
#undef TRY
}

static void drop_graphs(Ctx* c);
static void drop_solve_graphs(Ctx* c);

extern "C" int cpfem_destroy(cpfem_ctx* c) {
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    calib_free(c);
    free_schedule(c->sched);
    drop_graphs(c);
    drop_solve_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->red.host) (void)hipHostFree(c->red.host);
    if (c->nrhost) (void)hipHostFree(c->nrhost);
    for (int i = 0; i < Ctx::kEvents; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

// field-major <-> reference per-point layout
static void to_field(std::vector<double>& out, const double* in, size_t n, int k, int ld) {
    out.resize(n * k);
    for (size_t g = 0; g < n; ++g)
        for (int i = 0; i < k; ++i) out[(size_t)i * n + g] = in[(size_t)ld * g + i];
}
static void from_field(double* out, const std::vector<double>& in, size_t n, int k, int ld) {
    for (size_t g = 0; g < n; ++g)
        for (int i = 0; i < k; ++i) out[(size_t)ld * g + i] = in[(size_t)i * n + g];
}

extern "C" int cpfem_set_state(cpfem_ctx* c, const double* x, const double* dx, const double* eps,
                               const double* dep, const double* sv, const double* dsv, int32_t ldsv, const double* ep,
                               const double* de, const double* st, const double* ds, const double* rt,
                               const double* dr, const int32_t* twinsys, const int32_t* timestwinned) {
    if (!c || ldsv < c->nsv) return CPFEM_E_ARG;
    const size_t G = c->ngps, D = 3 * (size_t)c->nnod;
    hipStream_t s = c->stream;
    std::vector<double> tmp;
    if (x) CPFEM_HIP(hipMemcpy(c->x, x, D * 8, hipMemcpyHostToDevice));
    if (dx) CPFEM_HIP(hipMemcpy(c->dx, dx, D * 8, hipMemcpyHostToDevice));
    if (eps) CPFEM_HIP(hipMemcpy(c->eps, eps, G * 8, hipMemcpyHostToDevice));
    if (dep) CPFEM_HIP(hipMemcpy(c->dep, dep, G * 8, hipMemcpyHostToDevice));
    struct F { const double* h; double* d; int k, ld; } fs[] = {
        {sv, c->sv, c->nsv, ldsv}, {dsv, c->dsv, c->nsv, ldsv}, {ep, c->ep, 6, 6}, {de, c->de, 6, 6},
        {st, c->st, 6, 6}, {ds, c->ds, 6, 6}, {rt, c->rt, 3, 3}, {dr, c->dr, 3, 3}};
    for (auto& f : fs) {
        if (!f.h) continue;
        to_field(tmp, f.h, G, f.k, f.ld);
        CPFEM_HIP(hipMemcpy(f.d, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice));
    }
    if (twinsys) CPFEM_HIP(hipMemcpy(c->twinsys, twinsys, G * 4, hipMemcpyHostToDevice));
    if (timestwinned) CPFEM_HIP(hipMemcpy(c->timestwinned, timestwinned, G * 4, hipMemcpyHostToDevice));
    (void)s;
    return 0;
}

extern "C" int cpfem_get_state(cpfem_ctx* c, double* x, double* dx, double* fc, double* eps, double* dep,
                               double* sv, double* dsv, int32_t ldsv, double* ep, double* de, double* st, double* ds,
                               double* rt, double* dr, double* srate, double* cumslip, int32_t* twinsys,
                               int32_t* timestwinned) {
    if (!c || ldsv < c->nsv) return CPFEM_E_ARG;
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    const size_t G = c->ngps, D = 3 * (size_t)c->nnod;
    std::vector<double> tmp;
    if (x) CPFEM_HIP(hipMemcpy(x, c->x, D * 8, hipMemcpyDeviceToHost));
    if (dx) CPFEM_HIP(hipMemcpy(dx, c->dx, D * 8, hipMemcpyDeviceToHost));
    if (fc) CPFEM_HIP(hipMemcpy(fc, c->fc, D * 8, hipMemcpyDeviceToHost));
    if (eps) CPFEM_HIP(hipMemcpy(eps, c->eps, G * 8, hipMemcpyDeviceToHost));
    if (dep) CPFEM_HIP(hipMemcpy(dep, c->dep, G * 8, hipMemcpyDeviceToHost));
    struct F { double* h; const double* d; int k, ld; } fs[] = {
        {sv, c->sv, c->nsv, ldsv}, {dsv, c->dsv, c->nsv, ldsv}, {ep, c->ep, 6, 6}, {de, c->de, 6, 6},
        {st, c->st, 6, 6}, {ds, c->ds, 6, 6}, {rt, c->rt, 3, 3}, {dr, c->dr, 3, 3},
        {srate, c->srate, c->nss, CPFEM_MAXSS}, {cumslip, c->cumslip, c->nss, CPFEM_MAXSS}};
    for (auto& f : fs) {
        if (!f.h) continue;
        tmp.resize(G * f.k);
        CPFEM_HIP(hipMemcpy(tmp.data(), f.d, tmp.size() * 8, hipMemcpyDeviceToHost));
        from_field(f.h, tmp, G, f.k, f.ld);
    }
    if (twinsys) CPFEM_HIP(hipMemcpy(twinsys, c->twinsys, G * 4, hipMemcpyDeviceToHost));
    if (timestwinned) CPFEM_HIP(hipMemcpy(timestwinned, c->timestwinned, G * 4, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// REAL(4) binding of the FAS_COM / FAS_CPL arrays (fas-com-mod.f90:13-80):
// the reference's single-precision module arrays pass as they are; the
// library widens them to its FP64 device state (exactly) and narrows results
// to single precision with round-to-nearest, as the reference's own REAL(4)
// stores do.
// ---------------------------------------------------------------------------
namespace {
// FP64 copy of a single-precision host array (empty for a null pointer)
struct Wide {
    std::vector<double> v;
    Wide(const float* p, size_t n) {
        if (p) v.assign(p, p + n);
    }
    double* get() { return v.empty() ? nullptr : v.data(); }
    void narrow(float* p) const {
        if (p)
            for (size_t i = 0; i < v.size(); ++i) p[i] = (float)v[i];
    }
};
}  // namespace

extern "C" int cpfem_create_r4(const cpfem_mesh_r4* m, const cpfem_materials_r4* mt, const cpfem_control* ctl,
                               cpfem_ctx** out) {
    if (!m || !mt || !ctl || !out || mt->nmatl < 1 || mt->nmatl > CPFEM_MAXMDS || !mt->nss || !mt->e ||
        !mt->ssang || !mt->ab || !mt->an || !mt->ampar || !mt->smatrx || m->ninc1 < 1 || m->nxs < 0 ||
        m->nxs > 99 || m->nfs < 0 || m->nfs > 99)
        return CPFEM_E_ARG;
    const size_t ninc1 = (size_t)m->ninc1;
    Wide tiem(m->tiem, ninc1), xs(m->x_spec, ninc1 * 99 * 15), fs(m->f_spec, ninc1 * 99 * 3);
    cpfem_mesh md;
    md.nlmnt = m->nlmnt;
    md.nnod = m->nnod;
    md.ngps = m->ngps;
    md.nsv = m->nsv;
    md.le = m->le;
    md.lt = m->lt;
    md.nd = m->nd;
    md.np = m->np;
    md.ln = m->ln;
    md.lg = m->lg;
    md.ib = m->ib;
    md.ninc1 = m->ninc1;
    md.nxs = m->nxs;
    md.nfs = m->nfs;
    md.tiem = tiem.get();
    md.x_spec = xs.get();
    md.f_spec = fs.get();
    // material set k at its slice of the module arrays: e(9,k), ssang(60,k),
    // ab/an(3,60,k), ampar(10,60,k), smatrx(60,60,k); twinlim and twinprob are
    // module scalars shared by every set (fas-com-mod.f90:59-60)
    std::vector<cpfem_material> mats((size_t)mt->nmatl);
    for (int k = 0; k < mt->nmatl; ++k) {
        cpfem_material& d = mats[(size_t)k];
        memset(&d, 0, sizeof(d));
        d.nss = mt->nss[k];
        if (d.nss < 1 || d.nss > CPFEM_MAXSS) return CPFEM_E_NSS;
        d.twinlim = mt->twinlim;
        d.twinprob = mt->twinprob;
        for (int i = 0; i < 9; ++i) d.e[i] = mt->e[9 * k + i];
        const size_t b = (size_t)k * CPFEM_MAXSS;
        for (int i = 0; i < CPFEM_MAXSS; ++i) {
            d.ssang[i] = mt->ssang[b + i];
            for (int q = 0; q < 3; ++q) {
                d.ab[i][q] = mt->ab[3 * (b + i) + q];
                d.an[i][q] = mt->an[3 * (b + i) + q];
            }
            for (int q = 0; q < 10; ++q) d.ampar[i][q] = mt->ampar[10 * (b + i) + q];
            for (int j = 0; j < CPFEM_MAXSS; ++j)
                d.smatrx[i][j] = mt->smatrx[(size_t)k * CPFEM_MAXSS * CPFEM_MAXSS + (size_t)i * CPFEM_MAXSS + j];
        }
    }
    return cpfem_create(&md, mats.data(), mt->nmatl, ctl, out);
}

extern "C" int cpfem_set_state_r4(cpfem_ctx* c, const float* x, const float* dx, const float* eps,
                                  const float* dep, const float* sv, const float* dsv, int32_t ldsv, const float* ep,
                                  const float* de, const float* st, const float* ds, const float* rt, const float* dr,
                                  const int32_t* twinsys, const int32_t* timestwinned) {
    if (!c || ldsv < c->nsv) return CPFEM_E_ARG;
    const size_t G = c->ngps, D = 3 * (size_t)c->nnod;
    Wide wx(x, D), wdx(dx, D), weps(eps, G), wdep(dep, G), wsv(sv, G * ldsv), wdsv(dsv, G * ldsv);
    Wide wep(ep, 6 * G), wde(de, 6 * G), wst(st, 6 * G), wds(ds, 6 * G), wrt(rt, 3 * G), wdr(dr, 3 * G);
    return cpfem_set_state(c, wx.get(), wdx.get(), weps.get(), wdep.get(), wsv.get(), wdsv.get(), ldsv, wep.get(),
                           wde.get(), wst.get(), wds.get(), wrt.get(), wdr.get(), twinsys, timestwinned);
}

extern "C" int cpfem_get_state_r4(cpfem_ctx* c, float* x, float* dx, float* fc, float* eps, float* dep, float* sv,
                                  float* dsv, int32_t ldsv, float* ep, float* de, float* st, float* ds, float* rt,
                                  float* dr, float* srate, float* cumslip, int32_t* twinsys, int32_t* timestwinned) {
    if (!c || ldsv < c->nsv) return CPFEM_E_ARG;
    const size_t G = c->ngps, D = 3 * (size_t)c->nnod;
    // entries the library does not write (state variables past nsv, slip
    // columns past nss) keep the caller's values: the FP64 buffers start from them
    Wide wx(x, D), wdx(dx, D), wfc(fc, D), weps(eps, G), wdep(dep, G), wsv(sv, G * ldsv), wdsv(dsv, G * ldsv);
    Wide wep(ep, 6 * G), wde(de, 6 * G), wst(st, 6 * G), wds(ds, 6 * G), wrt(rt, 3 * G), wdr(dr, 3 * G);
    Wide wsr(srate, G * CPFEM_MAXSS), wcs(cumslip, G * CPFEM_MAXSS);
    const int rc = cpfem_get_state(c, wx.get(), wdx.get(), wfc.get(), weps.get(), wdep.get(), wsv.get(), wdsv.get(),
                                   ldsv, wep.get(), wde.get(), wst.get(), wds.get(), wrt.get(), wdr.get(), wsr.get(),
                                   wcs.get(), twinsys, timestwinned);
    if (rc) return rc;
    wx.narrow(x);
    wdx.narrow(dx);
    wfc.narrow(fc);
    weps.narrow(eps);
    wdep.narrow(dep);
    wsv.narrow(sv);
    wdsv.narrow(dsv);
    wep.narrow(ep);
    wde.narrow(de);
    wst.narrow(st);
    wds.narrow(ds);
    wrt.narrow(rt);
    wdr.narrow(dr);
    wsr.narrow(srate);
    wcs.narrow(cumslip);
    return 0;
}

extern "C" int cpfem_update_node(cpfem_ctx* c) {
    if (!c) return CPFEM_E_ARG;
    const int D = 3 * c->nnod;
    hipLaunchKernelGGL(k_axpy_nodes, dim3(nblk(D)), dim3(kBlock), 0, c->stream, c->x, (const double*)c->dx, 1.0, D);
    CPFEM_HIP(hipGetLastError());
    return 0;
}

// fc = 0, A = 0 and the element assembly (Elvals.f90:80-221), colour by
// colour (the elements of one colour share no node), on the context's stream.
static void enqueue_assembly(Ctx* c) {
    hipStream_t s = c->stream;
    const size_t D = 3 * (size_t)c->nnod, nv = c->A.nvals();
    if (c->A.gather) {
        // element blocks of every element in one launch, then every matrix
        // block and node force sums its contributions (all values written:
        // no zero fill)
        AsmArgs AA;
        AA.nel_color = c->nlmnt;
        AA.elems = nullptr;
        AA.elem_nodes = c->elem_nodes;
        AA.elem_gps = c->elem_gps;
        AA.elem_lt = c->elem_lt;
        AA.elem_np = c->elem_np;
        AA.scatter = nullptr;
        AA.scat_start = nullptr;
        AA.x = c->x;
        AA.chi = c->chi;
        AA.t = c->t;
        AA.ngps = c->ngps;
        AA.fc = c->fc;
        AA.efc = c->A.efc;
        const int epb = c->asm_epb;
        AA.per_xcd = ((c->nlmnt + epb - 1) / epb + kXcd - 1) / kXcd;
        GatherArgs G;
        G.nseg = c->A.ngroups * c->A.P;
        G.per_xcd = group_per_xcd(G.nseg);
        G.order = c->A.g_order;
        G.start = c->A.g_start;
        G.cnt = c->A.g_cnt;
        G.list = c->A.g_list;
        if (c->precision == 1) {
            hipLaunchKernelGGL((k_assemble<float, true>), dim3(AA.per_xcd * kXcd), dim3(64 * epb), 0, s, AA,
                               (float*)c->A.eblk);
            hipLaunchKernelGGL(k_gather_blocks<float>, dim3(G.per_xcd * kXcd), dim3(kBlock), 0, s, G,
                               (const float*)c->A.eblk, (float*)c->A.vals);
        } else {
            hipLaunchKernelGGL((k_assemble<double, true>), dim3(AA.per_xcd * kXcd), dim3(64 * epb), 0, s, AA,
                               (double*)c->A.eblk);
            hipLaunchKernelGGL(k_gather_blocks<double>, dim3(G.per_xcd * kXcd), dim3(kBlock), 0, s, G,
                               (const double*)c->A.eblk, (double*)c->A.vals);
        }
        hipLaunchKernelGGL(k_gather_forces, dim3(nblk(c->nnod)), dim3(kBlock), 0, s, c->nnod,
                           (const int*)c->A.f_start, (const int*)c->A.f_cnt, (const int*)c->A.f_list,
                           (const double*)c->A.efc, c->fc);
        return;
    }
    // zero fc and the matrix values
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(D)), dim3(kBlock), 0, s, c->fc, D, 0.0);
    if (c->precision == 1) {
        hipLaunchKernelGGL(k_fill<float>, dim3(std::min(nblk(nv), 65536)), dim3(kBlock), 0, s, (float*)c->A.vals, nv,
                           0.0f);
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(std::min(nblk(nv), 65536)), dim3(kBlock), 0, s, (double*)c->A.vals, nv,
                           0.0);
    }
    AsmArgs AA;
    AA.per_xcd = 0;
    AA.efc = nullptr;
    AA.elem_nodes = c->elem_nodes;
    AA.elem_gps = c->elem_gps;
    AA.elem_lt = c->elem_lt;
    AA.elem_np = c->elem_np;
    AA.scatter = c->A.scatter;
    AA.scat_start = c->A.scat_start;
    AA.x = c->x;
    AA.chi = c->chi;
    AA.t = c->t;
    AA.ngps = c->ngps;
    AA.fc = c->fc;
    const bool f32 = c->precision == 1;
    const int ncol_el = ((int)c->ekstart.size() - 1) / kKinds;
    for (int k = 0; k < ncol_el; ++k) {
        for (int q = 0; q < kKinds; ++q) {
            const int b0 = c->ekstart[kKinds * k + q], cnt = c->ekstart[kKinds * k + q + 1] - b0;
            if (cnt == 0) continue;
            AA.nel_color = cnt;
            AA.elems = c->elem_by_color + b0;
            if (q == kElem8) {
                const int epb = c->asm_epb;
                AA.per_xcd = ((cnt + epb - 1) / epb + kXcd - 1) / kXcd;
                const int nb = AA.per_xcd * kXcd;
                if (f32)
                    hipLaunchKernelGGL((k_assemble<float, false>), dim3(nb), dim3(64 * epb), 0, s, AA, (float*)c->A.vals);
                else
                    hipLaunchKernelGGL((k_assemble<double, false>), dim3(nb), dim3(64 * epb), 0, s, AA, (double*)c->A.vals);
            } else if (q == kElem20) {
                if (f32)
                    hipLaunchKernelGGL((k_assemble_tab<float, 20>), dim3(cnt), dim3(256), 0, s, AA, (float*)c->A.vals);
                else
                    hipLaunchKernelGGL((k_assemble_tab<double, 20>), dim3(cnt), dim3(256), 0, s, AA,
                                       (double*)c->A.vals);
            } else if (q == kElem10) {
                if (f32)
                    hipLaunchKernelGGL((k_assemble_tab<float, 10>), dim3(cnt), dim3(256), 0, s, AA, (float*)c->A.vals);
                else
                    hipLaunchKernelGGL((k_assemble_tab<double, 10>), dim3(cnt), dim3(256), 0, s, AA,
                                       (double*)c->A.vals);
            } else if (q == kElem15) {
                if (f32)
                    hipLaunchKernelGGL((k_assemble_tab<float, 15>), dim3(cnt), dim3(256), 0, s, AA, (float*)c->A.vals);
                else
                    hipLaunchKernelGGL((k_assemble_tab<double, 15>), dim3(cnt), dim3(256), 0, s, AA,
                                       (double*)c->A.vals);
            } else {
                if (f32)
                    hipLaunchKernelGGL(k_ties<float>, dim3(nblk(cnt)), dim3(kBlock), 0, s, cnt, AA.elems, c->elem_lt,
                                       c->A.scatter, c->A.scat_start, c->ctl.efix, (float*)c->A.vals);
                else
                    hipLaunchKernelGGL(k_ties<double>, dim3(nblk(cnt)), dim3(kBlock), 0, s, cnt, AA.elems, c->elem_lt,
                                       c->A.scatter, c->A.scat_start, c->ctl.efix, (double*)c->A.vals);
            }
        }
    }
}

// ELVALS: point kernel, then zero fc/matrix and the colour-ordered assembly.
static int elvals_impl(Ctx* c, int inc, double* sumcpiter, int32_t* tflag) {
    if (inc < 1 || inc > c->ninc1) return CPFEM_E_ARG;
    hipStream_t s = c->stream;
    const double dt = c->tiem[inc - 1];
    c->rng_epoch += 1;
    c->assembled = false;
    CPFEM_HIP(hipEventRecord(c->ev[0], s));
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, c->flags + 1, (size_t)1, 0);
    // point kinematics (strain and rotation increments, element checks)
    KinArgs K;
    K.n = c->ngps;
    K.gp_elem = c->gp_elem;
    K.gp_local = c->gp_local;
    K.elem_nodes = c->elem_nodes;
    K.elem_lt = c->elem_lt;
    K.elem_nd = c->elem_nd;
    K.x = c->x;
    K.dx = c->dx;
    K.smale = c->ctl.smale;
    K.smalr = c->ctl.smalr;
    K.de = c->de;
    K.r = c->dr;
    K.kerr = c->sched.ck_kerr;
    if (c->npts8 > 0) hipLaunchKernelGGL(k_kin<8>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    if (c->npts20 > 0) hipLaunchKernelGGL(k_kin<20>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    if (c->npts10 > 0) hipLaunchKernelGGL(k_kin<10>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    if (c->npts15 > 0) hipLaunchKernelGGL(k_kin<15>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    ConstArgs A;
    memset(&A, 0, sizeof(A));
    A.n = c->ngps;
    A.mesh = 1;
    A.gp_elem = c->gp_elem;
    A.elem_le = c->elem_le;
    A.de = c->de;
    A.r = c->dr;
    A.dep = c->dep;
    A.st = c->st;
    A.ds = c->ds;
    A.sv = c->sv;
    A.dsv = c->dsv;
    A.t = c->t;
    A.chi = c->chi;
    A.erepmd = c->erepmd;
    A.srate = c->srate;
    A.twinsys = c->twinsys;
    A.timestwinned = c->timestwinned;
    A.flops = c->flops;
    A.transc = c->transc;
    A.err = c->gperr;
    A.tflag = c->flags;
    A.errmax = c->flags + 1;
    A.rng_epoch = c->rng_epoch;
    const DevControl C = make_control(c->ctl);
    int counts[4] = {0, 0, 0, 0};
    {
        const hipError_t he = run_constitutive(c->knss, c->dmat, C, dt, A, c->sched, s, counts);
        if (he == hipErrorNotSupported) return CPFEM_E_WIDTH;
        CPFEM_HIP(he);
    }
    c->stats.pts_continued += counts[0];
    c->stats.pts_wide += counts[1];
    for (int q = 0; q < 4; ++q) c->stats.t_const_pass_ms[q] += c->sched.pass_ms[q];
    CPFEM_HIP(hipEventRecord(c->ev[1], s));
    enqueue_assembly(c);
    CPFEM_HIP(hipGetLastError());
    CPFEM_HIP(hipEventRecord(c->ev[2], s));
    // reductions for the report
    const int nb = std::min(nblk(c->ngps), c->red.nblocks);
    hipLaunchKernelGGL(k_point_sums, dim3(nb), dim3(kBlock), 0, s, (const double*)c->erepmd,
                       (const long long*)c->flops, (const long long*)c->transc, (const int*)c->gperr, c->ngps,
                       c->red.partial);
    double r[4];
    reduce_now(c, nb, r);
    int fl[2];
    CPFEM_HIP(hipMemcpy(fl, c->flags, 2 * sizeof(int), hipMemcpyDeviceToHost));
    c->stats.sumcpiter = r[0];
    c->stats.flops_const += (int64_t)llround(r[1]);
    c->stats.transc_const += (int64_t)llround(r[2]);
    c->stats.gp_updates += c->ngps;
    c->stats.pts_987 += (int64_t)llround(r[3]);
    c->stats.pts_987_last = (int32_t)llround(r[3]);
    c->stats.t_const_ms += elapsed(c->ev[0], c->ev[1]);
    c->stats.t_asm_ms += elapsed(c->ev[1], c->ev[2]);
    if (sumcpiter) *sumcpiter = r[0];
    if (tflag) *tflag = fl[0];
    if (fl[1] > c->stats.err) c->stats.err = fl[1];
    return fl[1];
}

// ELVALS enqueued with every decision and every report value on the device
// (graph capture): the twin-draw epoch is advanced and read on the device, the
// constitutive schedule reads its counts there, the sums land in nd.
static int elvals_enqueue_dev(Ctx* c, NrDev* nd) {
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_epoch_inc, dim3(1), dim3(64), 0, s, nd);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, c->flags + 1, (size_t)1, 0);
    // point kinematics (strain and rotation increments, element checks)
    KinArgs K;
    K.n = c->ngps;
    K.gp_elem = c->gp_elem;
    K.gp_local = c->gp_local;
    K.elem_nodes = c->elem_nodes;
    K.elem_lt = c->elem_lt;
    K.elem_nd = c->elem_nd;
    K.x = c->x;
    K.dx = c->dx;
    K.smale = c->ctl.smale;
    K.smalr = c->ctl.smalr;
    K.de = c->de;
    K.r = c->dr;
    K.kerr = c->sched.ck_kerr;
    if (c->npts8 > 0) hipLaunchKernelGGL(k_kin<8>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    if (c->npts20 > 0) hipLaunchKernelGGL(k_kin<20>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    if (c->npts10 > 0) hipLaunchKernelGGL(k_kin<10>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    if (c->npts15 > 0) hipLaunchKernelGGL(k_kin<15>, dim3(nblk(c->ngps)), dim3(kBlock), 0, s, K);
    ConstArgs A;
    memset(&A, 0, sizeof(A));
    A.n = c->ngps;
    A.mesh = 1;
    A.gp_elem = c->gp_elem;
    A.elem_le = c->elem_le;
    A.de = c->de;
    A.r = c->dr;
    A.dep = c->dep;
    A.st = c->st;
    A.ds = c->ds;
    A.sv = c->sv;
    A.dsv = c->dsv;
    A.t = c->t;
    A.chi = c->chi;
    A.erepmd = c->erepmd;
    A.srate = c->srate;
    A.twinsys = c->twinsys;
    A.timestwinned = c->timestwinned;
    A.flops = c->flops;
    A.transc = c->transc;
    A.err = c->gperr;
    A.tflag = c->flags;
    A.errmax = c->flags + 1;
    A.rng_epoch_dev = &nd->epoch;
    A.dt_dev = &nd->dt;
    const DevControl C = make_control(c->ctl);
    CPFEM_HIP(run_constitutive_dev(c->knss, c->dmat, C, 0.0, A, c->sched, s, nd->report));
    enqueue_assembly(c);
    CPFEM_HIP(hipGetLastError());
    // reductions for the report, into nd
    const int nb = std::min(nblk(c->ngps), c->red.nblocks);
    hipLaunchKernelGGL(k_point_sums, dim3(nb), dim3(kBlock), 0, s, (const double*)c->erepmd,
                       (const long long*)c->flops, (const long long*)c->transc, (const int*)c->gperr, c->ngps,
                       c->red.partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, s, (const double*)c->red.partial, nb, nd->sums);
    hipLaunchKernelGGL(k_flags_to_nr, dim3(1), dim3(64), 0, s, (const int*)c->flags, nd);
    return hipGetLastError() == hipSuccess ? 0 : CPFEM_E_HIP;
}

// SUVALS enqueued with the increment read on the device and the sums into nd.
static void suvals_enqueue_dev(Ctx* c, NrDev* nd) {
    hipStream_t s = c->stream;
    BcArgs B;
    B.nnod = c->nnod;
    B.ib = c->ib;
    B.diag_off = c->A.diag_off;
    B.x = c->x;
    B.dx = c->dx;
    B.fc = c->fc;
    B.xs = c->x_spec;
    B.fs = c->f_spec;
    B.efix = c->ctl.efix;
    B.rhs = c->rhs;
    B.inc_dev = &nd->inc;
    const int nb = nblk(c->nnod);
    if (c->precision == 1)
        hipLaunchKernelGGL(k_suvals<float>, dim3(nb), dim3(kBlock), 0, s, B, (float*)c->A.vals, c->red.partial);
    else
        hipLaunchKernelGGL(k_suvals<double>, dim3(nb), dim3(kBlock), 0, s, B, (double*)c->A.vals, c->red.partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, s, (const double*)c->red.partial, nb, nd->suv);
}

extern "C" int cpfem_elvals(cpfem_ctx* c, int32_t inc, double* sumcpiter, int32_t* tflag) {
    if (!c) return CPFEM_E_ARG;
    return elvals_impl(c, inc, sumcpiter, tflag);
}

static int suvals_impl(Ctx* c, int inc, double* sfc) {
    if (inc < 1 || inc > c->ninc1) return CPFEM_E_ARG;
    hipStream_t s = c->stream;
    CPFEM_HIP(hipEventRecord(c->ev[3], s));
    BcArgs B;
    B.nnod = c->nnod;
    B.ib = c->ib;
    B.diag_off = c->A.diag_off;
    B.x = c->x;
    B.dx = c->dx;
    B.fc = c->fc;
    B.xs = c->x_spec + (size_t)(inc - 1) * 99 * 15;
    B.fs = c->f_spec + (size_t)(inc - 1) * 99 * 3;
    B.efix = c->ctl.efix;
    B.rhs = c->rhs;
    B.inc_dev = nullptr;
    const int nb = nblk(c->nnod);
    if (c->precision == 1)
        hipLaunchKernelGGL(k_suvals<float>, dim3(nb), dim3(kBlock), 0, s, B, (float*)c->A.vals, c->red.partial);
    else
        hipLaunchKernelGGL(k_suvals<double>, dim3(nb), dim3(kBlock), 0, s, B, (double*)c->A.vals, c->red.partial);
    CPFEM_HIP(hipGetLastError());
    double r[4];
    reduce_now(c, nb, r);
    CPFEM_HIP(hipEventRecord(c->ev[4], s));
    CPFEM_HIP(hipEventSynchronize(c->ev[4]));
    c->stats.t_suvals_ms += elapsed(c->ev[3], c->ev[4]);
    c->stats.sfc = sqrt(r[0]);
    if (sfc) *sfc = c->stats.sfc;
    c->assembled = true;
    c->vals32_fresh = false;
    return 0;
}

extern "C" int cpfem_suvals(cpfem_ctx* c, int32_t inc, double* sfc) {
    if (!c) return CPFEM_E_ARG;
    return suvals_impl(c, inc, sfc);
}

// Number of sweeps queued between host checks (one graph replay).
static const int kSweepBatch = 16;
// Most batches between two residual checks of the residual-based stopping rule.
static const int kMaxCheckBatches = 64;


template <typename T, typename V, bool ACC>
static void sor_sweep(Ctx* c, const T* vals, const V* b, V* y, double* partial, const int* stop = nullptr,
                      const double* omega_dev = nullptr) {
    SorArgs S;
    S.P = c->A.P;
    S.cols = c->A.cols;
    S.omega = c->ctl.sor_omega;
    S.stop = stop ? stop : c->flags + 2;
    S.omega_dev = omega_dev;
    int slot = 0;
    for (int q = 0; q < c->A.ncolor; ++q) {
        const int k = c->corder[q];
        S.nrows = c->A.cstart[k + 1] - c->A.cstart[k];
        S.gbase = c->A.gstart[k];
        S.per_xcd = group_per_xcd(c->A.gstart[k + 1] - c->A.gstart[k], c->sor_wpb);
        const int nb = S.per_xcd * kXcd;
        if (S.P == 3)
            hipLaunchKernelGGL((k_sor_color<T, V, ACC, 3>), dim3(nb), dim3(64 * c->sor_wpb), 0, c->stream, S, vals,
                               b, y, partial, slot);
        else
            hipLaunchKernelGGL((k_sor_color<T, V, ACC, 0>), dim3(nb), dim3(64 * c->sor_wpb), 0, c->stream, S, vals,
                               b, y, partial, slot);
        slot += nb;
    }
}

static int sor_partial_blocks(Ctx* c) {
    int slot = 0;
    for (int k = 0; k < c->A.ncolor; ++k) slot += kXcd * group_per_xcd(c->A.gstart[k + 1] - c->A.gstart[k], c->sor_wpb);
    return slot;
}

// Matrix values the FP32 sweeps read: the FP32 matrix itself (precision 1) or
// the rounded copy of the FP64 matrix (precision 2).
static const float* sweep_vals32(const Ctx* c) {
    return c->precision == 2 ? c->A.vals32 : (const float*)c->A.vals;
}

// Precision 2: round the assembled FP64 matrix (penalties included) into the
// copy the FP32 sweeps read, once per assembled system.
static void round_sweep_matrix(Ctx* c) {
    if (c->precision != 2 || c->vals32_fresh) return;
    const size_t nv = c->A.nvals();
    hipLaunchKernelGGL(k_round_vals, dim3((unsigned)std::min<size_t>((nv + kBlock - 1) / kBlock, 65536)),
                       dim3(kBlock), 0, c->stream, (const double*)c->A.vals, c->A.vals32, nv);
    c->vals32_fresh = true;
}

// The kernels of one batch of kSweepBatch sweeps, enqueued on the context's stream.
static void enqueue_batch(Ctx* c, int which) {
    hipStream_t s = c->stream;
    if (which == 0) {
        // FP64 sweeps with the relative-update test after each one
        const double tol2 = c->ctl.sor_tol * c->ctl.sor_tol;
        const int npb = sor_partial_blocks(c);
        for (int k = 0; k < kSweepBatch; ++k) {
            sor_sweep<double, double, true>(c, (const double*)c->A.vals, c->rhs, c->ddx, c->red.partial);
            hipLaunchKernelGGL(k_sor_check, dim3(1), dim3(kBlock), 0, s, (const double*)c->red.partial, npb, tol2,
                               c->flags + 2, c->flags + 3, c->red.result + 4, c->ctl.sor_maxit);
        }
    } else if (which == 1) {
        // FP64 sweeps, residual test by the caller
        for (int k = 0; k < kSweepBatch; ++k)
            sor_sweep<double, double, false>(c, (const double*)c->A.vals, c->rhs, c->ddx, c->red.partial);
    } else {
        // FP32 sweeps on the correction equation A d = r (rhs32, del32); the
        // caller forms r before the first batch of a refinement cycle and adds
        // d to ddx after the last
        for (int k = 0; k < kSweepBatch; ++k)
            sor_sweep<float, float, false>(c, sweep_vals32(c), c->rhs32, c->del32, c->red.partial);
    }
}

// A batch of sweeps captured once as a graph and replayed (kernel nodes only).
// The captured graph is kept until the executable graph is dropped.
static hipError_t sweep_graph(Ctx* c, int which, hipGraphExec_t* exec_out) {
    if (c->graphs[which]) {
        *exec_out = c->graphs[which];
        return hipSuccess;
    }
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    enqueue_batch(c, which);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&c->graphs[which], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return e;
    }
    c->graph_src[which] = g;
    *exec_out = c->graphs[which];
    return hipSuccess;
}

// One batch of sweeps: a replay of its graph, or (graphs switched off with
// CPFEM_SOR_GRAPHS=0, e.g. under a tracing profiler) the same kernels
// launched one by one.
static hipError_t run_batch(Ctx* c, int which) {
    if (!c->use_graphs) {
        enqueue_batch(c, which);
        return hipGetLastError();
    }
    hipGraphExec_t gx;
    hipError_t e = sweep_graph(c, which, &gx);
    if (e != hipSuccess) return e;
    return hipGraphLaunch(gx, c->stream);
}

// r = b - A y in FP64 (res non-null: r kept there) with the partial sums of
// the Jacobi-scaled norms in red.partial; gate non-null: only while *gate.
// Returns the number of partial slots.
static int residual_enqueue(Ctx* c, const double* b, const double* y, double* res, const int* gate = nullptr) {
    SorArgs S;
    S.nrows = c->A.ngroups;
    S.gbase = 0;
    S.P = c->A.P;
    S.cols = c->A.cols;
    S.omega = c->ctl.sor_omega;
    S.stop = nullptr;
    S.gate = gate;
    S.per_xcd = group_per_xcd(c->A.ngroups);
    const int nb = S.per_xcd * kXcd;
    if (c->precision == 1) {
        if (S.P == 3)
            hipLaunchKernelGGL((k_residual<float, 3>), dim3(nb), dim3(kBlock), 0, c->stream, S,
                               (const float*)c->A.vals, b, y, res, c->red.partial);
        else
            hipLaunchKernelGGL((k_residual<float, 0>), dim3(nb), dim3(kBlock), 0, c->stream, S,
                               (const float*)c->A.vals, b, y, res, c->red.partial);
    } else {
        if (S.P == 3)
            hipLaunchKernelGGL((k_residual<double, 3>), dim3(nb), dim3(kBlock), 0, c->stream, S,
                               (const double*)c->A.vals, b, y, res, c->red.partial);
        else
            hipLaunchKernelGGL((k_residual<double, 0>), dim3(nb), dim3(kBlock), 0, c->stream, S,
                               (const double*)c->A.vals, b, y, res, c->red.partial);
    }
    return nb;
}

static double residual(Ctx* c, const double* b, const double* y, double* res) {
    const int nb = residual_enqueue(c, b, y, res);
    double r[4];
    reduce_now(c, nb, r);
    if (!std::isfinite(r[1])) return std::nan("");
    return r[1] > 0.0 ? sqrt(r[0] / r[1]) : 0.0;
}

// One unit of a residual-rule solve: a batch of kSweepBatch sweeps and the
// stopping-rule kernels after it (NrDev).  The sweeps read the relaxation
// factor from the record.  FP64: the sweeps relax A ddx = rhs.  Mixed: a
// stretch between two checks is one refinement cycle, the FP32 sweeps relax
// A d = r from d = 0 with r the FP64 residual in xold, and ddx += d at its end.
static void enqueue_unit(Ctx* c) {
    hipStream_t s = c->stream;
    NrDev* nd = c->nrdev;
    const int D = 3 * c->nnod;
    const bool mixed = c->precision != 0;
    if (mixed)
        hipLaunchKernelGGL(k_cycle_start, dim3(nblk(D)), dim3(kBlock), 0, s, (const NrDev*)nd,
                           (const double*)c->xold, c->rhs32, c->del32, D);
    for (int k = 0; k < kSweepBatch; ++k) {
        if (mixed)
            sor_sweep<float, float, false>(c, sweep_vals32(c), c->rhs32, c->del32, c->red.partial, &nd->stop,
                                           &nd->omega);
        else
            sor_sweep<double, double, false>(c, (const double*)c->A.vals, c->rhs, c->ddx, c->red.partial, &nd->stop,
                                             &nd->omega);
    }
    hipLaunchKernelGGL(k_solve_after, dim3(1), dim3(64), 0, s, nd, kSweepBatch);
    if (mixed)
        hipLaunchKernelGGL(k_cycle_end, dim3(nblk(D)), dim3(kBlock), 0, s, (const NrDev*)nd, c->ddx,
                           (const float*)c->del32, D);
    const int nb = residual_enqueue(c, c->rhs, c->ddx, mixed ? c->xold : nullptr, &nd->check_now);
    hipLaunchKernelGGL(k_solve_check, dim3(1), dim3(kBlock), 0, s, (const double*)c->red.partial, nb, nd,
                       c->ctl.sor_tol, c->ctl.sor_maxit, kSweepBatch, kMaxCheckBatches);
}

// Start of a residual-rule solve: ddx = 0 (unless warm), the record's
// stopping-rule state, the relaxation factor, and (mixed) the first
// refinement cycle's residual r = rhs.  flags: see k_solve_init.
static void enqueue_solve_init(Ctx* c, bool warm, const int* flags) {
    hipStream_t s = c->stream;
    const int D = 3 * c->nnod;
    if (!warm) hipLaunchKernelGGL(k_fill<double>, dim3(nblk(D)), dim3(kBlock), 0, s, c->ddx, (size_t)D, 0.0);
    hipLaunchKernelGGL(k_solve_init, dim3(1), dim3(64), 0, s, c->nrdev, flags, c->ctl.sor_maxit);
    if (c->precision != 0)
        hipLaunchKernelGGL(k_copy, dim3(nblk(D)), dim3(kBlock), 0, s, c->xold, (const double*)c->rhs, D);
}

// The damped Newton update; gated: only after the solve has stopped and once.
static void enqueue_update(Ctx* c, bool gated) {
    hipStream_t s = c->stream;
    NrDev* nd = c->nrdev;
    const NrDev* gate = gated ? nd : nullptr;
    const int D = 3 * c->nnod;
    const int nb = std::min(nblk(D), c->red.nblocks);
    hipLaunchKernelGGL(k_norms2_g, dim3(nb), dim3(kBlock), 0, s, (const double*)c->dx, (const double*)c->ddx, D,
                       c->red.partial, gate);
    hipLaunchKernelGGL(k_nr_alpha, dim3(1), dim3(kBlock), 0, s, (const double*)c->red.partial, nb, nd,
                       c->ctl.toler, c->ctl.alpha, gate);
    hipLaunchKernelGGL(k_nr_update_g, dim3(nblk(D)), dim3(kBlock), 0, s, c->x, c->dx, (const double*)c->ddx,
                       (const NrDev*)nd, D, gate);
    if (gated) hipLaunchKernelGGL(k_nr_mark, dim3(1), dim3(64), 0, s, nd);
}

// Capture of the kernels enqueue(c) puts on the context's stream as an
// executable graph (kernel nodes only).
template <typename F>
static hipError_t capture(Ctx* c, F enqueue, hipGraph_t* src, hipGraphExec_t* exec) {
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(c->stream, &g);
    if (e == hipSuccess && rc != 0) e = hipErrorInvalidValue;
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return e;
    }
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        *exec = nullptr;
        return e;
    }
    *src = g;
    return hipSuccess;
}

static void check_solve_graphs(Ctx* c);

// One unit: a replay of its graph or (CPFEM_SOR_GRAPHS=0) its kernels.
static hipError_t run_unit(Ctx* c) {
    if (!c->use_graphs) {
        enqueue_unit(c);
        return hipGetLastError();
    }
    check_solve_graphs(c);
    if (!c->g_unit) {
        hipError_t e = capture(
            c, [&]() { enqueue_unit(c); return hipGetLastError() == hipSuccess ? 0 : 1; }, &c->g_unit_src,
            &c->g_unit);
        if (e != hipSuccess) return e;
    }
    return hipGraphLaunch(c->g_unit, c->stream);
}

static hipError_t read_nrdev(Ctx* c) {
    hipError_t e = hipMemcpyAsync(c->nrhost, c->nrdev, sizeof(NrDev), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

static int autotune_impl(Ctx* c, int flags);

// Relaxation factor that follows the solves.  The ladder tuning of the first
// solve sees the first (elastic) system; the stiffness softens as the points
// yield and the best factor moves.  Each residual-rule solve measures its
// average rate (log of its final relative residual per sweep, k_solve_check),
// and a coordinate search moves omega: after a solve at the best factor, the
// next runs at best + dir*step; a trial that converges faster per sweep becomes
// the best and the search goes on in that direction, one that does not turns
// it round, halves the step (not below 0.0025) and returns to the best factor,
// whose rate the next solve measures again (the systems drift).  Two
// successes in a row in one direction double the step (up to 0.02), so the
// search keeps up with the factor's drift while the specimen yields.  Near the
// optimum the sweep count rises steeply on either side (64^3 plane strain, a
// steady increment's system: 2432 sweeps at 1.965, 1760 at 1.970, 2032 at
// 1.975).  Any factor in (0, 2) converges on these SPD systems, so a trial
// costs sweeps, never the solution.  The rates are deterministic reductions:
// so is the sequence.
static void drop_graphs(Ctx* c);
static void refine_omega(Ctx* c, double rate) {
    if (!c->om_adapt || !(rate < 0.0) || !std::isfinite(rate)) return;
    const double lo = 1.0, hi = 1.995, min_step = 0.0025, max_step = 0.02;
    double next;
    if (!c->om_trial) {
        c->om_rate = rate;
        next = c->om_best + c->om_dir * c->om_step;
        if (next < lo || next > hi) {
            c->om_dir = -c->om_dir;
            next = c->om_best + c->om_dir * c->om_step;
        }
        c->om_trial = true;
    } else if (rate < c->om_rate) {
        c->om_best = c->ctl.sor_omega;
        c->om_rate = rate;
        if (++c->om_wins >= 2) {
            c->om_wins = 0;
            c->om_step = std::min(2.0 * c->om_step, max_step);
        }
        next = c->om_best + c->om_dir * c->om_step;
        if (next < lo || next > hi) {
            c->om_dir = -c->om_dir;
            c->om_step = std::max(0.5 * c->om_step, min_step);
            next = c->om_best;
            c->om_trial = false;
        }
    } else {
        c->om_wins = 0;
        c->om_dir = -c->om_dir;
        c->om_step = std::max(0.5 * c->om_step, min_step);
        next = c->om_best;
        c->om_trial = false;
    }
    next = std::min(hi, std::max(lo, next));
    if (next != c->ctl.sor_omega) {
        c->ctl.sor_omega = next;
        drop_graphs(c);
    }
}

// FACSLS: solve A ddx = rhs from a zero start.
// warm: start from the ddx in place instead of zero (the reference starts from
// zero, Stif_sol.f90:179; the calibration's solves of consecutive increments
// start from the previous solution)
static int facsls_impl(Ctx* c, int* sweeps_out, double* err_out, bool warm = false) {
    round_sweep_matrix(c);
    if (c->omega_auto) {
        const int rc = autotune_impl(c, CPFEM_TUNE_LAUNCH | CPFEM_TUNE_OMEGA);
        if (rc) return rc;
        c->omega_auto = false;
        c->om_adapt = true;
        c->om_best = c->ctl.sor_omega;
        c->om_trial = false;
        c->om_step = 0.01;
        c->om_dir = 1;
        c->om_wins = 0;
    }
    hipStream_t s = c->stream;
    const int D = 3 * c->nnod;
    const int maxit = c->ctl.sor_maxit;
    const double omega_used = c->ctl.sor_omega;
    double final_rate = 0.0;
    CPFEM_HIP(hipEventRecord(c->ev[5], s));
    if (!warm) hipLaunchKernelGGL(k_fill<double>, dim3(nblk(D)), dim3(kBlock), 0, s, c->ddx, (size_t)D, 0.0);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, c->flags + 2, (size_t)2, 0);
    hipLaunchKernelGGL(k_set_omega, dim3(1), dim3(64), 0, s, c->nrdev, c->ctl.sor_omega);
    int sweeps = 0, code = 0;
    double err = 0.0;
    if (c->precision == 0 && c->ctl.sor_kind == 0) {
        // the reference's relative-update test after every sweep, decided on the device
        const double tol2 = c->ctl.sor_tol * c->ctl.sor_tol;
        int hf[2] = {0, 0};
        while (true) {
            CPFEM_HIP(hipEventRecord(c->ev[7], s));
            CPFEM_HIP(run_batch(c, 0));
            CPFEM_HIP(hipEventRecord(c->ev[8], s));
            CPFEM_HIP(hipMemcpyAsync(c->red.host + 4, c->flags + 2, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
            CPFEM_HIP(hipStreamSynchronize(s));
            c->stats.t_sweep_ms += elapsed(c->ev[7], c->ev[8]);
            c->stats.sweep_launches += (int64_t)kSweepBatch * c->A.ncolor;
            memcpy(hf, c->red.host + 4, 2 * sizeof(int));
            if (hf[0]) break;
        }
        sweeps = hf[1];
        CPFEM_HIP(hipMemcpy(&err, c->red.result + 4, sizeof(double), hipMemcpyDeviceToHost));
        if (!(err <= 10.0 * tol2)) code = 904;
    } else {
        // relative Jacobi-scaled residual, checked after a number of 16-sweep
        // batches predicted from the observed geometric convergence rate (at
        // most 80 % of the predicted remainder, so the test is not passed by
        // more than one batch).  In mixed precision each stretch between two
        // residuals is one refinement cycle: the sweeps run in FP32 on the
        // correction equation A d = r from d = 0, r formed in FP64 by the
        // previous residual, then ddx += d.  The decisions are taken on the
        // device (enqueue_unit); the host replays the stretch's units and
        // reads the record after each check.
        enqueue_solve_init(c, true, nullptr);
        CPFEM_HIP(hipGetLastError());
        const NrDev& h = *c->nrhost;
        int left = 1;
        while (true) {
            CPFEM_HIP(hipEventRecord(c->ev[7], s));
            for (int k = 0; k < left; ++k) CPFEM_HIP(run_unit(c));
            CPFEM_HIP(hipEventRecord(c->ev[8], s));
            CPFEM_HIP(read_nrdev(c));
            c->stats.t_sweep_ms += elapsed(c->ev[7], c->ev[8]);
            c->stats.sweep_launches += (int64_t)left * kSweepBatch * c->A.ncolor;
            if (h.stop) break;
            left = h.left;
        }
        sweeps = h.sweeps;
        err = h.rel;
        final_rate = h.final_rate;
        c->last_batches = sweeps / kSweepBatch;
        if (!(err <= 10.0 * c->ctl.sor_tol)) code = 904;
    }
    CPFEM_HIP(hipEventRecord(c->ev[6], s));
    CPFEM_HIP(hipEventSynchronize(c->ev[6]));
    c->stats.t_sor_ms += elapsed(c->ev[5], c->ev[6]);
    c->stats.sor_sweeps = sweeps;
    c->stats.sor_sweeps_total += sweeps;
    c->stats.sor_error = err;
    c->stats.sor_omega = omega_used;
    c->stats.sor_wpb = c->sor_wpb;
    if (code == 0 && !warm) refine_omega(c, final_rate);
    if (sweeps_out) *sweeps_out = sweeps;
    if (err_out) *err_out = err;
    if (code > c->stats.err) c->stats.err = code;
    return code;
}

// ---------------------------------------------------------------------------
// Sweep tuning.  The launch shape changes only how row groups are packed into
// workgroups (each wave relaxes its own rows), so the iterates are the same at
// every shape; it is chosen by timing replays of the solve's own sweep graph.
// The relaxation factor changes the iterates: each candidate of a ladder runs
// two windows of W sweeps from ddx = 0 (in mixed precision each window is one
// refinement cycle, as in the solve), and the sweeps to the solve's tolerance
// are extrapolated from the residuals after each window at the second
// window's geometric rate.  The ladder is walked upwards and stops two steps
// past the best candidate (the sweep count is unimodal in omega for SOR on an
// SPD system).
static void drop_graphs(Ctx* c) {
    for (int i = 0; i < 3; ++i) {
        if (c->graphs[i]) {
            (void)hipGraphExecDestroy(c->graphs[i]);
            c->graphs[i] = nullptr;
        }
        if (c->graph_src[i]) {
            (void)hipGraphDestroy(c->graph_src[i]);
            c->graph_src[i] = nullptr;
        }
    }
}

static void drop_solve_graphs(Ctx* c) {
    hipGraphExec_t* ex[3] = {&c->g_unit, &c->g_it, &c->g_more};
    hipGraph_t* src[3] = {&c->g_unit_src, &c->g_it_src, &c->g_more_src};
    for (int i = 0; i < 3; ++i) {
        if (*ex[i]) (void)hipGraphExecDestroy(*ex[i]);
        if (*src[i]) (void)hipGraphDestroy(*src[i]);
        *ex[i] = nullptr;
        *src[i] = nullptr;
    }
}

// The unit and iteration graphs bake in the launch shapes: captured again
// when one of them has changed since.
static void check_solve_graphs(Ctx* c) {
    const int key[4] = {c->sor_wpb, c->asm_epb, c->sched.first_iters, c->A.ncolor};
    if (memcmp(key, c->g_key, sizeof(key)) != 0) {
        drop_solve_graphs(c);
        memcpy(c->g_key, key, sizeof(key));
    }
}

// Graph the solve replays for this context's settings.
static int solve_graph_kind(const Ctx* c) { return c->precision != 0 ? 2 : (c->ctl.sor_kind == 0 ? 0 : 1); }

// Device time (ms) of `reps` replays of the solve's sweep graph from a zero start.
static hipError_t time_sweep_graph(Ctx* c, int reps, float* ms) {
    hipStream_t s = c->stream;
    const int D = 3 * c->nnod;
    const int which = solve_graph_kind(c);
    hipError_t e;
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        if (which == 2) {
            hipLaunchKernelGGL(k_to_float, dim3(nblk(D)), dim3(kBlock), 0, s, (const double*)c->rhs, c->rhs32, D);
            hipLaunchKernelGGL(k_fill<float>, dim3(nblk(D)), dim3(kBlock), 0, s, c->del32, (size_t)D, 0.0f);
        } else {
            hipLaunchKernelGGL(k_fill<double>, dim3(nblk(D)), dim3(kBlock), 0, s, c->ddx, (size_t)D, 0.0);
        }
        hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, c->flags + 2, (size_t)2, 0);
        if ((e = hipEventRecord(c->ev[7], s)) != hipSuccess) return e;
        for (int k = 0; k < reps; ++k)
            if ((e = run_batch(c, which)) != hipSuccess) return e;
        if ((e = hipEventRecord(c->ev[8], s)) != hipSuccess) return e;
        if ((e = hipEventSynchronize(c->ev[8])) != hipSuccess) return e;
        if (rep > 0) best = std::min(best, elapsed(c->ev[7], c->ev[8]));   // the first is a warm-up
    }
    *ms = best;
    return hipSuccess;
}

// One window of W sweeps continuing from the current ddx (and, in mixed
// precision, the residual held in xold); returns the relative residual.
static hipError_t tune_window(Ctx* c, int W, double* rel) {
    hipStream_t s = c->stream;
    const int D = 3 * c->nnod;
    const bool mixed = c->precision != 0;
    hipError_t e;
    if (mixed) {
        hipLaunchKernelGGL(k_to_float, dim3(nblk(D)), dim3(kBlock), 0, s, (const double*)c->xold, c->rhs32, D);
        hipLaunchKernelGGL(k_fill<float>, dim3(nblk(D)), dim3(kBlock), 0, s, c->del32, (size_t)D, 0.0f);
    }
    for (int k = 0; k < W / kSweepBatch; ++k)
        if ((e = run_batch(c, mixed ? 2 : 1)) != hipSuccess) return e;
    if (mixed) hipLaunchKernelGGL(k_add_float, dim3(nblk(D)), dim3(kBlock), 0, s, c->ddx, (const float*)c->del32, D);
    *rel = residual(c, c->rhs, c->ddx, mixed ? c->xold : nullptr);
    return hipGetLastError();
}

// Sweeps to the tolerance predicted from the residuals r1, r2 after W and 2W sweeps.
static double predicted_sweeps(double r1, double r2, int W, double tol) {
    if (!(r1 > 0.0) || !std::isfinite(r1)) return r1 == 0.0 ? 0.0 : 1e300;
    if (r1 <= tol) return W * log(tol) / log(std::min(r1, 0.999));
    if (!(r2 > 0.0) || !std::isfinite(r2) || r2 >= r1) return r2 == 0.0 ? W : 1e300;
    const double rate = log(r2 / r1) / W;
    return r2 <= tol ? W + log(tol / r1) / rate : 2.0 * W + log(tol / r2) / rate;
}

// Candidate colour orders of the order tuning: the order in force and, with
// 8 colours, the six binary-reflected Gray codes of the colour numbers (the
// bit flipped at step k is the lowest set bit of k, with the three bits given
// the roles of a permutation).  The greedy colouring of a brick mesh numbered
// x fastest gives colour px + 2 py + 4 pz (node parities), so these visit the
// parity classes one axis flip at a time; measured on the 64^3 plane-strain
// increments they need 20-30 % fewer sweeps than the colour-number order.
static std::vector<std::vector<int>> order_candidates(const Ctx* c) {
    std::vector<std::vector<int>> out{c->corder};
    if (c->A.ncolor != 8) return out;
    const int roles[6][3] = {{1, 2, 4}, {1, 4, 2}, {2, 1, 4}, {2, 4, 1}, {4, 1, 2}, {4, 2, 1}};
    for (const auto& r : roles) {
        std::vector<int> o(8);
        for (int k = 0; k < 8; ++k) {
            const int g = k ^ (k >> 1);   // bits b0 b1 b2 of the reflected code
            o[k] = ((g & 1) ? r[0] : 0) | ((g & 2) ? r[1] : 0) | ((g & 4) ? r[2] : 0);
        }
        if (o != c->corder) out.push_back(o);
    }
    return out;
}

static int autotune_impl(Ctx* c, int flags) {
    if (!c->assembled) return CPFEM_E_ARG;
    round_sweep_matrix(c);
    hipStream_t s = c->stream;
    const int D = 3 * c->nnod;
    const auto t0 = std::chrono::steady_clock::now();
    double* saved = nullptr;
    CPFEM_HIP(hipMalloc((void**)&saved, D * sizeof(double)));
    int rc = 0;
    auto fail = [&](hipError_t e) {
        rc = e == hipSuccess ? 0 : CPFEM_E_HIP;
        return rc;
    };
    (void)hipMemcpyAsync(saved, c->ddx, D * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (flags & CPFEM_TUNE_LAUNCH) {
        int best_w = c->sor_wpb;
        float best_ms = 1e30f;
        for (int w : {4, 2, 1}) {
            c->sor_wpb = w;
            if (sor_partial_blocks(c) > c->red.nblocks) continue;
            drop_graphs(c);
            float ms = 0.0f;
            if (fail(time_sweep_graph(c, 4, &ms))) break;
            c->stats.tune_sweeps += 3 * 4 * kSweepBatch;
            if (ms < 0.98f * best_ms) {
                best_ms = ms;
                best_w = w;
            }
        }
        c->sor_wpb = best_w;
        drop_graphs(c);
    }
    if (!rc && (flags & CPFEM_TUNE_OMEGA)) {
        static const double ladder[] = {1.0, 1.3, 1.5, 1.65, 1.75, 1.82, 1.87, 1.91, 1.94, 1.96, 1.98};
        // window: an eighth of the sweep limit, 64 to 256 sweeps, whole batches
        const int W = std::max(64, std::min(256, c->ctl.sor_maxit / 8)) / kSweepBatch * kSweepBatch;
        const double tol = c->precision != 0 ? std::max(c->ctl.sor_tol, 1e-12) : std::max(c->ctl.sor_tol, 1e-14);
        double best_om = c->ctl.sor_omega, best_n = 1e300;
        int since_best = 0;
        for (double om : ladder) {
            c->ctl.sor_omega = om;
            drop_graphs(c);
            hipLaunchKernelGGL(k_fill<double>, dim3(nblk(D)), dim3(kBlock), 0, s, c->ddx, (size_t)D, 0.0);
            if (c->precision != 0)
                (void)hipMemcpyAsync(c->xold, c->rhs, D * sizeof(double), hipMemcpyDeviceToDevice, s);
            double r1 = 0.0, r2 = 0.0;
            if (fail(tune_window(c, W, &r1))) break;
            if (r1 > c->ctl.sor_tol && std::isfinite(r1))
                if (fail(tune_window(c, W, &r2))) break;
            c->stats.tune_sweeps += (r2 > 0.0 ? 2 : 1) * W;
            const double n = predicted_sweeps(r1, r2, W, tol);
            if (n < best_n) {
                best_n = n;
                best_om = om;
                since_best = 0;
            } else if (++since_best >= 2) {
                break;
            }
        }
        c->ctl.sor_omega = best_om;
        drop_graphs(c);
    }
    if (!rc && (flags & CPFEM_TUNE_ORDER)) {
        // each candidate order: two windows from a zero start at the current
        // omega; a candidate replaces the order in force only when it is
        // predicted to need at least 3 % fewer sweeps
        const int W = std::max(64, std::min(256, c->ctl.sor_maxit / 8)) / kSweepBatch * kSweepBatch;
        const double tol = c->precision != 0 ? std::max(c->ctl.sor_tol, 1e-12) : std::max(c->ctl.sor_tol, 1e-14);
        const auto cands = order_candidates(c);
        std::vector<int> best = c->corder;
        double best_n = 1e300;
        for (size_t q = 0; q < cands.size() && !rc; ++q) {
            c->corder = cands[q];
            drop_graphs(c);
            hipLaunchKernelGGL(k_fill<double>, dim3(nblk(D)), dim3(kBlock), 0, s, c->ddx, (size_t)D, 0.0);
            if (c->precision != 0)
                (void)hipMemcpyAsync(c->xold, c->rhs, D * sizeof(double), hipMemcpyDeviceToDevice, s);
            double r1 = 0.0, r2 = 0.0;
            if (fail(tune_window(c, W, &r1))) break;
            if (r1 > c->ctl.sor_tol && std::isfinite(r1))
                if (fail(tune_window(c, W, &r2))) break;
            c->stats.tune_sweeps += (r2 > 0.0 ? 2 : 1) * W;
            const double n = predicted_sweeps(r1, r2, W, tol);
            if (q == 0 || n < 0.97 * best_n) {
                best_n = n;
                best = cands[q];
            }
        }
        c->corder = best;
        drop_graphs(c);
        drop_solve_graphs(c);
    }
    (void)hipMemcpyAsync(c->ddx, saved, D * sizeof(double), hipMemcpyDeviceToDevice, s);
    const hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(saved);
    if (!rc && e != hipSuccess) rc = CPFEM_E_HIP;
    c->stats.t_tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->stats.sor_omega = c->ctl.sor_omega;
    c->stats.sor_wpb = c->sor_wpb;
    return rc;
}

extern "C" int cpfem_autotune(cpfem_ctx* c, int32_t flags, double* omega, int32_t* waves_per_block) {
    if (!c || (flags & ~(CPFEM_TUNE_LAUNCH | CPFEM_TUNE_OMEGA | CPFEM_TUNE_ORDER))) return CPFEM_E_ARG;
    if (flags == 0) {
        // apply the given settings
        const double om = omega ? *omega : 0.0;
        const int w = waves_per_block ? *waves_per_block : 0;
        if (!(om < 2.0) || (w != 0 && w != 1 && w != 2 && w != 4)) return CPFEM_E_ARG;
        if (w > 0) {
            const int keep = c->sor_wpb;
            c->sor_wpb = w;
            if (sor_partial_blocks(c) > c->red.nblocks) {
                c->sor_wpb = keep;
                return CPFEM_E_ARG;
            }
        }
        if (om > 0.0) {
            c->ctl.sor_omega = om;
            c->omega_auto = false;
            c->om_adapt = false;
        }
        drop_graphs(c);
    } else {
        const int rc = autotune_impl(c, flags);
        if (rc) return rc;
        if ((flags & CPFEM_TUNE_OMEGA) && c->om_adapt) {
            c->om_best = c->ctl.sor_omega;
            c->om_trial = false;
            c->om_wins = 0;
        }
    }
    if (flags & CPFEM_TUNE_OMEGA) c->omega_auto = false;
    if (omega) *omega = c->ctl.sor_omega;
    if (waves_per_block) *waves_per_block = c->sor_wpb;
    return 0;
}

extern "C" int cpfem_sweep_order(cpfem_ctx* c, int32_t* order, int32_t n, int32_t set) {
    if (!c || !order) return CPFEM_E_ARG;
    const int nc = c->A.ncolor;
    if (n < nc) return CPFEM_E_ARG;
    if (!set) {
        for (int k = 0; k < nc; ++k) order[k] = c->corder[k];
        return nc;
    }
    if (n != nc) return CPFEM_E_ARG;
    std::vector<int> seen(nc, 0), o(nc);
    for (int k = 0; k < nc; ++k) {
        if (order[k] < 0 || order[k] >= nc || seen[order[k]]++) return CPFEM_E_ARG;
        o[k] = order[k];
    }
    if (o != c->corder) {
        c->corder = o;
        drop_graphs(c);
        drop_solve_graphs(c);
    }
    return nc;
}

// Launch shapes of the point kernels: the outer iterations every point gets in
// the constitutive kernel's first pass (the rest resume in later passes) and
// the elements per workgroup of the 8-node assembly.  Neither changes a
// result (every point and every element is computed by the same code), only
// how the work is packed, so they are chosen by timing ELVALS on the current
// state: four runs with first-pass budgets 4, 8, 16, 32 and 1, 2, 4, 4
// elements per workgroup, each timed by its events, the state ELVALS
// overwrites restored after each run.
static int autotune_points_impl(Ctx* c, int inc) {
    hipStream_t s = c->stream;
    const size_t G = c->ngps;
    struct Keep { void* d; size_t bytes; void* saved; } keep[] = {
        {c->ds, 6 * G * 8, nullptr}, {c->dsv, (size_t)c->nsv * G * 8, nullptr},
        {c->srate, (size_t)c->nss * G * 8, nullptr}, {c->flops, G * 8, nullptr},
        {c->twinsys, G * 4, nullptr}, {c->flags, 2 * 4, nullptr}};
    int rc = 0;
    for (auto& k : keep)
        if (hipMalloc(&k.saved, k.bytes) != hipSuccess) rc = CPFEM_E_HIP;
    if (!rc)
        for (auto& k : keep) (void)hipMemcpyAsync(k.saved, k.d, k.bytes, hipMemcpyDeviceToDevice, s);
    const uint64_t epoch = c->rng_epoch;
    const cpfem_stats st0 = c->stats;
    const int fis[4] = {4, 8, 16, 32}, epbs[4] = {1, 2, 4, 4};
    double tc[4] = {0, 0, 0, 0}, ta[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4 && !rc; ++k) {
        c->sched.first_iters = fis[k];
        c->asm_epb = epbs[k];
        c->rng_epoch = epoch;
        memset(&c->stats, 0, sizeof(c->stats));
        const int code = elvals_impl(c, inc, nullptr, nullptr);
        if (code < 0) rc = code;
        tc[k] = c->stats.t_const_ms;
        ta[k] = c->stats.t_asm_ms;
        for (auto& q : keep) (void)hipMemcpyAsync(q.d, q.saved, q.bytes, hipMemcpyDeviceToDevice, s);
    }
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = CPFEM_E_HIP;
    for (auto& k : keep)
        if (k.saved) (void)hipFree(k.saved);
    c->rng_epoch = epoch;
    c->stats = st0;
    c->assembled = false;   // fc and the matrix hold a candidate's ELVALS: the caller runs ELVALS again
    if (rc) {
        c->sched.first_iters = 8;
        c->asm_epb = 4;
        return rc;
    }
    int bf = 0;
    for (int k = 1; k < 4; ++k)
        if (tc[k] < tc[bf]) bf = k;
    c->sched.first_iters = fis[bf];
    const double a1 = ta[0], a2 = ta[1], a4 = 0.5 * (ta[2] + ta[3]);
    c->asm_epb = (a1 <= a2 && a1 <= a4) ? 1 : (a2 <= a4 ? 2 : 4);
    return 0;
}

extern "C" int cpfem_autotune_points(cpfem_ctx* c, int32_t inc, int32_t* first_iters, int32_t* elems_per_block) {
    if (!c) return CPFEM_E_ARG;
    if (inc == 0) {
        const int fi = first_iters ? *first_iters : 0, ep = elems_per_block ? *elems_per_block : 0;
        if (fi < 0 || fi > 1000 || (ep != 0 && ep != 1 && ep != 2 && ep != 4)) return CPFEM_E_ARG;
        if (fi > 0) c->sched.first_iters = fi;
        if (ep > 0) c->asm_epb = ep;
    } else {
        if (inc < 1 || inc > c->ninc1) return CPFEM_E_ARG;
        const int rc = autotune_points_impl(c, inc);
        if (rc) return rc;
    }
    if (first_iters) *first_iters = c->sched.first_iters;
    if (elems_per_block) *elems_per_block = c->asm_epb;
    return 0;
}

extern "C" int cpfem_facsls(cpfem_ctx* c, int32_t* sweeps, double* error) {
    if (!c) return CPFEM_E_ARG;
    return facsls_impl(c, sweeps, error);
}

static int nr_update_impl(Ctx* c, double* sdx_out, double* alp2_out) {
    hipStream_t s = c->stream;
    CPFEM_HIP(hipEventRecord(c->ev[5], s));
    enqueue_update(c, false);
    CPFEM_HIP(hipGetLastError());
    CPFEM_HIP(hipEventRecord(c->ev[6], s));
    CPFEM_HIP(read_nrdev(c));
    c->stats.t_update_ms += elapsed(c->ev[5], c->ev[6]);
    c->stats.sdx = c->nrhost->sdx;
    c->stats.alp2 = c->nrhost->alp2;
    if (sdx_out) *sdx_out = c->stats.sdx;
    if (alp2_out) *alp2_out = c->stats.alp2;
    return 0;
}

extern "C" int cpfem_nr_update(cpfem_ctx* c, double* sdx, double* alp2) {
    if (!c) return CPFEM_E_ARG;
    return nr_update_impl(c, sdx, alp2);
}

// Graph mode.  One Newton iteration is one launch of the iteration graph:
// ELVALS (kinematics, the constitutive passes and tiers with their counts on
// the device, assembly, point sums), SUVALS, precision 2's rounded matrix,
// the solve's start, kItUnits units and the gated update.  A solve that needs
// more batches continues in replays of the continuation graph (kMoreUnits
// units and the gated update), as many as the last solve's length predicts
// queued at once, then one at a time; the host reads the record once they
// have run.  Units after the solve's stop are skipped kernel by kernel.  The
// kernels and their order are those of the call-by-call path, so both give
// the same iterates.
static const int kItUnits = 2;
static const int kMoreUnits = 4;

static int enqueue_iteration(Ctx* c) {
    NrDev* nd = c->nrdev;
    const int rc = elvals_enqueue_dev(c, nd);
    if (rc) return rc;
    suvals_enqueue_dev(c, nd);
    if (c->precision == 2) {
        const size_t nv = c->A.nvals();
        hipLaunchKernelGGL(k_round_vals, dim3((unsigned)std::min<size_t>((nv + kBlock - 1) / kBlock, 65536)),
                           dim3(kBlock), 0, c->stream, (const double*)c->A.vals, c->A.vals32, nv);
    }
    enqueue_solve_init(c, false, c->flags);
    for (int u = 0; u < kItUnits; ++u) enqueue_unit(c);
    enqueue_update(c, true);
    return hipGetLastError() == hipSuccess ? 0 : CPFEM_E_HIP;
}

static int enqueue_more(Ctx* c) {
    for (int u = 0; u < kMoreUnits; ++u) enqueue_unit(c);
    enqueue_update(c, true);
    return hipGetLastError() == hipSuccess ? 0 : CPFEM_E_HIP;
}

// One Newton iteration in graph mode; the report values go to the stats.
// Returns the iteration's code as the call-by-call sequence would.
static int graph_iteration(Ctx* c, int inc, double* sdx) {
    hipStream_t s = c->stream;
    check_solve_graphs(c);
    if (!c->g_it) CPFEM_HIP(capture(c, [&]() { return enqueue_iteration(c); }, &c->g_it_src, &c->g_it));
    if (!c->g_more) CPFEM_HIP(capture(c, [&]() { return enqueue_more(c); }, &c->g_more_src, &c->g_more));
    hipLaunchKernelGGL(k_nr_set, dim3(1), dim3(64), 0, s, c->nrdev, inc, c->tiem[inc - 1], c->ctl.sor_omega,
                       c->rng_epoch);
    const double omega_used = c->ctl.sor_omega;
    CPFEM_HIP(hipEventRecord(c->ev[0], s));
    CPFEM_HIP(hipGraphLaunch(c->g_it, s));
    int launches = 1;
    const int predicted = (c->last_batches - kItUnits + kMoreUnits - 1) / kMoreUnits;
    for (int k = 0; k < predicted - 1; ++k, ++launches) CPFEM_HIP(hipGraphLaunch(c->g_more, s));
    CPFEM_HIP(read_nrdev(c));
    const NrDev& h = *c->nrhost;
    while (!h.updated) {
        CPFEM_HIP(hipGraphLaunch(c->g_more, s));
        ++launches;
        CPFEM_HIP(read_nrdev(c));
    }
    CPFEM_HIP(hipEventRecord(c->ev[1], s));
    CPFEM_HIP(hipEventSynchronize(c->ev[1]));
    c->rng_epoch += 1;
    c->assembled = !h.fatal;
    c->vals32_fresh = c->precision == 2 && !h.fatal;
    cpfem_stats& S = c->stats;
    S.graph_launches += launches;
    S.t_graph_ms += elapsed(c->ev[0], c->ev[1]);
    S.sumcpiter = h.sums[0] / (double)c->ngps;
    S.flops_const += (int64_t)llround(h.sums[1]);
    S.transc_const += (int64_t)llround(h.sums[2]);
    S.gp_updates += c->ngps;
    S.pts_987 += (int64_t)llround(h.sums[3]);
    S.pts_987_last = (int32_t)llround(h.sums[3]);
    S.pts_continued += h.report[0];
    S.pts_wide += h.report[1];
    if (h.report[2] > 0) return CPFEM_E_WIDTH;
    const int ecode = h.flags[1];
    if (ecode > S.err) S.err = ecode;
    if (h.fatal) return ecode;
    S.sfc = sqrt(h.suv[0]);
    S.sor_sweeps = h.sweeps;
    S.sor_sweeps_total += h.sweeps;
    S.sor_error = h.rel;
    S.sor_omega = omega_used;
    S.sor_wpb = c->sor_wpb;
    S.sweep_launches += (int64_t)h.sweeps * c->A.ncolor;
    c->last_batches = h.sweeps / kSweepBatch;
    int code = 0;
    if (!(h.rel <= 10.0 * c->ctl.sor_tol)) code = 904;
    if (code > S.err) S.err = code;
    if (code == 0) refine_omega(c, h.final_rate);
    S.sdx = h.sdx;
    S.alp2 = h.alp2;
    *sdx = h.sdx;
    return code ? code : ecode;
}

extern "C" int cpfem_set_graph_mode(cpfem_ctx* c, int32_t on) {
    if (!c) return CPFEM_E_ARG;
    if (on && (c->ctl.sor_kind != 1)) return CPFEM_E_ARG;
    c->graph_mode = on != 0;
    return 0;
}

extern "C" int cpfem_nriter(cpfem_ctx* c, int32_t inc, cpfem_stats* st) {
    if (!c) return CPFEM_E_ARG;
    if (inc < 1 || inc > c->ninc1) return CPFEM_E_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    memset(&c->stats, 0, sizeof(c->stats));
    int nit = 0, code = 0;
    bool more = true;
    while (more) {
        ++nit;
        if (c->graph_mode && c->ctl.sor_kind == 1 && !c->omega_auto) {
            // the first solve of an automatic-omega context tunes on the host path
            double sdx = 0.0;
            code = graph_iteration(c, inc, &sdx);
            if (code < 0 || (code > 500 && code != 987 && code != 910 && code != 904)) break;
            if (!std::isfinite(c->stats.sor_error)) break;   // no update was applied
            more = (sdx >= c->ctl.toler) && (nit < c->ctl.maxit);
            continue;
        }
        double sumcp = 0.0, sfc = 0.0, sdx = 0.0, alp2 = 0.0;
        int tflag = 0, sweeps = 0;
        double serr = 0.0;
        // 987 (slip Newton), 910 (singular slip matrix) and 904 (SOR not converged)
        // are reported and the iteration goes on; the caller decides whether they
        // are fatal, as the reference's FAFERR does.  Other codes stop the loop.
        code = elvals_impl(c, inc, &sumcp, &tflag);
        if (code < 0 || (code > 500 && code != 987 && code != 910)) break;
        code = suvals_impl(c, inc, &sfc);
        if (code) break;
        code = facsls_impl(c, &sweeps, &serr);
        if (code < 0 || (code > 500 && code != 904)) break;
        // a solution with NaN or Inf entries is not applied: NRITER ends with 904
        if (!std::isfinite(serr)) break;
        code = nr_update_impl(c, &sdx, &alp2);
        if (code) break;
        c->stats.sumcpiter = sumcp / (double)c->ngps;
        more = (sdx >= c->ctl.toler) && (nit < c->ctl.maxit);
    }
    c->stats.nit = nit;
    code = std::max(code, c->stats.err);
    if (code == 0 || code == 987 || code == 910 || code == 904) {
        hipLaunchKernelGGL(k_cumslip, dim3(nblk(c->ngps)), dim3(kBlock), 0, c->stream, c->ngps, c->nss, c->cumslip,
                           (const double*)c->srate, c->tiem[inc - 1]);
        CPFEM_HIP(hipStreamSynchronize(c->stream));
    }
    c->stats.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (code > c->stats.err) c->stats.err = code;
    if (st) *st = c->stats;
    return code;
}

extern "C" int cpfem_cumslip(cpfem_ctx* c, int32_t inc) {
    if (!c || inc < 1 || inc > c->ninc1) return CPFEM_E_ARG;
    hipLaunchKernelGGL(k_cumslip, dim3(nblk(c->ngps)), dim3(kBlock), 0, c->stream, c->ngps, c->nss, c->cumslip,
                       (const double*)c->srate, c->tiem[inc - 1]);
    CPFEM_HIP(hipGetLastError());
    return 0;
}

extern "C" int cpfem_power(cpfem_ctx* c, double* fdx) {
    if (!c || !fdx) return CPFEM_E_ARG;
    const int D = 3 * c->nnod;
    const int nb = std::min(nblk(D), c->red.nblocks);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kBlock), 0, c->stream, (const double*)c->fc, (const double*)c->dx, D,
                       c->red.partial);
    double r[4];
    reduce_now(c, nb, r);
    *fdx = r[0];
    return 0;
}

extern "C" int cpfem_update_gptv(cpfem_ctx* c) {
    if (!c) return CPFEM_E_ARG;
    hipLaunchKernelGGL(k_commit, dim3(nblk(c->ngps)), dim3(kBlock), 0, c->stream, c->ngps, c->nsv, c->eps,
                       (const double*)c->dep, c->rt, (const double*)c->dr, c->ep, (const double*)c->de, c->st,
                       (const double*)c->ds, c->sv, (const double*)c->dsv);
    CPFEM_HIP(hipGetLastError());
    return 0;
}

// Twin tables of material 1 (DoReorient uses material 1 only,
// Fasolt3.f90:103,124-125), staged on the device at the first reorientation.
static int twin_tables(Ctx* c) {
    if (c->twin_tab) return 0;
    static_assert(sizeof(TwinTables) == 60 * 8 * 7, "layout");
    TwinTables h;
    memset(&h, 0, sizeof(h));
    for (int i = 0; i < c->mat0.nss; ++i) {
        h.ssang[i] = c->mat0.ssang[i];
        for (int k = 0; k < 3; ++k) {
            h.ab[i][k] = c->mat0.ab[i][k];
            h.an[i][k] = c->mat0.an[i][k];
        }
    }
    TwinTables* dT = nullptr;
    int* dcount = nullptr;
    if (dalloc(c, &dT, 1) || dalloc(c, &dcount, 1)) return CPFEM_E_HIP;
    CPFEM_HIP(hipMemcpy(dT, &h, sizeof(h), hipMemcpyHostToDevice));
    c->twin_tab = dT;
    c->twin_count = dcount;
    return 0;
}

extern "C" int cpfem_reorient(cpfem_ctx* c, int32_t* n) {
    if (!c) return CPFEM_E_ARG;
    int fl = 0;
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    CPFEM_HIP(hipMemcpy(&fl, c->flags, sizeof(int), hipMemcpyDeviceToHost));
    int cnt = 0;
    if (fl) {
        const int rc = twin_tables(c);
        if (rc) return rc;
        int* dcount = (int*)c->twin_count;
        hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, c->stream, dcount, (size_t)1, 0);
        hipLaunchKernelGGL(k_reorient, dim3(nblk(c->ngps)), dim3(kBlock), 0, c->stream, c->ngps, c->sv, c->twinsys,
                           c->timestwinned, (const TwinTables*)c->twin_tab, dcount);
        hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, c->stream, c->flags, (size_t)1, 0);
        CPFEM_HIP(hipGetLastError());
        CPFEM_HIP(hipMemcpyAsync(c->red.host + 6, dcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        CPFEM_HIP(hipStreamSynchronize(c->stream));
        memcpy(&cnt, c->red.host + 6, sizeof(int));
    }
    if (n) *n = cnt;
    return 0;
}

extern "C" int cpfem_volume_average(cpfem_ctx* c, double* out) {
    if (!c || !out) return CPFEM_E_ARG;
    KinArgs K;
    memset(&K, 0, sizeof(K));
    K.n = c->ngps;
    K.gp_elem = c->gp_elem;
    K.gp_local = c->gp_local;
    K.elem_nodes = c->elem_nodes;
    K.elem_lt = c->elem_lt;
    K.elem_nd = c->elem_nd;
    K.x = c->x;
    const int nb = nblk(c->ngps);
    if (nb > c->red.nblocks) return CPFEM_E_ARG;
    double sums[16];
    for (int group = 0; group < 4; ++group) {
        hipLaunchKernelGGL(k_vol_sums, dim3(nb), dim3(kBlock), 0, c->stream, K, (const double*)c->st,
                           (const double*)c->ep, (const double*)c->eps, group, c->red.partial, 0);
        CPFEM_HIP(hipGetLastError());
        reduce_now(c, nb, sums + 4 * group);
    }
    const double vol = sums[0];
    out[0] = vol;
    for (int i = 0; i < 6; ++i) out[1 + i] = vol > 0.0 ? sums[2 + i] / vol : 0.0;
    for (int i = 0; i < 6; ++i) out[7 + i] = vol > 0.0 ? sums[8 + i] / vol : 0.0;
    out[13] = vol > 0.0 ? sums[1] / vol : 0.0;
    return 0;
}

hipError_t cpfem::kin_transpose(Ctx* c, const double* x, const double* cot, double* out) {
    KinArgs K;
    memset(&K, 0, sizeof(K));
    K.n = c->ngps;
    K.gp_elem = c->gp_elem;
    K.gp_local = c->gp_local;
    K.elem_nodes = c->elem_nodes;
    K.elem_lt = c->elem_lt;
    K.elem_nd = c->elem_nd;
    K.x = x;
    const size_t D = 3 * (size_t)c->nnod;
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(D)), dim3(kBlock), 0, c->stream, out, D, 0.0);
    const int ncol_el = ((int)c->ekstart.size() - 1) / kKinds;
    for (int k = 0; k < ncol_el; ++k) {
        const int b0 = c->ekstart[kKinds * k + kElem8], cnt = c->ekstart[kKinds * k + kElem8 + 1] - b0;
        if (cnt == 0) continue;
        hipLaunchKernelGGL(k_kin_transpose8, dim3(nblk(cnt)), dim3(kBlock), 0, c->stream, K, cnt,
                           (const int*)(c->elem_by_color + b0), (const int*)c->elem_gps, cot, out);
    }
    return hipGetLastError();
}

hipError_t cpfem::kin_apply(Ctx* c, const double* x, const double* dx, double* de, double* r) {
    KinArgs K;
    memset(&K, 0, sizeof(K));
    K.n = c->ngps;
    K.gp_elem = c->gp_elem;
    K.gp_local = c->gp_local;
    K.elem_nodes = c->elem_nodes;
    K.elem_lt = c->elem_lt;
    K.elem_nd = c->elem_nd;
    K.x = x;
    K.dx = dx;
    K.smale = c->ctl.smale;
    K.smalr = c->ctl.smalr;
    K.de = de;
    K.r = r;
    K.kerr = c->sched.ck_kerr;
    hipLaunchKernelGGL(k_kin<8>, dim3(nblk(c->ngps)), dim3(kBlock), 0, c->stream, K);
    return hipGetLastError();
}

int cpfem::solve_with(Ctx* c, const double* vals, const double* rhs, int* sweeps, double* err, bool warm) {
    if (c->precision != 0) return CPFEM_E_ARG;
    CPFEM_HIP(hipMemcpyAsync(c->A.vals, vals, sizeof(double) * c->A.nvals(), hipMemcpyDeviceToDevice, c->stream));
    CPFEM_HIP(hipMemcpyAsync(c->rhs, rhs, sizeof(double) * 3 * (size_t)c->nnod, hipMemcpyDeviceToDevice, c->stream));
    c->assembled = true;
    return facsls_impl(c, sweeps, err, warm);
}

hipError_t cpfem::point_weights(Ctx* c, double* w) {
    KinArgs K;
    memset(&K, 0, sizeof(K));
    K.n = c->ngps;
    K.gp_elem = c->gp_elem;
    K.gp_local = c->gp_local;
    K.elem_nodes = c->elem_nodes;
    K.elem_lt = c->elem_lt;
    K.elem_nd = c->elem_nd;
    K.x = c->x;
    hipLaunchKernelGGL(k_point_weights, dim3(nblk(c->ngps)), dim3(kBlock), 0, c->stream, K, w);
    return hipGetLastError();
}

extern "C" int cpfem_twin_stream(cpfem_ctx* c, uint64_t next_epoch) {
    if (c)
        c->rng_epoch = next_epoch - 1;   // ELVALS advances it before use
    else
        cpfem::set_batch_epoch(next_epoch);
    return 0;
}

extern "C" int cpfem_set_tflag(cpfem_ctx* c, int32_t tflag) {
    if (!c) return CPFEM_E_ARG;
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, c->stream, c->flags, (size_t)1, tflag ? 1 : 0);
    CPFEM_HIP(hipGetLastError());
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int cpfem_increment(cpfem_ctx* c, int32_t inc, cpfem_stats* st) {
    if (!c) return CPFEM_E_ARG;
    int rc = cpfem_update_node(c);
    if (rc) return rc;
    rc = cpfem_nriter(c, inc, st);
    if (rc < 0 || (rc > 500 && rc != 987 && rc != 910 && rc != 904)) return rc;
    int r2 = cpfem_update_gptv(c);
    if (r2) return r2;
    r2 = cpfem_reorient(c, nullptr);
    if (r2) return r2;
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    return rc;
}

extern "C" int cpfem_get_system(cpfem_ctx* c, double* rhs, double* ddx) {
    if (!c) return CPFEM_E_ARG;
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    const size_t D = 3 * (size_t)c->nnod;
    if (rhs) CPFEM_HIP(hipMemcpy(rhs, c->rhs, D * 8, hipMemcpyDeviceToHost));
    if (ddx) CPFEM_HIP(hipMemcpy(ddx, c->ddx, D * 8, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int cpfem_get_matrix(cpfem_ctx* c, int64_t* nnz, int32_t* rows, int32_t* cols, double* vals) {
    if (!c || !nnz) return CPFEM_E_ARG;
    CPFEM_HIP(hipStreamSynchronize(c->stream));
    const int P = c->A.P, ng = c->A.ngroups;
    const int nslot = P * kRowLanes;
    std::vector<int> cl((size_t)ng * P * 64);
    CPFEM_HIP(hipMemcpy(cl.data(), c->A.cols, cl.size() * 4, hipMemcpyDeviceToHost));
    // real blocks: slot 0 plus the slots whose column differs from the row's node
    int64_t nb = 0;
    for (int q = 0; q < ng * kGroupRows; ++q) {
        const int nd = cl[col_index(q, P, 0)];
        if (nd < 0) continue;
        for (int s = 0; s < nslot; ++s)
            if (s == 0 || cl[col_index(q, P, s)] != nd) ++nb;
    }
    if (!rows || !cols || !vals) {
        *nnz = nb * 9;
        return 0;
    }
    if (*nnz < nb * 9) return CPFEM_E_ARG;
    const size_t nv = c->A.nvals();
    std::vector<double> v(nv);
    if (c->precision == 1) {
        std::vector<float> f(nv);
        CPFEM_HIP(hipMemcpy(f.data(), c->A.vals, nv * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nv; ++i) v[i] = f[i];
    } else {
        CPFEM_HIP(hipMemcpy(v.data(), c->A.vals, nv * 8, hipMemcpyDeviceToHost));
    }
    int64_t k = 0;
    for (int q = 0; q < ng * kGroupRows; ++q) {
        const int nd = cl[col_index(q, P, 0)];
        if (nd < 0) continue;
        for (int s = 0; s < nslot; ++s) {
            const int cn = cl[col_index(q, P, s)];
            if (s != 0 && cn == nd) continue;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    rows[k] = 3 * nd + a;
                    cols[k] = 3 * cn + b;
                    vals[k] = v[val_index(q, P, s, 3 * a + b)];
                    ++k;
                }
        }
    }
    *nnz = k;
    return 0;
}

extern "C" int cpfem_get_stats(cpfem_ctx* c, cpfem_stats* st) {
    if (!c || !st) return CPFEM_E_ARG;
    *st = c->stats;
    return 0;
}
