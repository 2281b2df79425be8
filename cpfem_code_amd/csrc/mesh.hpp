// Device-resident FASOLT3 problem: layout of the context and kernel launchers.
//
// HBM layout (all point arrays field-major, stride ngps; nodal vectors in the
// reference's DOF order 3*(node-1)+k):
//   x, dx, fc, rhs, ddx            [3*nnod]  FP64
//   st, ep, ds, de, t              [6][ngps] FP64
//   rt, dr                         [3][ngps]
//   sv, dsv                        [3+nss][ngps]
//   srate, cumslip                 [nss][ngps]
//   chi                            [36][ngps] tangent of the last pass
//   matrix (block ELL over nodes, row groups; see below)
// Matrix rows are the nodes permuted colour-major.  Each colour's rows are cut
// into groups of kGroupRows rows, one group per 64-lane wave, kRowLanes lanes
// per row (lane = l*kGroupRows + i for row i of the group, lane 63 idle).
// Slot s of a row (slot 0 = the diagonal 3x3 block, then the neighbour nodes in
// ascending order, padded with the row's own node and zero values) belongs to
// lane l = s % kRowLanes at depth p = s / kRowLanes, and a group stores
//   vals[group][p][e][lane]  (e = 3*a + b of the 3x3 block)
//   cols[group][p][lane]
// so one load instruction of a wave reads 64 consecutive values: a sweep
// streams each group's P*9*64 values as whole contiguous segments.  Padded rows
// of a colour's last group have column -1 and zero values.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/cpfem.h"
#include "constitutive_launch.hpp"

namespace cpfem {

struct Reducer {
    double* partial = nullptr;   // [nblocks][4]
    double* result = nullptr;    // [4] device
    double* host = nullptr;      // [4] pinned host
    int nblocks = 0;
};

constexpr int kMaxNodes = 20;                    // nodes per element (20-node brick)
// Element kinds: 8-node bricks, the tabulated quadratic elements (20-node
// bricks, 10-node tetrahedra, 15-node pentahedra) and two-node ties.
enum ElemKind { kElem8 = 0, kElem20 = 1, kTie = 2, kElem10 = 3, kElem15 = 4 };
constexpr int kKinds = 5;

constexpr int kRowLanes = 9;                     // lanes per matrix row
constexpr int kGroupRows = 64 / kRowLanes;       // rows per group (= per wave)

// Index of entry e of slot s of the row at group position q (q = group *
// kGroupRows + row within the group), for P slots per lane.
__host__ __device__ inline size_t val_index(int q, int P, int s, int e) {
    return (size_t)(q / kGroupRows) * P * 576 + (size_t)((s / kRowLanes) * 9 + e) * 64 +
           (s % kRowLanes) * kGroupRows + q % kGroupRows;
}
__host__ __device__ inline size_t col_index(int q, int P, int s) {
    return (size_t)(q / kGroupRows) * P * 64 + (size_t)(s / kRowLanes) * 64 + (s % kRowLanes) * kGroupRows +
           q % kGroupRows;
}

struct Matrix {
    int R = 0;            // rows (= nodes)
    int S = 0;            // slots per row (max neighbours incl. self)
    int P = 0;            // slots per lane, ceil(S / kRowLanes)
    int ncolor = 0;
    int ngroups = 0;
    std::vector<int> cstart;   // [ncolor+1] row range of each colour
    std::vector<int> gstart;   // [ncolor+1] group range of each colour
    int* cols = nullptr;       // [ngroups][P][64] neighbour node of each slot
    void* vals = nullptr;      // [ngroups][P][9][64] double or float
    float* vals32 = nullptr;   // precision 2: FP32 copy of the FP64 vals that the sweeps read
    int* diag_off = nullptr;   // [nnod] index of entry 0 of each node's diagonal block
    int* scatter = nullptr;    // per element nd*nd entries from scat_start[e]: index of entry 0 of the
                               // block of node pair (a, b) at a*nd + b
    int* scat_start = nullptr; // [nel]
    // gather assembly (meshes of 8-node bricks): the element blocks are stored
    // element by element, then every matrix block and node force sums its
    // contributions from lists (in element-colour order)
    bool gather = false;
    void* eblk = nullptr;      // [nel][8][9][8] double or float: entry q of node pair (a, b) at a*72 + q*8 + b
    double* efc = nullptr;     // [nel][8][3] forces of each element node
    int* g_order = nullptr;    // [ngroups*P] segment of each wave of the block gather
    int* g_start = nullptr;    // [ngroups*P] first list entry of each segment
    int* g_cnt = nullptr;      // [ngroups*P] list entries per lane of each segment
    int* g_list = nullptr;     // element-block indices, 64 lanes per entry
    int* f_start = nullptr;    // [ceil(nnod/64)] the same for the node forces
    int* f_cnt = nullptr;
    int* f_list = nullptr;     // efc indices
    size_t nnzb = 0;           // true (unpadded) blocks
    size_t nvals() const { return (size_t)ngroups * P * 576; }
};

// Device record of the Newton iterations: the solve's stopping-rule state
// (every decision of the residual-rule SOR is taken on the device, so a
// solve runs the same kernels whether the host replays them batch by batch
// or a whole Newton iteration is one graph launch), the Newton update's
// norms and damping, and the report values the host reads after an
// iteration.  Written by the kernels of mesh.hip; the host reads it back.
struct NrDev {
    int stop;          // nonzero: the sweeps and the checks are skipped
    int sweeps;        // sweeps of the solve so far
    int left;          // batches left before the next residual check
    int check_now;     // the batch just run ends a stretch: refine, residual, check
    int cycle;         // the next batch starts a stretch (mixed: r -> FP32, d = 0)
    int prev_sweeps;   // sweeps at the previous check
    int updated;       // the iteration's Newton update has run (or is skipped)
    int fatal;         // the iteration's ELVALS code stops NRITER: no solve, no update
    int inc;           // increment of the iteration (load tables)
    int flags[2];      // tflag, largest point code of the ELVALS
    int report[3];     // constitutive: continued points, wider tiers, too wide
    uint64_t epoch;    // twin-draw stream epoch of the ELVALS
    double dt;         // time step of the increment
    double omega;      // relaxation factor the sweeps read
    double rel, prev_rel, final_rate;
    double sdx, alp2;
    double sums[4];    // ELVALS point sums: sumcpiter, flops, transcendentals, 987 points
    double suv[4];     // SUVALS sums
    double nrm[4];     // |dx|^2, |ddx|^2
};

struct Ctx {
    int nlmnt = 0, nnod = 0, ngps = 0, nsv = 0, nss = 0, ninc1 = 0, nmat = 0;   // nss: largest of the materials
    int precision = 0;           // cpfem_control::precision (0 FP64, 1 FP32 matrix, 2 FP32 sweeps on an FP64 matrix)
    int knss = 0;                // slip-system count the constitutive kernels are compiled for
    cpfem_control ctl;
    cpfem_material mat0;         // host copy of material 1
    void* dmat = nullptr;
    hipStream_t stream = nullptr;
    // mesh
    int* elem_nodes = nullptr;   // [nel][kMaxNodes] 0-based
    int* elem_gps = nullptr;     // [nel][8] 0-based (unused slots -1)
    int* elem_nd = nullptr;      // [nel] nodes of each element
    int* elem_lt = nullptr;      // [nel]
    int* elem_le = nullptr;      // [nel]
    int* elem_np = nullptr;      // [nel]
    int* gp_elem = nullptr;      // [ngps]
    int* gp_local = nullptr;     // [ngps]
    int* ib = nullptr;           // [nnod]
    // elements by colour and, within a colour, by kind (ElemKind):
    // kind q of colour k is elem_by_color[ekstart[kKinds*k+q] .. ekstart[kKinds*k+q+1])
    std::vector<int> ekstart;
    std::vector<int> corder;     // order in which a sweep relaxes the node colours (a permutation of 0..ncolor-1)
    int npts8 = 0, npts20 = 0;   // points of 8-node and of 20-node elements
    int npts10 = 0, npts15 = 0;  // points of 10-node and of 15-node elements
    int* elem_by_color = nullptr;
    double* x_spec = nullptr;    // [ninc1][99][15]
    double* f_spec = nullptr;    // [ninc1][99][3]
    std::vector<double> tiem;
    // nodal vectors
    double *x = nullptr, *dx = nullptr, *fc = nullptr, *rhs = nullptr, *ddx = nullptr;
    double *xold = nullptr;      // SOR scratch (FP64) / residual
    float *rhs32 = nullptr, *del32 = nullptr;   // mixed-precision inner solve
    // point state
    double *eps = nullptr, *dep = nullptr, *st = nullptr, *ep = nullptr, *ds = nullptr, *de = nullptr;
    double *rt = nullptr, *dr = nullptr, *sv = nullptr, *dsv = nullptr, *srate = nullptr, *cumslip = nullptr;
    double *t = nullptr, *chi = nullptr, *deff = nullptr, *erepmd = nullptr;
    int *twinsys = nullptr, *timestwinned = nullptr, *gperr = nullptr;
    long long *flops = nullptr, *transc = nullptr;
    int* flags = nullptr;        // [0] tflag, [1] errmax
    uint64_t rng_epoch = 0;
    void* twin_tab = nullptr;    // TwinTables of material 1 (first reorientation)
    void* twin_count = nullptr;  // [1] reoriented points
    Matrix A;
    Reducer red;
    ConstSchedule sched;
    hipGraphExec_t graphs[3] = {nullptr, nullptr, nullptr};  // sweep batches: FP64 update test, FP64, FP32
    hipGraph_t graph_src[3] = {nullptr, nullptr, nullptr};   // their captured graphs
    bool use_graphs = true;      // CPFEM_SOR_GRAPHS=0: launch the batches kernel by kernel
    // residual-rule solves: the batch of sweeps with its stopping-rule kernels
    // (one unit) as a graph; graph mode: each Newton iteration as one graph
    // (ELVALS, SUVALS, solve start, it_units units, gated update) and its
    // continuation (more_units units, gated update)
    hipGraphExec_t g_unit = nullptr, g_it = nullptr, g_more = nullptr;
    hipGraph_t g_unit_src = nullptr, g_it_src = nullptr, g_more_src = nullptr;
    int it_units = 0, more_units = 0;
    int g_key[4] = {0, 0, 0, 0};  // launch shapes the iteration graphs were captured with
    bool graph_mode = false;     // cpfem_set_graph_mode
    int last_batches = 0;        // batches of the last residual-rule solve
    NrDev* nrdev = nullptr;      // device record
    NrDev* nrhost = nullptr;     // pinned host copy
    int sor_wpb = 4;             // waves per workgroup of the colour sweeps (cpfem_autotune)
    int asm_epb = 4;             // elements (waves) per workgroup of the 8-node assembly (cpfem_autotune)
    bool omega_auto = false;     // sor_omega <= 0 at creation: tuned at the first solve
    // sor_omega <= 0 at creation: after the first solve's tuning, omega follows
    // the convergence rates of the solves (coordinate search, see refine_omega)
    bool om_adapt = false;
    bool om_trial = false;       // the last solve ran at a trial factor
    double om_best = 0.0, om_rate = 0.0, om_step = 0.01;
    int om_dir = 1;
    int om_wins = 0;             // successful trials in a row (two double the step)
    bool assembled = false;      // A and rhs hold a system (set by SUVALS)
    bool vals32_fresh = false;   // precision 2: A.vals32 holds the rounded copy of the current A.vals
    // timing and counters of the last calls
    static constexpr int kEvents = 10;   // [7], [8]: around each sweep batch
    hipEvent_t ev[kEvents] = {};
    cpfem_stats stats;
    std::vector<void*> allocs;
    void* calib = nullptr;       // replay and adjoint state (calib.hip)
};

// Integration weight of every point at the current coordinates, w[ngps]
// (mesh.hip), enqueued on the context's stream.
hipError_t point_weights(Ctx* c, double* w);
// Transpose of the 8-node-brick point kinematics at coordinates x: nodal
// vector out (zeroed first) from the cotangents cot[9][ngps] (de, r).
hipError_t kin_transpose(Ctx* c, const double* x, const double* cot, double* out);
// Point kinematics (DSTRNS) of the nodal vector dx at coordinates x (8-node bricks).
hipError_t kin_apply(Ctx* c, const double* x, const double* dx, double* de, double* r);
// FACSLS on the FP64 matrix values vals (layout of A.vals) and right-hand side
// rhs, with the context's solver settings, from zero or (warm) from the ddx in
// place; the solution is left in ddx and A.vals, rhs are overwritten.
int solve_with(Ctx* c, const double* vals, const double* rhs, int* sweeps, double* err, bool warm);
// Releases the replay and adjoint state (calib.hip).
void calib_free(Ctx* c);

}  // namespace cpfem
