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
//   matrix (block ELL over nodes)  vals[slot][9][R] (FP64 or FP32), cols[slot][R]
// Matrix rows are the nodes permuted colour-major: rows of one colour are
// contiguous, so a colour sweep reads vals/cols fully coalesced.  Slot 0 of a
// row is its diagonal 3x3 block; padded slots point at the row's own node with
// zero values.
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

struct Matrix {
    int R = 0;            // rows (= nodes)
    int S = 0;            // slots per row (max neighbours incl. self)
    int ncolor = 0;
    std::vector<int> cstart;   // [ncolor+1] row range of each colour
    int* row_node = nullptr;   // [R] permuted row -> node (0-based)
    int* node_row = nullptr;   // [R] node -> row
    int* cols = nullptr;       // [S][R] neighbour node of each slot
    void* vals = nullptr;      // [S][9][R] double or float
    double* vals64 = nullptr;  // FP64 copy for residuals in mixed mode (null in FP64 mode)
    int* scatter = nullptr;    // [nel][64] slot*9*R + row of each element node pair
    size_t nnzb = 0;           // true (unpadded) blocks
};

struct Ctx {
    int nlmnt = 0, nnod = 0, ngps = 0, nsv = 0, nss = 0, ninc1 = 0, nmat = 0;
    int precision = 0;
    cpfem_control ctl;
    void* dmat = nullptr;
    hipStream_t stream = nullptr;
    // mesh
    int* elem_nodes = nullptr;   // [nel][8] 0-based
    int* elem_gps = nullptr;     // [nel][8] 0-based (unused slots -1)
    int* elem_lt = nullptr;      // [nel]
    int* elem_le = nullptr;      // [nel]
    int* elem_np = nullptr;      // [nel]
    int* gp_elem = nullptr;      // [ngps]
    int* gp_local = nullptr;     // [ngps]
    int* ib = nullptr;           // [nnod]
    std::vector<int> ecstart;    // element colour ranges
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
    Matrix A;
    Reducer red;
    ConstSchedule sched;
    hipGraphExec_t graphs[3] = {nullptr, nullptr, nullptr};  // sweep batches: FP64 update test, FP64, FP32
    // timing and counters of the last calls
    hipEvent_t ev[8];
    cpfem_stats stats;
    std::vector<void*> allocs;
};

}  // namespace cpfem
