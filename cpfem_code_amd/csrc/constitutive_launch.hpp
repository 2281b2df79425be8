// Host-side view of the constitutive kernels: the multi-pass schedule.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cpfem.h"
#include "constitutive.hpp"

namespace cpfem {

// Widest active set of the compiled kernels: 12 for the kernels of 12 and 18
// systems (twin systems never enter the active set), 24 for the generic ones.
constexpr int kConstMaxWidth = 24;

// Device scratch of the schedule: checkpoints, per-point work arrays and the
// point lists.
struct ConstSchedule {
    double* ck_s = nullptr;    // [6][n]
    int* ck_iter = nullptr;    // [n]
    int* ck_kerr = nullptr;    // [n]
    int* ck_inner = nullptr;   // [n] slip-Newton iterations so far (ordering key)
    int* sorted = nullptr;     // [n] a point list in resume order
    int* prev_bin = nullptr;   // [n] sort bin of each point's previous-update work (mesh mode)
    int* bins = nullptr;       // [1024] counting-sort bins
    int* list_a = nullptr;     // [n] unfinished after the first pass
    int* list_b = nullptr;     // [n] spare
    int* list_ovf = nullptr;   // [n] active set wider than the first kernel
    double* el = nullptr;      // [21][n] compliance in specimen axes of the current pass
    double* gp = nullptr;      // [kConstMaxWidth][n] slips of the last hardening update
    int* counts = nullptr;     // [8] device: list counts [0..3], work-queue counter [4]
    int* host_counts = nullptr;// [4] pinned
    hipEvent_t ev[5] = {};     // around the passes of the last run (timing)
    int first_iters = 8;       // outer iterations of the first pass over every point (cpfem_autotune)
    int wide_iters = 0;        // outer iterations (counted from the point's start) before the 8-wide tier's
                               // unfinished points are compacted and resumed (0: first_iters)
    double pass_ms[4] = {};    // device time of the last run: pass 1, pass 2, 8-wide, 12(+)-wide tiers
    bool queue = true;         // persistent launches draining the point lists (CPFEM_CONST_QUEUE=0: one
                               // workgroup per 64 list entries, drained by the hardware dispatcher)
};

DevControl make_control(const cpfem_control& c);
size_t material_bytes(int nss);
void build_material(int nss, void* host, const cpfem_material& m, double toler);
hipError_t alloc_schedule(ConstSchedule& S, int n);
void free_schedule(ConstSchedule& S);
// Runs the constitutive update of all A.n points; counts[0] = points continued
// after the first pass, counts[1] = points needing the full-width kernel.
hipError_t run_constitutive(int nss, const void* dmat, const DevControl& C, double dt, const ConstArgs& A,
                            ConstSchedule& S, hipStream_t stream, int counts[4]);
// The same with every decision on the device (capturable in a graph); report
// [0] continued points, [1] points of the wider tiers, [2] points too wide.
hipError_t run_constitutive_dev(int nss, const void* dmat, const DevControl& C, double dt, const ConstArgs& A,
                                ConstSchedule& S, hipStream_t stream, int* report);
// Stream epoch of the twin draws of the next cpfem_dstrss_batch call.
void set_batch_epoch(uint64_t next);

}  // namespace cpfem
