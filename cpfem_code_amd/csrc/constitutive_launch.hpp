// Host-side view of the constitutive kernels: SoA point arrays and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cpfem.h"

namespace cpfem {

// Device pointers of the per-point state, field-major with stride n.
struct PointArrays {
    int n;
    int gp_offset;            // index of point 0 in the mesh (twin draw key)
    const int* iel;           // [n]
    const double* de;         // [6][n]
    const double* r;          // [3][n]
    const double* st;         // [6][n]
    double* ds;               // [6][n]  guess in, increment out
    const double* sv;         // [3+nss][n]
    double* dsv;              // [3+nss][n]
    double* t;                // [6][n]
    double* chi;              // [36][n] Fortran chi(6,6) order
    double* deff;             // [n]
    double* erepmd;           // [n]
    double* srate;            // [nss][n]
    int* twinsys;             // [n]
    const int* timestwinned;  // [n]
    long long* flops;         // [n] or null
    long long* transc;        // [n] or null
    int* err;                 // [n] or null
    int* tflag;               // scalar
    int* errmax;              // scalar
    uint64_t rng_epoch;
};

struct DevControl;
DevControl make_control(const cpfem_control& c);
size_t material_bytes(int nss);
void build_material(int nss, void* host, const cpfem_material& m, double toler);
hipError_t launch_dstrss(int nss, const void* dmat, const DevControl& C, double dt, const PointArrays& A,
                         hipStream_t stream);

}  // namespace cpfem
