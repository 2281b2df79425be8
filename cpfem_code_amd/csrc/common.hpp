// Shared host helpers of the HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../../include/cpfem.h"

#define CPFEM_HIP(call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            fprintf(stderr, "cpfem: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_),   \
                    __FILE__, __LINE__);                                                       \
            return CPFEM_E_HIP;                                                                \
        }                                                                                      \
    } while (0)

namespace cpfem {

// RAII device buffer used by the host-pointer entry points.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace cpfem
