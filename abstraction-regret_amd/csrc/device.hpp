// Device helpers shared by the kernel translation units (not by the host-only layout code,
// which also builds with a plain C++ compiler for the host-logic checks).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gfsm {

// Block-wide copy of `vecs` 16-byte vectors from global memory into LDS (no barrier). Each
// thread keeps 8 loads in flight before its LDS writes: a loop that waits for every load in
// turn pays one memory round trip per vector per thread. Loads and writes are unconditional:
// indexes past the end are clamped to the last vector, which is then written again with its
// own value (a guarded load gets sunk into its branch and waited for there). vecs >= 1.
__device__ __forceinline__ void copy_to_lds(const uint4 *__restrict__ src, uint4 *dst,
                                            uint32_t vecs) {
    constexpr int kBatch = 8;
    for (uint32_t k0 = threadIdx.x; k0 < vecs; k0 += kBatch * blockDim.x) {
        uint4 v[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const uint32_t k = k0 + j * blockDim.x;
            v[j] = src[k < vecs ? k : vecs - 1];
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const uint32_t k = k0 + j * blockDim.x;
            dst[k < vecs ? k : vecs - 1] = v[j];
        }
    }
}

}  // namespace gfsm
