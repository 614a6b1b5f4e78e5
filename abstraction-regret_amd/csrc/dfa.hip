// DFA walk kernels for gfx950.
//
// Semantics: `src/gpufsm/reference.py:68-81` (simulate_dfa), latch-first-match. The
// upstream CUDA kernel is `native/dfa.cu:9-29` (one thread per string, int32 symbols,
// separate int8 accept array: two dependent loads per byte).
//
// Here one lane still owns one input (the per-input chain is sequential), but
//   * the input is read as bytes with 16-byte vector loads (the upstream path reads int32
//     symbols, four times the traffic);
//   * each table entry holds the successor with the successor's accept bit folded into
//     its top bit, so a byte costs exactly one dependent load;
//   * tables of <= 32767 states are uint16 (a 4096-state table is 2 MiB and stays in each
//     XCD's 4 MB L2); tables of <= 320 states are staged in LDS by `dfa_lds`.
#include "internal.hpp"

namespace gfsm {

template <typename T>
struct EntryBits;
template <>
struct EntryBits<uint16_t> {
    static constexpr uint32_t acc = 0x8000u, state = 0x7fffu;
};
template <>
struct EntryBits<uint32_t> {
    static constexpr uint32_t acc = 0x80000000u, state = 0x7fffffffu;
};

// Walk `len` bytes from `cur`; return the 1-based position of the first accepting state
// reached, or 0. The caller has already handled an accepting start state.
template <typename T>
__device__ __forceinline__ int64_t dfa_walk(const T *__restrict__ tab, uint32_t cur,
                                            const uint8_t *__restrict__ p, int64_t len) {
    constexpr uint32_t ACC = EntryBits<T>::acc, ST = EntryBits<T>::state;
    int64_t i = 0;
    const int64_t head = ((16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15);
    const int64_t h = head < len ? head : len;
    for (; i < h; ++i) {
        const uint32_t e = tab[cur * kAlphabet + p[i]];
        cur = e & ST;
        if (e & ACC) return i + 1;
    }
    for (; i + 16 <= len; i += 16) {
        const uint4 v = *reinterpret_cast<const uint4 *>(p + i);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t b = (w[k >> 2] >> ((k & 3) * 8)) & 0xffu;
            const uint32_t e = tab[cur * kAlphabet + b];
            cur = e & ST;
            if (e & ACC) return i + k + 1;
        }
    }
    for (; i < len; ++i) {
        const uint32_t e = tab[cur * kAlphabet + p[i]];
        cur = e & ST;
        if (e & ACC) return i + 1;
    }
    return 0;
}

__global__ void fill_result_kernel(int32_t *flags, int32_t *lens, int32_t count, int32_t flag,
                                   int32_t len) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        flags[i] = flag;
        lens[i] = len;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void dfa_gather_kernel(const T *__restrict__ tab,
                                                         uint32_t start,
                                                         const uint8_t *__restrict__ data,
                                                         const int64_t *__restrict__ off,
                                                         int32_t count, int32_t *flags,
                                                         int32_t *lens) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t b = off[i];
    const int64_t m = dfa_walk<T>(tab, start, data + b, off[i + 1] - b);
    flags[i] = m > 0;
    lens[i] = static_cast<int32_t>(m);
}

// Table staged in LDS once per block; blocks then stride over the inputs.
__global__ __launch_bounds__(256) void dfa_lds_kernel(const uint16_t *__restrict__ tab,
                                                      int32_t entries, uint32_t start,
                                                      const uint8_t *__restrict__ data,
                                                      const int64_t *__restrict__ off,
                                                      int32_t count, int32_t *flags,
                                                      int32_t *lens) {
    extern __shared__ uint4 lds_raw[];
    uint16_t *lt = reinterpret_cast<uint16_t *>(lds_raw);
    const uint4 *src = reinterpret_cast<const uint4 *>(tab);
    const int32_t vecs = entries / 8;  // entries is a multiple of 256
    for (int32_t k = threadIdx.x; k < vecs; k += blockDim.x) lds_raw[k] = src[k];
    __syncthreads();
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += gridDim.x * blockDim.x) {
        const int64_t b = off[i];
        const int64_t m = dfa_walk<uint16_t>(lt, start, data + b, off[i + 1] - b);
        flags[i] = m > 0;
        lens[i] = static_cast<int32_t>(m);
    }
}

void launch_fill_result(int32_t *flags, int32_t *lens, int32_t count, int32_t flag,
                        int32_t len, hipStream_t stream) {
    if (count <= 0) return;
    fill_result_kernel<<<(count + 255) / 256, 256, 0, stream>>>(flags, lens, count, flag, len);
    GFSM_HIP(hipGetLastError());
}

void launch_dfa_gather(const DfaDev &d, const uint8_t *data, const int64_t *off, int32_t count,
                       int32_t *flags, int32_t *lens, hipStream_t stream) {
    if (count <= 0) return;
    const int grid = (count + 255) / 256;
    if (d.wide)
        dfa_gather_kernel<uint32_t><<<grid, 256, 0, stream>>>(
            static_cast<const uint32_t *>(d.table), d.start, data, off, count, flags, lens);
    else
        dfa_gather_kernel<uint16_t><<<grid, 256, 0, stream>>>(
            static_cast<const uint16_t *>(d.table), d.start, data, off, count, flags, lens);
    GFSM_HIP(hipGetLastError());
}

size_t dfa_lds_bytes(int32_t n) { return static_cast<size_t>(n) * kAlphabet * sizeof(uint16_t); }

void launch_dfa_lds(const DfaDev &d, const uint8_t *data, const int64_t *off, int32_t count,
                    int32_t *flags, int32_t *lens, hipStream_t stream) {
    if (count <= 0) return;
    if (d.wide) throw Error(3, "dfa lds: table entries are not uint16");
    const size_t bytes = dfa_lds_bytes(d.n);
    static bool attr_set = false;
    if (!attr_set) {
        GFSM_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&dfa_lds_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    int grid = (count + 255) / 256;
    if (grid > 2048) grid = 2048;
    dfa_lds_kernel<<<grid, 256, bytes, stream>>>(static_cast<const uint16_t *>(d.table),
                                                 d.n * kAlphabet, d.start, data, off, count,
                                                 flags, lens);
    GFSM_HIP(hipGetLastError());
}

}  // namespace gfsm
