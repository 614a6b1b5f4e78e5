// NFA step kernels for gfx950 (wave64).
//
// Semantics: `src/gpufsm/reference.py:36-65` (simulate) with the epsilon closure of
// `reference.py:20-33`. Latch-first-match; an empty successor set ends the input with
// (0, 0). The start closure is computed once per automaton on the host (as the upstream
// bitmap spec does, `core/bitmap.py:84-86`); an accepting start closure never reaches
// these kernels.
//
// worklist  — general NFAs (any labels, wildcards, epsilon cycles). One wave64 per input.
//             The active set is a compacted list of state ids; a per-wave LDS bitmap dedups
//             successors with ds_or atomics and is cleared through the list (O(active) per
//             byte, never O(states)). Rows are label-sorted, and long rows carry a per-label
//             index, so a step on byte c touches only the edges labelled c or ANY. The
//             epsilon closure is a breadth-first frontier over the same list. The accept
//             test is fused into the append. Upstream counterparts: worklist_warp_kernel
//             and worklist_compact_kernel (`native/worklist_global.cu:128-310`).
//
// bitpacked — chain-structured NFAs (every non-start state has at most one non-start
//             predecessor and successor, no epsilon edges, homogeneous labels: the shape
//             of pattern unions). States are renumbered into 4096-state tiles, one 64-bit
//             word per lane, so one step for 4096 states is
//                 D = ((D & hassucc) << 1 | carry-in | (start active ? first : 0)) & M[c]
//             with the 128 KiB class-mask table M of the tile in LDS. Every tile of an
//             input runs independently and the input's result is the minimum first-accept
//             position over its tiles (global atomicMin, also used to stop tiles early).
//             Upstream counterpart: the bit-packed multistream kernels
//             (`native/bitpacked.cu:12-166`), which OR per-state masks word by word.
#include "device.hpp"
#include "internal.hpp"

#include <climits>

namespace gfsm {

// ---------------------------------------------------------------- worklist ---------------

struct WlShared {
    int32_t cnt;
};

__device__ __forceinline__ bool wl_add(uint64_t *mark, int32_t *nxt, int32_t *cnt,
                                       const uint64_t *__restrict__ accw, uint32_t t) {
    const uint64_t bit = 1ull << (t & 63u);
    const uint64_t old = atomicOr(reinterpret_cast<unsigned long long *>(&mark[t >> 6]),
                                  static_cast<unsigned long long>(bit));
    if (old & bit) return false;
    const int32_t p = atomicAdd(cnt, 1);
    nxt[p] = static_cast<int32_t>(t);
    return (accw[t >> 6] & bit) != 0;
}

__global__ __launch_bounds__(64) void nfa_worklist_kernel(NfaWorklistDev d,
                                                          const uint8_t *__restrict__ data,
                                                          const int64_t *__restrict__ off,
                                                          int32_t count, int32_t *flags,
                                                          int32_t *lens, int32_t *scratch,
                                                          int32_t lds_lists) {
    extern __shared__ uint64_t smem[];
    __shared__ int32_t cnt;
    uint64_t *mark = smem;
    int32_t *cur;
    int32_t *nxt;
    if (lds_lists) {
        cur = reinterpret_cast<int32_t *>(smem + d.nwords);
    } else {
        cur = scratch + static_cast<int64_t>(blockIdx.x) * 2 * d.n;
    }
    nxt = cur + d.n;
    const int32_t lane = threadIdx.x;
    for (int32_t w = lane; w < d.nwords; w += 64) mark[w] = 0;
    __syncthreads();

    for (int32_t in = blockIdx.x; in < count; in += gridDim.x) {
        const int64_t base = off[in];
        const int64_t len = off[in + 1] - base;
        const uint8_t *p = data + base;
        for (int32_t j = lane; j < d.k0; j += 64) cur[j] = d.start_list[j];
        int32_t k = d.k0;
        int32_t res_f = 0, res_l = 0;
        if (lane == 0) cnt = 0;
        __syncthreads();
        for (int64_t pos = 0; pos < len; ++pos) {
            const uint32_t b = p[pos];
            bool hit = false;
            for (int32_t j = lane; j < k; j += 64) {
                const int32_t s = cur[j];
                const int32_t bi = d.bigidx[s];
                if (bi >= 0) {
                    const int32_t *bo = d.bigoff + static_cast<int64_t>(bi) * kBigStride;
                    for (int32_t e = bo[b], ee = bo[b + 1]; e < ee; ++e)
                        hit |= wl_add(mark, nxt, &cnt, d.accw, d.edge[e] >> kLabelBits);
                    for (int32_t e = bo[kAnySymbol], ee = bo[kAnySymbol + 1]; e < ee; ++e)
                        hit |= wl_add(mark, nxt, &cnt, d.accw, d.edge[e] >> kLabelBits);
                } else {
                    for (int32_t e = d.row[s], ee = d.row[s + 1]; e < ee; ++e) {
                        const uint32_t v = d.edge[e];
                        const uint32_t lab = v & ((1u << kLabelBits) - 1u);
                        if (lab == b || lab == static_cast<uint32_t>(kAnySymbol))
                            hit |= wl_add(mark, nxt, &cnt, d.accw, v >> kLabelBits);
                    }
                }
            }
            __syncthreads();
            int32_t nk = cnt;
            if (d.has_eps) {
                int32_t head = 0;
                while (head < nk) {
                    __syncthreads();
                    for (int32_t j = head + lane; j < nk; j += 64) {
                        const int32_t s = nxt[j];
                        for (int32_t e = d.erow[s], ee = d.erow[s + 1]; e < ee; ++e)
                            hit |= wl_add(mark, nxt, &cnt, d.accw,
                                          static_cast<uint32_t>(d.etgt[e]));
                    }
                    __syncthreads();
                    head = nk;
                    nk = cnt;
                }
            }
            for (int32_t j = lane; j < nk; j += 64) mark[nxt[j] >> 6] = 0;
            const int any_hit = __syncthreads_or(hit ? 1 : 0);
            if (any_hit) {
                res_f = 1;
                res_l = static_cast<int32_t>(pos + 1);
                break;
            }
            if (nk == 0) break;
            int32_t *t = cur;
            cur = nxt;
            nxt = t;
            k = nk;
            if (lane == 0) cnt = 0;
            __syncthreads();
        }
        if (lane == 0) {
            flags[in] = res_f;
            lens[in] = res_l;
        }
        __syncthreads();
    }
}

static size_t wl_lds_bytes(const NfaWorklistDev &d, bool lds_lists) {
    size_t b = static_cast<size_t>(d.nwords) * 8;
    if (lds_lists) b += static_cast<size_t>(d.n) * 2 * 4;
    return b;
}

static constexpr size_t kWorklistLdsBudget = 48 * 1024;
static constexpr int64_t kWorklistScratchBudget = 256ll << 20;  // bytes of global lists

void nfa_worklist_geometry(const NfaWorklistDev &d, int32_t count, int32_t *grid,
                           int64_t *scratch_elems) {
    const bool lds_lists = wl_lds_bytes(d, true) <= kWorklistLdsBudget;
    int64_t g = count;
    if (!lds_lists) {
        const int64_t per_block = static_cast<int64_t>(d.n) * 2;
        int64_t cap = kWorklistScratchBudget / (per_block * 4);
        if (cap < 256) cap = 256;
        if (g > cap) g = cap;
        *scratch_elems = g * per_block;
    } else {
        if (g > 65536) g = 65536;
        *scratch_elems = 0;
    }
    if (g < 1) g = 1;
    *grid = static_cast<int32_t>(g);
}

void launch_nfa_worklist(const NfaWorklistDev &d, const uint8_t *data, const int64_t *off,
                         int32_t count, int32_t *flags, int32_t *lens, int32_t *scratch,
                         int32_t grid, hipStream_t stream) {
    if (count <= 0) return;
    const bool lds_lists = scratch == nullptr;
    const size_t bytes = wl_lds_bytes(d, lds_lists);
    ensure_dynamic_lds(reinterpret_cast<const void *>(&nfa_worklist_kernel), bytes);
    nfa_worklist_kernel<<<grid, 64, bytes, stream>>>(d, data, off, count, flags, lens, scratch,
                                                     lds_lists ? 1 : 0);
    GFSM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- bitpacked --------------

constexpr int kBpThreads = 1024;
constexpr size_t kBpLdsWords = 256 * kTileWords + 3 * kTileWords;

__global__ void bp_init_kernel(int32_t *best, int32_t count) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) best[i] = INT_MAX;
}

__global__ void bp_finish_kernel(const int32_t *best, int32_t count, int32_t *flags,
                                 int32_t *lens) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const int32_t v = best[i];
        flags[i] = v != INT_MAX;
        lens[i] = v != INT_MAX ? v : 0;
    }
}

// Four input bytes at 4-aligned offset `a`, never reading at or past `data_end` (the end
// of the caller's buffer) and skipping words that start at or past `stop`.
__device__ __forceinline__ uint32_t bp_load_word(const uint8_t *__restrict__ data, int64_t a,
                                                 int64_t stop, int64_t data_end) {
    if (a >= stop) return 0;
    if (a + 4 <= data_end) return *reinterpret_cast<const uint32_t *>(data + a);
    uint32_t w = 0;
    for (int q = 0; q < 4; ++q)
        if (a + q < data_end) w |= static_cast<uint32_t>(data[a + q]) << (8 * q);
    return w;
}

// One step of the chain automaton for one tile word (this lane's 64 states).
__device__ __forceinline__ uint64_t bp_step(uint64_t D, uint64_t hs, uint64_t first_if_active,
                                            uint64_t m, int lane) {
    const uint64_t x = D & hs;
    const uint32_t top = static_cast<uint32_t>(x >> 63);
    uint32_t carry = __shfl_up(top, 1, 64);
    if (lane == 0) carry = 0;
    return ((x << 1) | carry | first_if_active) & m;
}

__global__ __launch_bounds__(kBpThreads) void nfa_bitpacked_kernel(
    NfaTileDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, int32_t *best, const int32_t *__restrict__ only,
    const unsigned int *only_count) {
    extern __shared__ uint64_t lds[];
    // no input handed over (the common case behind the chain tiers): leave before staging
    // the tile's 128 KiB mask table (uniform over the block: every thread reads the count)
    if (only_count && __atomic_load_n(only_count, __ATOMIC_RELAXED) == 0u) return;
    const int32_t tile = blockIdx.x % d.tiles;
    const int32_t g = blockIdx.x / d.tiles;
    const int32_t G = gridDim.x / d.tiles;
    {
        const uint64_t *src = d.mask + static_cast<size_t>(tile) * 256 * kTileWords;
        copy_to_lds(reinterpret_cast<const uint4 *>(src), reinterpret_cast<uint4 *>(lds),
                    256 * kTileWords / 2);
        if (threadIdx.x < kTileWords) {
            lds[256 * kTileWords + threadIdx.x] = d.first[tile * kTileWords + threadIdx.x];
            lds[257 * kTileWords + threadIdx.x] = d.hassucc[tile * kTileWords + threadIdx.x];
            lds[258 * kTileWords + threadIdx.x] = d.acc[tile * kTileWords + threadIdx.x];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int32_t wave = threadIdx.x >> 6;
    const int32_t nwaves = blockDim.x >> 6;
    const uint64_t first = lds[256 * kTileWords + lane];
    const uint64_t hs = lds[257 * kTileWords + lane];
    const uint64_t acc = lds[258 * kTileWords + lane];
    const int64_t data_end = off[count];

    for (int32_t in = g * nwaves + wave; in < count; in += G * nwaves) {
        if (only && !only[in]) continue;
        const int64_t base = off[in];
        const int64_t len = off[in + 1] - base;
        int64_t limit = len;
        {
            const int32_t bst = __atomic_load_n(&best[in], __ATOMIC_RELAXED);
            if (bst != INT_MAX && bst - 1 < limit) limit = bst - 1;
        }
        uint64_t D = 0;
        bool s0 = true;
        int32_t found = 0;
        // Input is consumed in windows of 256 bytes: lane j holds bytes [4j, 4j + 4) of
        // the window (one coalesced 4-byte load per lane, guarded by the buffer end), and
        // each byte is broadcast with a uniform readlane.
        const int64_t abase = base & ~static_cast<int64_t>(3);
        int64_t wstart = abase;
        int64_t stop = base + limit;  // lowered when another tile reports an earlier match
        uint32_t word = bp_load_word(data, wstart + 4 * lane, stop, data_end);
        while (wstart < stop && !found) {
            const int64_t next = wstart + 256;
            const uint32_t nword = bp_load_word(data, next + 4 * lane, stop, data_end);
            // the input's best match so far (other tiles, other tiers), read with the next
            // window and applied after this one: the load has returned by then
            const int32_t bnow = __atomic_load_n(&best[in], __ATOMIC_RELAXED);
            for (int32_t j = 0; j < 64 && !found; ++j) {
                const uint32_t w4 = __builtin_amdgcn_readlane(word, j);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int64_t at = wstart + 4 * j + q;
                    if (at < base || at >= stop) continue;
                    const uint32_t b = (w4 >> (8 * q)) & 0xffu;
                    const uint64_t m = lds[b * kTileWords + lane];
                    D = bp_step(D, hs, s0 ? first : 0ull, m, lane);
                    if (!d.start_always) s0 = s0 && ((d.selfmask[b >> 6] >> (b & 63)) & 1ull);
                    if (__any((D & acc) != 0ull)) {
                        found = static_cast<int32_t>(at - base + 1);
                        break;
                    }
                }
                if (!found && !s0 && !__any(D != 0ull)) {
                    found = -1;  // this tile can no longer match
                }
            }
            word = nword;
            wstart = next;
            // a match ending at byte base + bnow - 1 is known: only an earlier one matters
            // here, so the scan limit drops to it (this tile's bytes before it still run)
            if (bnow != INT_MAX && base + bnow - 1 < stop) stop = base + bnow - 1;
        }
        if (found > 0 && lane == 0) atomicMin(&best[in], found);
    }
}

// The tiled bit-parallel pass alone (best[] must hold INT_MAX or earlier results): every
// input, or only those with only[i] != 0.
void launch_nfa_bitpacked_core(const NfaTileDev &d, const uint8_t *data, const int64_t *off,
                               int32_t count, int32_t *best, const int32_t *only,
                               const unsigned int *only_count, hipStream_t stream) {
    const size_t bytes = kBpLdsWords * sizeof(uint64_t);
    ensure_dynamic_lds(reinterpret_cast<const void *>(&nfa_bitpacked_kernel), bytes);
    const int32_t waves_per_block = kBpThreads / 64;
    int32_t G = (count + waves_per_block - 1) / waves_per_block;
    const int32_t target = (512 + d.tiles - 1) / d.tiles;  // about two blocks per CU
    if (G > target) G = target;
    if (G < 1) G = 1;
    nfa_bitpacked_kernel<<<G * d.tiles, kBpThreads, bytes, stream>>>(d, data, off, count, best,
                                                                    only, only_count);
    GFSM_HIP(hipGetLastError());
}

void launch_nfa_bitpacked(const NfaTileDev &d, const uint8_t *data, const int64_t *off,
                          int32_t count, int32_t *flags, int32_t *lens, int32_t *best,
                          hipStream_t stream) {
    if (count <= 0) return;
    bp_init_kernel<<<(count + 255) / 256, 256, 0, stream>>>(best, count);
    GFSM_HIP(hipGetLastError());
    launch_nfa_bitpacked_core(d, data, off, count, best, nullptr, nullptr, stream);
    launch_nfa_best_finish(best, count, flags, lens, stream);
}

void launch_nfa_best_finish(const int32_t *best, int32_t count, int32_t *flags, int32_t *lens,
                            hipStream_t stream) {
    bp_finish_kernel<<<(count + 255) / 256, 256, 0, stream>>>(best, count, flags, lens);
    GFSM_HIP(hipGetLastError());
}

}  // namespace gfsm
