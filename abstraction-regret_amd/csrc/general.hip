// NFA step kernels for general automata (any labels, wildcards, epsilon edges) on gfx950.
//
// Semantics: `src/gpufsm/reference.py:36-65` (simulate), latch-first-match. Every kernel
// here runs the automaton with its epsilon closure folded into the symbol edges on the host
// (layout.cpp, fold_epsilon): the set after byte i is the union of the folded successors of
// the set before it, which is closure(successors) of the reference, and the start set is
// closure(start). An accepting start closure never reaches these kernels; an empty set ends
// the input with (0, 0).
//
// gather    one wave64 per input; the state set is a bitset in LDS, one 64-bit word per
//           lane (lane j owns words j, j + 64, ...). A step on byte c walks the folded
//           edges labelled c or ANY, stored byte-major as (source << 16 | target), 64 at a
//           time: an edge whose source bit is set ORs its target bit into the next set
//           (ds_or_b64, no return). The work per byte is the number of edges labelled c,
//           whatever the size of the active set, so it is the technique for dense active
//           sets. Accept and "set empty" are one 64-bit ballot each. When the lists and the
//           accept words fit, the block stages them in LDS once. Upstream counterparts: the
//           bit-packed kernels `simulate_one<NW>` / `bitpacked_multistream_kernel` /
//           `bitpacked_multistream_shared_kernel` (`native/bitpacked.cu:12-166`), which scan
//           every state's CSR row per byte on one thread (up to 512 states).
//
// worklist  one wave64 per input over the compacted list of active states. Each lane takes
//           list entries, finds the state's edges on c (label-sorted rows; rows longer than
//           16 edges carry a per-label index) and ORs the targets into an LDS bitmap
//           (ds_or_b64, no return). The bitmap is then compacted into the next list: every
//           lane owns a run of words, counts their bits, a wave prefix sum gives its offset,
//           and it writes its state ids there and clears the words, so the bitmap is zero
//           again for the next byte. The accept test is fused into the compaction. Upstream
//           counterparts: `worklist_warp_kernel`, `worklist_shared_kernel` and
//           `worklist_compact_kernel` (`native/worklist_global.cu:128-342`).
//
// register  one lane per input, the state set in registers (NW 64-bit words, up to 512
//           states). The lane walks the set bits of its set and the label-sorted rows of
//           those states. Upstream counterpart: `worklist_multistream_kernel<NW>`
//           (`native/worklist_register.cu:16-128`).
#include "internal.hpp"

namespace gfsm {

namespace {

constexpr int kGenWaves = 4;  // waves per block (gather, worklist)
constexpr int kMaxDiag = 4;   // diagonal masks of the gather layout (NfaGeneralDev::doff)

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bytes [a, a + 4) of the batch as one little-endian word (a 4-aligned); bytes outside
// [0, data_end) read as zero.
__device__ __forceinline__ uint32_t gen_load_word(const uint8_t *__restrict__ data, int64_t a,
                                                  int64_t data_end) {
    if (a >= 0 && a + 4 <= data_end) return *reinterpret_cast<const uint32_t *>(data + a);
    uint32_t w = 0;
    for (int q = 0; q < 4; ++q)
        if (a + q >= 0 && a + q < data_end) w |= static_cast<uint32_t>(data[a + q]) << (8 * q);
    return w;
}

// Byte r (0..255) of a 256-byte window whose 4-byte words are spread over the lanes.
__device__ __forceinline__ uint32_t gen_window_byte(uint32_t word, int32_t r) {
    return (static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(word), r >> 2)) >>
            ((r & 3) << 3)) & 0xffu;
}

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ int32_t wave_inclusive_sum(int32_t x, int lane) {
#pragma unroll
    for (int dlt = 1; dlt < 64; dlt <<= 1) {
        const int32_t y = __shfl_up(x, dlt, 64);
        if (lane >= dlt) x += y;
    }
    return x;
}

// The wave's input bytes: a 256-byte window per 64 lanes, the next window in flight.
struct Window {
    int64_t wstart;   // absolute position of the window's first byte (4-aligned)
    uint32_t word, nword;
};

__device__ __forceinline__ void window_open(Window &w, const uint8_t *data, int64_t base,
                                            int64_t data_end, int lane) {
    w.wstart = base & ~static_cast<int64_t>(3);
    w.word = gen_load_word(data, w.wstart + 4 * lane, data_end);
    w.nword = gen_load_word(data, w.wstart + 256 + 4 * lane, data_end);
}

// Byte at absolute position `at` (>= the window start, wave-uniform and increasing).
__device__ __forceinline__ uint32_t window_at(Window &w, const uint8_t *data, int64_t at,
                                              int64_t data_end, int lane) {
    while (at >= w.wstart + 256) {
        w.wstart += 256;
        w.word = w.nword;
        w.nword = gen_load_word(data, w.wstart + 256 + 4 * lane, data_end);
    }
    return gen_window_byte(w.word, static_cast<int32_t>(at - w.wstart));
}

// Calls f(t) for every folded edge of state s labelled b or ANY.
template <typename F>
__device__ __forceinline__ void for_edges(const NfaGeneralDev &d, int32_t s, uint32_t b, F &&f) {
    const int32_t bi = d.bigidx[s];
    if (bi >= 0) {
        const int32_t *bo = d.bigoff + static_cast<int64_t>(bi) * kBigStride;
        for (int32_t e = bo[b], ee = bo[b + 1]; e < ee; ++e) f(d.edge[e] >> kLabelBits);
        for (int32_t e = bo[kAnySymbol], ee = bo[kAnySymbol + 1]; e < ee; ++e)
            f(d.edge[e] >> kLabelBits);
    } else {
        for (int32_t e = d.row[s], ee = d.row[s + 1]; e < ee; ++e) {
            const uint32_t v = d.edge[e];
            const uint32_t lab = v & ((1u << kLabelBits) - 1u);
            if (lab == b || lab == static_cast<uint32_t>(kAnySymbol)) f(v >> kLabelBits);
        }
    }
}

// Word w of the state set shifted by dd positions (dd in [-63, 63], not 0): bit s of the
// set lands on bit s + dd.
__device__ __forceinline__ uint64_t set_shifted(const uint64_t *set, int32_t w, int32_t nwords,
                                                int32_t dd) {
    if (dd > 0) {
        uint64_t x = set[w] << dd;
        if (w > 0) x |= set[w - 1] >> (64 - dd);
        return x;
    }
    uint64_t x = set[w] >> -dd;
    if (w + 1 < nwords) x |= set[w + 1] << (64 + dd);
    return x;
}

// The masks' part of the next set on byte c, for the words w = first, first + stride, ...:
// the first masks (successors of an always-active start) and, for each diagonal offset d,
// (set shifted by d) & mask (the masks are indexed by target). Written with plain stores
// (each word by its owner), so the gather lists' atomics come after a barrier. Four words
// per round: their mask loads are all issued before the first is used, one round trip per
// four words.
__device__ __forceinline__ void gather_masks_into(const NfaGeneralDev &d, const uint64_t *cur,
                                                  uint64_t *nxt, uint32_t c, int32_t first,
                                                  int32_t stride) {
    const int32_t nwords = d.nwords;
    if (!d.fmask && d.ndiag == 0) {
        for (int32_t w = first; w < nwords; w += stride) nxt[w] = 0;
        return;
    }
    constexpr int kW = 4;
    const uint64_t *fm = d.fmask ? d.fmask + static_cast<size_t>(c) * nwords : nullptr;
    const uint64_t *dm = d.dmask + static_cast<size_t>(c) * d.ndiag * nwords;
    for (int32_t w0 = first; w0 < nwords; w0 += kW * stride) {
        uint64_t f[kW], m[kW][kMaxDiag];
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int32_t w = w0 + i * stride;
            const bool in = w < nwords;
            f[i] = fm && in ? fm[w] : 0ull;
#pragma unroll
            for (int k = 0; k < kMaxDiag; ++k)
                m[i][k] = in && k < d.ndiag ? dm[static_cast<size_t>(k) * nwords + w] : 0ull;
        }
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int32_t w = w0 + i * stride;
            if (w >= nwords) break;
            uint64_t x = f[i];
#pragma unroll
            for (int k = 0; k < kMaxDiag; ++k) {
                if (k >= d.ndiag) break;
                x |= set_shifted(cur, w, nwords, d.doff[k]) & m[i][k];
            }
            nxt[w] = x;
        }
    }
}

// Edges of the byte's gather list (c, then ANY) applied to the next set: an edge whose
// source bit is set ORs its target bit in (ds_or_b64, no return).
__device__ __forceinline__ void gather_apply(const uint64_t *cur, uint64_t *nxt, uint32_t v) {
    const uint32_t s = v >> 16, t = v & 0xffffu;
    if ((cur[s >> 6] >> (s & 63)) & 1ull)
        atomicOr(reinterpret_cast<unsigned long long *>(&nxt[t >> 6]), 1ull << (t & 63));
}

}  // namespace

// ---------------------------------------------------------------- gather (bit-parallel) --

// LDS of a gather block: [staged: boff (258 int32, padded to 8-byte words) | accw | bedge
// (level 2 only)] then the bitsets. Staging level 1 holds the list offsets and accept words,
// level 2 also the lists.
__host__ __device__ static int64_t gather_stage_words(const NfaGeneralDev &d, int32_t level) {
    if (level <= 0) return 0;
    return (258 + 1) / 2 + d.nwords + (level >= 2 ? (static_cast<int64_t>(d.nedges) + 1) / 2 : 0);
}

// Copies the staged part of the layout into LDS (no barrier) and points the views at it.
__device__ __forceinline__ int64_t gather_stage(const NfaGeneralDev &d, int32_t level,
                                                uint64_t *lds, int32_t tid, int32_t nthreads,
                                                const int32_t *&boff, const uint64_t *&accw,
                                                const uint32_t *&bedge) {
    if (level <= 0) return 0;
    int32_t *sb = reinterpret_cast<int32_t *>(lds);
    uint64_t *sa = lds + (258 + 1) / 2;
    for (int32_t j = tid; j < 258; j += nthreads) sb[j] = d.boff[j];
    for (int32_t j = tid; j < d.nwords; j += nthreads) sa[j] = d.accw[j];
    boff = sb;
    accw = sa;
    if (level >= 2) {
        uint32_t *se = reinterpret_cast<uint32_t *>(sa + d.nwords);
        for (int32_t j = tid; j < d.nedges; j += nthreads) se[j] = d.bedge[j];
        bedge = se;
    }
    return gather_stage_words(d, level);
}

// Eight edges of the byte's list per thread, at j, j + stride, ..., loads issued together
// (the lists of a dense automaton are read from L2 or the MALL, one round trip per round);
// entries past the list's end read as 0 and are skipped when applied.
__device__ __forceinline__ void gather_load_round(const uint32_t *bedge, int32_t e0, int32_t nc,
                                                  int32_t any0, int32_t total, int32_t j,
                                                  int32_t stride, uint32_t (&v)[8]) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int32_t x = j + stride * q;
        v[q] = x < total ? bedge[x < nc ? e0 + x : any0 + (x - nc)] : 0u;
    }
}

__device__ __forceinline__ void gather_apply_round(const uint64_t *cur, uint64_t *nxt,
                                                   int32_t total, int32_t j, int32_t stride,
                                                   const uint32_t (&v)[8]) {
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (j + stride * q < total) gather_apply(cur, nxt, v[q]);
}

__global__ __launch_bounds__(64 * kGenWaves) void nfa_gather_kernel(
    NfaGeneralDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, int32_t *flags, int32_t *lens, int32_t staged) {
    extern __shared__ uint64_t lds[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int32_t nwords = d.nwords;
    const int32_t *boff = d.boff;
    const uint32_t *bedge = d.bedge;
    const uint64_t *accw = d.accw;
    const int64_t stage = gather_stage(d, staged, lds, threadIdx.x, blockDim.x, boff, accw, bedge);
    __syncthreads();
    uint64_t *cur = lds + stage + static_cast<int64_t>(wave) * 2 * nwords;
    uint64_t *nxt = cur + nwords;
    const int64_t data_end = off[count];
    const int32_t any0 = boff[kAnySymbol], any1 = boff[kAnySymbol + 1];

    for (int32_t in = blockIdx.x * kGenWaves + wave; in < count; in += gridDim.x * kGenWaves) {
        const int64_t base = off[in];
        const int64_t len = off[in + 1] - base;
        for (int32_t w = lane; w < nwords; w += 64) cur[w] = d.start_bits[w];
        wave_sync();
        int32_t res = 0;
        Window win;
        window_open(win, data, base, data_end, lane);
        for (int64_t pos = 0; pos < len; ++pos) {
            const uint32_t c = window_at(win, data, base + pos, data_end, lane);
            const int32_t e0 = boff[c];
            const int32_t nc = boff[c + 1] - e0;
            const int32_t total = nc + (any1 - any0);
            // the list's first round is loaded with the masks, one round trip for both
            uint32_t v[8];
            gather_load_round(bedge, e0, nc, any0, total, lane, 64, v);
            gather_masks_into(d, cur, nxt, c, lane, 64);
            wave_sync();
            gather_apply_round(cur, nxt, total, lane, 64, v);
            for (int32_t j = lane + 512; j < total; j += 512) {
                gather_load_round(bedge, e0, nc, any0, total, j, 64, v);
                gather_apply_round(cur, nxt, total, j, 64, v);
            }
            wave_sync();
            uint64_t any = 0, acc = 0;
            for (int32_t w = lane; w < nwords; w += 64) {
                const uint64_t x = nxt[w];
                any |= x;
                acc |= x & accw[w];
            }
            uint64_t *t = cur;
            cur = nxt;
            nxt = t;
            if (__any(acc != 0ull)) {
                res = static_cast<int32_t>(pos + 1);
                break;
            }
            if (!__any(any != 0ull)) break;
        }
        if (lane == 0) {
            flags[in] = res > 0;
            lens[in] = res;
        }
        wave_sync();  // every lane is done with cur/nxt before the next input's start set
    }
}

// The same step with kGenWaves waves of one block on one input: the block shares the
// input's two bitsets and splits the byte's edge list between its 256 lanes, with three
// barriers per byte (sets cleared; edges applied; accept and any-live reduced). One wave per
// input leaves a batch of as many inputs as SIMDs at one wave per SIMD, each waiting on its
// edge loads (61% of its cycles on the 65,536-state control); four waves per input keep four
// of them in flight per SIMD, and a block holds one input's bitsets (16 KiB at 65,536 states)
// instead of four.
__global__ __launch_bounds__(64 * kGenWaves) void nfa_gather_block_kernel(
    NfaGeneralDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, int32_t *flags, int32_t *lens, int32_t staged) {
    extern __shared__ uint64_t lds[];
    __shared__ unsigned int red[2][2];  // per byte parity: any live, accept
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int32_t nthreads = 64 * kGenWaves;
    const int32_t nwords = d.nwords;
    const int32_t *boff = d.boff;
    const uint32_t *bedge = d.bedge;
    const uint64_t *accw = d.accw;
    const int64_t stage = gather_stage(d, staged, lds, tid, nthreads, boff, accw, bedge);
    uint64_t *const set0 = lds + stage;
    const int64_t data_end = off[count];
    __syncthreads();
    const int32_t any0 = boff[kAnySymbol], any1 = boff[kAnySymbol + 1];

    for (int32_t in = blockIdx.x; in < count; in += gridDim.x) {  // block-uniform
        const int64_t base = off[in];
        const int64_t len = off[in + 1] - base;
        uint64_t *cur = set0, *nxt = set0 + nwords;
        for (int32_t w = tid; w < nwords; w += nthreads) cur[w] = d.start_bits[w];
        if (tid < 4) red[tid >> 1][tid & 1] = 0u;
        int32_t res = 0;
        Window win;
        window_open(win, data, base, data_end, lane);
        __syncthreads();
        for (int64_t pos = 0; pos < len; ++pos) {
            const int p = static_cast<int>(pos & 1);
            const uint32_t c = window_at(win, data, base + pos, data_end, lane);
            const int32_t e0 = boff[c];
            const int32_t nc = boff[c + 1] - e0;
            const int32_t total = nc + (any1 - any0);
            // the list's first round is loaded with the masks, one round trip for both
            uint32_t v[8];
            gather_load_round(bedge, e0, nc, any0, total, tid, nthreads, v);
            gather_masks_into(d, cur, nxt, c, tid, nthreads);
            __syncthreads();  // (a) nxt holds the masks' part, red[p] clear
            gather_apply_round(cur, nxt, total, tid, nthreads, v);
            for (int32_t j = tid + 8 * nthreads; j < total; j += 8 * nthreads) {
                gather_load_round(bedge, e0, nc, any0, total, j, nthreads, v);
                gather_apply_round(cur, nxt, total, j, nthreads, v);
            }
            __syncthreads();  // (b) nxt complete
            uint64_t any = 0, acc = 0;
            for (int32_t w = tid; w < nwords; w += nthreads) {
                const uint64_t x = nxt[w];
                any |= x;
                acc |= x & accw[w];
            }
            const bool wany = __any(any != 0ull), wacc = __any(acc != 0ull);
            if (lane == 0) {
                if (wany) atomicOr(&red[p][0], 1u);
                if (wacc) atomicOr(&red[p][1], 1u);
            }
            // the other parity was last read before barrier (a) of this byte: clear it for
            // the next byte
            if (tid < 2) red[p ^ 1][tid] = 0u;
            __syncthreads();  // (c) red[p] complete
            const bool live = red[p][0] != 0u, hit = red[p][1] != 0u;
            uint64_t *t = cur;
            cur = nxt;
            nxt = t;
            if (hit) {
                res = static_cast<int32_t>(pos + 1);
                break;
            }
            if (!live) break;
        }
        if (tid == 0) {
            flags[in] = res > 0;
            lens[in] = res;
        }
        __syncthreads();  // every thread is done with the sets before the next input
    }
}

static constexpr size_t kGatherStageBudget = 96 * 1024;
static constexpr size_t kGatherLdsMax = 160 * 1024;

bool nfa_gather_fits(const NfaGeneralDev &d) {
    return d.boff != nullptr &&
           static_cast<size_t>(kGenWaves) * 2 * d.nwords * 8 <= kGatherLdsMax;
}

void launch_nfa_gather(const NfaGeneralDev &d, const uint8_t *data, const int64_t *off,
                       int32_t count, int32_t *flags, int32_t *lens, hipStream_t stream) {
    if (count <= 0) return;
    if (!nfa_gather_fits(d))
        throw Error(3, "bitpacked (general): the bitsets of " + std::to_string(d.n) +
                           " states do not fit LDS");
    const size_t sets = static_cast<size_t>(kGenWaves) * 2 * d.nwords * 8;
    // stage the lists too when they are small, else the offsets and accept words only
    const size_t all = static_cast<size_t>(gather_stage_words(d, 2)) * 8;
    const size_t idx = static_cast<size_t>(gather_stage_words(d, 1)) * 8;
    const int32_t staged = all <= kGatherStageBudget && all + sets <= kGatherLdsMax ? 2
                           : (idx + sets <= kGatherLdsMax ? 1 : 0);
    const size_t stage = static_cast<size_t>(gather_stage_words(d, staged)) * 8;
    const size_t bytes = sets + stage;
    // One wave per input while the batch gives every SIMD four inputs of its own; below
    // that, a block of four waves per input (a wave per input would leave SIMDs with one
    // latency-bound wave, or none).
    const int64_t simds = 4 * static_cast<int64_t>(device_cu_count());
    if (static_cast<int64_t>(count) < 4 * simds || env_int("GFSM_GATHER_BLOCK", 0) == 1) {
        const size_t bbytes = static_cast<size_t>(2) * d.nwords * 8 + stage;
        ensure_dynamic_lds(reinterpret_cast<const void *>(&nfa_gather_block_kernel), bbytes);
        const int32_t grid = count < 65536 ? count : 65536;  // blocks loop over the rest
        nfa_gather_block_kernel<<<grid, 64 * kGenWaves, bbytes, stream>>>(
            d, data, off, count, flags, lens, staged);
        GFSM_HIP(hipGetLastError());
        return;
    }
    ensure_dynamic_lds(reinterpret_cast<const void *>(&nfa_gather_kernel), bytes);
    int32_t grid = (count + kGenWaves - 1) / kGenWaves;
    if (grid > 65536) grid = 65536;  // waves loop over the inputs beyond that
    nfa_gather_kernel<<<grid, 64 * kGenWaves, bytes, stream>>>(d, data, off, count, flags, lens,
                                                               staged);
    GFSM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- worklist (ballot) ------

// LDS of a worklist block: per wave the bitmap (nwords words), then, when they fit, the two
// lists (n int32 each).
static size_t wl2_wave_words(const NfaGeneralDev &d, bool lds_lists) {
    return static_cast<size_t>(d.nwords) + (lds_lists ? static_cast<size_t>(d.n) : 0);
}

__global__ __launch_bounds__(64 * kGenWaves) void nfa_wl2_kernel(
    NfaGeneralDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, int32_t *flags, int32_t *lens, int32_t *scratch, int32_t lds_lists) {
    extern __shared__ uint64_t lds[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int32_t nwords = d.nwords, n = d.n;
    const size_t per_wave = static_cast<size_t>(nwords) + (lds_lists ? static_cast<size_t>(n) : 0);
    uint64_t *bits = lds + static_cast<size_t>(wave) * per_wave;
    int32_t *la = lds_lists ? reinterpret_cast<int32_t *>(bits + nwords)
                            : scratch + (static_cast<int64_t>(blockIdx.x) * kGenWaves + wave) *
                                            2 * static_cast<int64_t>(n);
    int32_t *lb = la + n;
    for (int32_t w = lane; w < nwords; w += 64) bits[w] = 0;
    // each lane compacts a contiguous run of words
    const int32_t wpl = (nwords + 63) >> 6;
    const int32_t w0 = lane * wpl;
    const int32_t w1 = (w0 + wpl) < nwords ? (w0 + wpl) : nwords;
    const int64_t data_end = off[count];
    wave_sync();

    for (int32_t in = blockIdx.x * kGenWaves + wave; in < count; in += gridDim.x * kGenWaves) {
        const int64_t base = off[in];
        const int64_t len = off[in + 1] - base;
        int32_t *cur = la, *nxt = lb;
        for (int32_t j = lane; j < d.k0; j += 64) cur[j] = d.start_list[j];
        int32_t k = d.k0;
        int32_t res = 0;
        Window win;
        window_open(win, data, base, data_end, lane);
        wave_sync();
        for (int64_t pos = 0; pos < len; ++pos) {
            const uint32_t c = window_at(win, data, base + pos, data_end, lane);
            for (int32_t j = lane; j < k; j += 64) {
                for_edges(d, cur[j], c, [&](uint32_t t) {
                    atomicOr(reinterpret_cast<unsigned long long *>(&bits[t >> 6]),
                             1ull << (t & 63));
                });
            }
            wave_sync();
            int32_t cnt = 0;
            uint64_t acc = 0;
            for (int32_t w = w0; w < w1; ++w) {
                const uint64_t x = bits[w];
                cnt += __popcll(x);
                acc |= x & d.accw[w];
            }
            const int32_t incl = wave_inclusive_sum(cnt, lane);
            const int32_t total = __builtin_amdgcn_readlane(incl, 63);
            int32_t at = incl - cnt;
            for (int32_t w = w0; w < w1; ++w) {
                uint64_t x = bits[w];
                if (!x) continue;
                bits[w] = 0;
                while (x) {
                    nxt[at++] = (w << 6) + __builtin_ctzll(x);
                    x &= x - 1;
                }
            }
            wave_sync();
            if (__any(acc != 0ull)) {
                res = static_cast<int32_t>(pos + 1);
                break;
            }
            if (total == 0) break;
            int32_t *t = cur;
            cur = nxt;
            nxt = t;
            k = total;
        }
        if (lane == 0) {
            flags[in] = res > 0;
            lens[in] = res;
        }
        wave_sync();
    }
}

static constexpr size_t kWl2LdsBudget = 64 * 1024;           // per block
static constexpr int64_t kWl2ScratchBudget = 256ll << 20;    // bytes of global lists

void nfa_wl2_geometry(const NfaGeneralDev &d, int32_t count, int32_t *grid,
                      int64_t *scratch_elems) {
    const bool lds_lists = wl2_wave_words(d, true) * 8 * kGenWaves <= kWl2LdsBudget;
    int64_t g = (static_cast<int64_t>(count) + kGenWaves - 1) / kGenWaves;
    if (!lds_lists) {
        const int64_t per_block = static_cast<int64_t>(d.n) * 2 * kGenWaves;
        int64_t cap = kWl2ScratchBudget / (per_block * 4);
        if (cap < 64) cap = 64;
        if (g > cap) g = cap;
        *scratch_elems = g * per_block;
    } else {
        if (g > 65536) g = 65536;
        *scratch_elems = 0;
    }
    if (g < 1) g = 1;
    *grid = static_cast<int32_t>(g);
}

void launch_nfa_wl2(const NfaGeneralDev &d, const uint8_t *data, const int64_t *off,
                    int32_t count, int32_t *flags, int32_t *lens, int32_t *scratch,
                    int32_t grid, hipStream_t stream) {
    if (count <= 0) return;
    const bool lds_lists = scratch == nullptr;
    const size_t bytes = wl2_wave_words(d, lds_lists) * 8 * kGenWaves;
    if (bytes > kGatherLdsMax)
        throw Error(3, "worklist: the bitmap of " + std::to_string(d.n) +
                           " states does not fit LDS");
    ensure_dynamic_lds(reinterpret_cast<const void *>(&nfa_wl2_kernel), bytes);
    nfa_wl2_kernel<<<grid, 64 * kGenWaves, bytes, stream>>>(d, data, off, count, flags, lens,
                                                            scratch, lds_lists ? 1 : 0);
    GFSM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- register (lane/input) --

constexpr int kRegThreads = 256;

template <int NW>
__device__ __forceinline__ void reg_set(uint64_t (&x)[NW], uint32_t t) {
    const uint32_t w = t >> 6;
    const uint64_t bit = 1ull << (t & 63);
#pragma unroll
    for (int q = 0; q < NW; ++q) x[q] |= (w == static_cast<uint32_t>(q)) ? bit : 0ull;
}

template <int NW>
__global__ __launch_bounds__(kRegThreads) void nfa_reg_kernel(
    NfaGeneralDev d, const uint8_t *__restrict__ data, const int64_t *__restrict__ off,
    int32_t count, int32_t *flags, int32_t *lens) {
    uint64_t start[NW], accw[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        start[q] = q < d.nwords ? d.start_bits[q] : 0ull;
        accw[q] = q < d.nwords ? d.accw[q] : 0ull;
    }
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += gridDim.x * blockDim.x) {
        const int64_t base = off[i];
        const int64_t len = off[i + 1] - base;
        const uint8_t *p = data + base;
        uint64_t cur[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) cur[q] = start[q];
        int32_t res = 0;
        // bytes until the lane's position is 4-aligned, then aligned 4-byte words, then the
        // tail bytewise
        const int64_t lead = (4 - (base & 3)) & 3;
        const int64_t head = lead < len ? lead : len;
        int64_t pos = 0;
        bool stop = false;
        while (pos < len && !stop) {
            uint32_t word;
            int nb;
            if (pos >= head && pos + 4 <= len) {
                word = *reinterpret_cast<const uint32_t *>(p + pos);
                nb = 4;
            } else {
                word = p[pos];
                nb = 1;
            }
            for (int q = 0; q < nb && !stop; ++q, ++pos) {
                const uint32_t c = (word >> (8 * q)) & 0xffu;
                uint64_t nxt[NW];
#pragma unroll
                for (int w = 0; w < NW; ++w) nxt[w] = 0ull;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    uint64_t b = cur[w];
                    while (b) {
                        const int32_t st = w * 64 + __builtin_ctzll(b);
                        b &= b - 1;
                        for_edges(d, st, c, [&](uint32_t t) { reg_set<NW>(nxt, t); });
                    }
                }
                uint64_t any = 0, acc = 0;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    any |= nxt[w];
                    acc |= nxt[w] & accw[w];
                    cur[w] = nxt[w];
                }
                if (acc) {
                    res = static_cast<int32_t>(pos + 1);
                    stop = true;
                } else if (!any) {
                    stop = true;
                }
            }
        }
        flags[i] = res > 0;
        lens[i] = res;
    }
}

void launch_nfa_register(const NfaGeneralDev &d, const uint8_t *data, const int64_t *off,
                         int32_t count, int32_t *flags, int32_t *lens, hipStream_t stream) {
    if (count <= 0) return;
    const int32_t nw = d.nwords;
    if (nw > kRegMaxWords)
        throw Error(3, "register needs at most " + std::to_string(kRegMaxWords * 64) +
                           " states, the automaton has " + std::to_string(d.n));
    int32_t grid = (count + kRegThreads - 1) / kRegThreads;
    if (grid > 65536) grid = 65536;
#define GFSM_REG(NW) \
    nfa_reg_kernel<NW><<<grid, kRegThreads, 0, stream>>>(d, data, off, count, flags, lens)
    switch (nw) {
        case 1: GFSM_REG(1); break;
        case 2: GFSM_REG(2); break;
        case 3: GFSM_REG(3); break;
        case 4: GFSM_REG(4); break;
        case 5: GFSM_REG(5); break;
        case 6: GFSM_REG(6); break;
        case 7: GFSM_REG(7); break;
        default: GFSM_REG(8); break;
    }
#undef GFSM_REG
    GFSM_HIP(hipGetLastError());
}

}  // namespace gfsm
