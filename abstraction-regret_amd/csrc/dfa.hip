// DFA walk kernels for gfx950.
//
// Semantics: `src/gpufsm/reference.py:68-81` (simulate_dfa), latch-first-match. The
// upstream CUDA kernel is `native/dfa.cu:9-29`: one thread per string, int32 symbols, a
// separate int8 accept array, so two dependent loads and a branch per byte.
//
// Table layout (built by layout.cpp): states renumbered so that accepting states come last,
// table stored by byte (column c = successors of all states on byte c), each entry the
// successor's byte offset in a column. One step is
//     e = *(T *)(table + byte * col_bytes + e)       (v_mad_u32_u24 + one load)
// and "an accepting state was entered" is "e >= accept_from", tracked with one v_max per
// byte and tested once per block of bytes; a block with a hit is replayed byte by byte to
// find the exact position (rare by construction: it happens once per accepting input).
//
// Techniques
//   gather  one lane per input, table in global memory (uint16 entries up to 32767 states:
//           a 4096-state table is 2 MiB and stays in each XCD's 4 MB L2);
//   lds     the same walk with the table staged in LDS (tables up to 160 KiB: 320 states);
//   split   in-input parallel walk over fixed segments of the batch (see below).
#include "device.hpp"
#include "internal.hpp"

#include <climits>
#include <cstdlib>
#include <map>
#include <mutex>
#include <utility>

namespace gfsm {

// ---------------------------------------------------------------- step primitives -------

// One step from entry e on byte b. Byte tables (at most 256 states, 256-entry columns):
// the entry is the state id and the successor sits at (b << 8) | e. Wider tables: the entry
// is the successor's byte offset in a column, at b * col + e.
template <typename T>
__device__ __forceinline__ uint32_t dstep(const char *tab, uint32_t col, uint32_t e,
                                          uint32_t b) {
    if constexpr (sizeof(T) == 1)
        return *reinterpret_cast<const uint8_t *>(tab + ((b << 8) | e));
    else if constexpr (sizeof(T) == 2)
        return *reinterpret_cast<const T *>(tab + (__umul24(b, col) + e));
    else
        return *reinterpret_cast<const T *>(tab + (b * col + e));
}

// One step on byte k (0..3, a constant) of the input word w. For a byte table the address
// is one v_perm_b32: byte k of w into bits 8-15 and the state (byte 0 of e, whose other
// bytes are zero) into bits 0-7; the other bytes select zero, so the address is below 64 KiB.
template <typename T>
__device__ __forceinline__ uint32_t dstepw(const char *tab, uint32_t col, uint32_t e,
                                           uint32_t w, int k) {
    if constexpr (sizeof(T) == 1)
        return *reinterpret_cast<const uint8_t *>(
            tab + __builtin_amdgcn_perm(w, e, 0x0c0c0400u + (static_cast<uint32_t>(k) << 8)));
    else
        return dstep<T>(tab, col, e, (w >> (k * 8)) & 0xffu);
}

// 16 bytes of a vector, tracking the largest entry reached.
template <typename T>
__device__ __forceinline__ uint32_t step16(const char *tab, uint32_t col, uint32_t e,
                                           const uint4 &v, uint32_t &mx) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        e = dstepw<T>(tab, col, e, w[k >> 2], k & 3);
        mx = e > mx ? e : mx;
    }
    return e;
}

// Replay 16 bytes from e; index of the first step entering an accepting state (0..15).
template <typename T>
__device__ __forceinline__ int32_t locate16(const char *tab, uint32_t col, uint32_t e,
                                            const uint4 &v, uint32_t acc) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        e = dstepw<T>(tab, col, e, w[k >> 2], k & 3);
        if (e >= acc) return k;
    }
    return 15;  // unreachable when the block's maximum was >= acc
}

// Checkpoints: the state an interval ends in, with bit 0 set when an accepting state was
// entered in the interval. Entries of 2- and 4-byte tables are even (byte offsets of 2- or
// 4-byte entries), so bit 0 is free; byte-table entries are state ids, stored shifted left
// by one in a 16-bit checkpoint.
template <typename T>
struct Ck {
    using type = T;
    __device__ static uint32_t encode(uint32_t e, bool flag) { return e | (flag ? 1u : 0u); }
    __device__ static uint32_t state(uint32_t c) { return c & ~1u; }
};
template <>
struct Ck<uint8_t> {
    using type = uint16_t;
    __device__ static uint32_t encode(uint32_t e, bool flag) { return (e << 1) | (flag ? 1u : 0u); }
    __device__ static uint32_t state(uint32_t c) { return c >> 1; }
};

__device__ __forceinline__ uint4 ld16(const uint8_t *p) {
    return *reinterpret_cast<const uint4 *>(p);
}

// Exact walk of [a, b) from e: absolute position of the first byte after which an
// accepting state is active, or -1 (then e is the state after b). Never reads outside.
// R: bytes per round. 128 reads a whole cache line per round, so a line is not evicted
// between the rounds that read its halves when thousands of lanes stream at once.
template <typename T, int R = 64>
__device__ __forceinline__ int64_t walk_exact(const char *tab, uint32_t col, uint32_t acc,
                                              uint32_t &e, const uint8_t *__restrict__ data,
                                              int64_t a, int64_t b) {
    int64_t q = a;
    for (; q < b && (reinterpret_cast<uintptr_t>(data + q) & 15); ++q) {
        e = dstep<T>(tab, col, e, data[q]);
        if (e >= acc) return q;
    }
    // R-byte rounds with the next round's bytes in flight: 64 dependent steps outlast a
    // load, 16 do not. The read-ahead is unconditional (the last round re-reads its own
    // bytes) and held ahead of the round by a scheduling barrier, so the wait for the
    // current bytes counts the later loads instead of draining them.
    constexpr int kV = R / 16;
    if (q + R <= b) {
        uint4 nv[kV];
#pragma unroll
        for (int j = 0; j < kV; ++j) nv[j] = ld16(data + q + 16 * j);
        for (; q + R <= b; q += R) {
            uint4 v[kV];
#pragma unroll
            for (int j = 0; j < kV; ++j) v[j] = nv[j];
            const int64_t at = q + 2 * R <= b ? q + R : q;
#pragma unroll
            for (int j = 0; j < kV; ++j) nv[j] = ld16(data + at + 16 * j);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < kV; ++j) {
                const uint32_t e0 = e;
                uint32_t mx = 0;
                e = step16<T>(tab, col, e, v[j], mx);
                if (mx >= acc) return q + 16 * j + locate16<T>(tab, col, e0, v[j], acc);
            }
        }
    }
    for (; q + 16 <= b; q += 16) {
        const uint4 v = ld16(data + q);
        const uint32_t e0 = e;
        uint32_t mx = 0;
        e = step16<T>(tab, col, e, v, mx);
        if (mx >= acc) return q + locate16<T>(tab, col, e0, v, acc);
    }
    for (; q < b; ++q) {
        e = dstep<T>(tab, col, e, data[q]);
        if (e >= acc) return q;
    }
    return -1;
}

__global__ void fill_result_kernel(int32_t *flags, int32_t *lens, int32_t count, int32_t flag,
                                   int32_t len) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        flags[i] = flag;
        lens[i] = len;
    }
}

template <typename T, bool kLds>
__device__ __forceinline__ const char *stage_table(const T *gtab, uint32_t bytes,
                                                   uint4 *lds_raw) {
    if constexpr (kLds) {
        // tables are n * 256 entries: a multiple of 16 bytes
        copy_to_lds(reinterpret_cast<const uint4 *>(gtab), lds_raw, bytes / 16);
        __syncthreads();
        return reinterpret_cast<const char *>(lds_raw);
    } else {
        (void)bytes;
        (void)lds_raw;
        return reinterpret_cast<const char *>(gtab);
    }
}

// ---------------------------------------------------------------- one lane per input ----

template <typename T, bool kLds>
__global__ __launch_bounds__(1024) void dfa_walk_kernel(const T *__restrict__ gtab,
                                                       uint32_t table_bytes, uint32_t col,
                                                       uint32_t start, uint32_t acc,
                                                       const uint8_t *__restrict__ data,
                                                       const int64_t *__restrict__ off,
                                                       int32_t count, int32_t *flags,
                                                       int32_t *lens) {
    extern __shared__ uint4 lds_raw[];
    const char *tab = stage_table<T, kLds>(gtab, table_bytes, lds_raw);
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += gridDim.x * blockDim.x) {
        const int64_t ib = off[i];
        uint32_t e = start;
        const int64_t hit = walk_exact<T, kLds ? 128 : 64>(tab, col, acc, e, data, ib, off[i + 1]);
        flags[i] = hit >= 0;
        lens[i] = hit >= 0 ? static_cast<int32_t>(hit - ib + 1) : 0;
    }
}

// ---------------------------------------------------------------- split walk -------------
//
// Exact in-input parallelism (the chunked DFA run of Mytkowicz, Musuvathi and Schulte,
// "Data-parallel finite-state machines", ASPLOS 2014, specialised to latch-first-match).
// The concatenated batch is cut into fixed segments of C bytes, one lane per segment.
//
//   pass 1 (walk)     Inputs that start inside the segment are walked exactly from the
//                     start state; their accepts are final. A segment that begins inside an
//                     input (its "head") is walked from the start state as a guess, storing
//                     every R bytes a checkpoint: the state reached, with bit 0 set when an
//                     accepting state was entered in that interval. A head stops early once
//                     its input is known to have accepted.
//   pass 2 (fix)      Each head is re-walked from its probable true entry (the previous
//                     segment's exit for the same input: exact if the input started there,
//                     else that segment's head, which has merged with the true path unless
//                     it never met it) until its state equals a checkpoint. From there both
//                     paths coincide and the checkpoints give the rest of the verdict.
//                     Random DFAs merge after about num_states bytes.
//   pass 3 (event)    Per head: did pass 2 assume the entry that the previous segment
//                     actually produced? A wrong assumption or an accept is an event; the
//                     earliest event of each input is kept with a 64-bit atomicMin.
//   pass 4 (finish)   Per input: an accept event gives the verdict; a wrong assumption
//                     (rare) is resolved by walking on from there, again merging with the
//                     checkpoints. The verdict is therefore always the sequential one.

struct SplitArgs {
    const uint8_t *data;
    const int64_t *off;
    int32_t count;
    int32_t rshift;  // log2 R
    int64_t total, seg, nck, segments;
    uint32_t start, acc, col, table_bytes;
    int32_t *flags, *lens, *done;
    int32_t *seg_input;  // [segments] input holding the segment's first byte (walk pass)
    unsigned long long *first_event;
    uint32_t *head_exit, *tail_exit, *fixed_entry, *fixed_exit;
    int32_t *fixed_acc;
    // [segments] when nck <= 64: bit k set when interval k of the head entered an
    // accepting state on its speculative path (the flag bits of the checkpoints, gathered)
    unsigned long long *flagmask;
    // Heads resolved by their full transition map (map_idx null when not in use):
    int32_t *map_idx;          // [segments] the head's map, -1 = none
    int32_t *map_list;         // [map_cap] heads to map, in flag order
    unsigned int *map_count;   // heads flagged (may exceed map_cap)
    uint16_t *map_exit;        // [map_cap][parts][n] exit entry of each part, every entry
    int32_t *map_acc;          // [map_cap][parts][n] first accept in the part (offset in the
                               // head), -1 = none
    int32_t map_cap, n;        // n = states
    void *ckpt;  // T [segments * nck]
};

constexpr unsigned long long kNoEvent = ~0ull;

__device__ __forceinline__ int32_t split_find_input(const int64_t *__restrict__ off,
                                                    int32_t count, int64_t pos) {
    int32_t lo = 0, hi = count;  // off[lo] <= pos < off[count]
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (off[mid] <= pos) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void split_init_kernel(int32_t *flags, int32_t *lens, int32_t *done,
                                  unsigned long long *first_event, int32_t count,
                                  unsigned int *map_count) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && map_count) *map_count = 0u;
    if (i < count) {
        flags[i] = 0;
        lens[i] = 0;
        done[i] = 0;
        first_event[i] = kNoEvent;
    }
}

// ---- interleaved walking: K independent chains per lane -----------------------------------
//
// One dependent table lookup per byte makes a single chain latency-bound (a wave waits for
// every LDS read). Each lane therefore advances K chains together, byte by byte, so K
// independent loads are in flight per lane; a chain that finishes its part takes the next
// segment of the lane's list. A round moves every chain by 32 bytes when all of them sit on
// a 32-byte boundary with 32 bytes left in their part (the common case inside long inputs);
// otherwise the chains that are not aligned advance byte by byte to the next boundary.

constexpr int kRound = 64;               // bytes per chain per round (default)
constexpr int kVec = kRound / 16;        // 16-byte vectors per round
// Round of the walk kernel with K chains per lane over entries of esz bytes (the checkpoint
// interval is a multiple): 128 for byte tables with one chain per lane; the wider steps and
// the interleaved chains would spill registers at 128.
constexpr int split_walk_round(int K, int esz) { return K == 1 && esz == 1 ? 128 : 64; }

template <typename T, int K, int R = kRound>
__device__ __forceinline__ void step_round(const char *tab, uint32_t col, uint32_t (&e)[K],
                                           uint32_t (&mx)[K], const uint4 (&v)[K][R / 16]) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const uint4 &x = v[c][k >> 4];
            const int kk = k & 15;
            const uint32_t w = kk < 4 ? x.x : (kk < 8 ? x.y : (kk < 12 ? x.z : x.w));
            e[c] = dstepw<T>(tab, col, e[c], w, kk & 3);
            mx[c] = e[c] > mx[c] ? e[c] : mx[c];
        }
    }
}

// Replay 16 bytes from e; index of the first step entering an accepting state, or -1.
template <typename T>
__device__ __forceinline__ int32_t find16(const char *tab, uint32_t col, uint32_t &e,
                                          const uint4 &v, uint32_t acc) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        e = dstepw<T>(tab, col, e, w[k >> 2], k & 3);
        if (e >= acc) return k;
    }
    return -1;
}

// Index (0..R-1) of the first accepting step of a round of R bytes replayed from e.
template <typename T, int R = kRound>
__device__ __forceinline__ int32_t find_round(const char *tab, uint32_t col, uint32_t e,
                                              const uint4 (&v)[R / 16], uint32_t acc) {
    for (int j = 0; j < R / 16; ++j) {
        const int32_t k = find16<T>(tab, col, e, v[j], acc);
        if (k >= 0) return 16 * j + k;
    }
    return R - 1;  // unreachable when the round's maximum was >= acc
}

// Checkpoints are stored interval-major ([interval][segment]) so that the lanes of a wave,
// which hold consecutive segments, write and read whole contiguous lines.
template <typename T>
struct CkView {
    const typename Ck<T>::type *p;  // &ckpt[segment]
    int64_t step;                   // number of segments
    __device__ __forceinline__ uint32_t operator[](int64_t k) const {
        return static_cast<uint32_t>(p[k * step]);
    }
    __device__ __forceinline__ uint32_t state(int64_t k) const { return Ck<T>::state((*this)[k]); }
};

// Threads per block of the walk (one block per CU: the table fills the LDS). The walk keeps
// kWalkThreads * K chains in flight per CU; each step is one dependent LDS read, so the
// number of chains in flight sets the walk's rate.
constexpr int kWalkThreads = 1024;
constexpr int kFixThreads = 1024;
// The fix pass walks in rounds of R bytes: 64 up to 512 threads per CU, 32 above (hoisting
// the byte offsets of a 64-byte round takes more VGPRs than 1024 threads leave).
constexpr int fix_threads_for(int round) { return round >= 64 ? 512 : 1024; }

template <typename T, bool kLds, int K>
__global__ __launch_bounds__(kWalkThreads / K) void dfa_split_walk_kernel(
    const T *__restrict__ gtab, SplitArgs a) {
    extern __shared__ uint4 lds_raw[];
    // bytes per chain per round: a whole 128-byte line for byte tables with one chain per
    // lane (a line read in two rounds was often evicted from L2 in between when 100,000+
    // chains stream at once), else 64
    constexpr int kR = split_walk_round(K, sizeof(T));
    constexpr int kV = kR / 16;
    const char *tab = stage_table<T, kLds>(gtab, a.table_bytes, lds_raw);
    const uint32_t col = a.col, acc = a.acc, start = a.start;
    const int64_t rmask = (1ll << a.rshift) - 1;
    using CkT = typename Ck<T>::type;
    CkT *ckpt = static_cast<CkT *>(a.ckpt);
    const int64_t lanes = static_cast<int64_t>(gridDim.x) * blockDim.x;
    int64_t next_g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint4 zero = make_uint4(0, 0, 0, 0);

    const bool masks = a.nck <= 64;
    int64_t g[K], q[K], pend[K];
    int32_t in[K], mode[K];  // mode: 0 idle, 1 head (speculative), 2 exact
    uint32_t e[K], mx[K];
    unsigned long long fm[K];  // head: flagged intervals so far
    bool have[K];
    uint4 nv[K][kV];

    // chain c takes the next segment of this lane (or goes idle)
    auto pick = [&](int c) {
        have[c] = false;
        e[c] = start;
        mx[c] = 0;
        fm[c] = 0;
        if (next_g >= a.segments) {
            mode[c] = 0;
            return;
        }
        g[c] = next_g;
        next_g += lanes;
        const int64_t s0 = g[c] * a.seg;
        const int64_t s1 = (s0 + a.seg) < a.total ? (s0 + a.seg) : a.total;
        const int32_t i = split_find_input(a.off, a.count, s0);
        a.seg_input[g[c]] = i;
        in[c] = i;
        q[c] = s0;
        pend[c] = a.off[i + 1] < s1 ? a.off[i + 1] : s1;
        mode[c] = a.off[i] < s0 ? 1 : 2;
    };
    // chain c has reached the end of its part (q == pend); `normal`: walked to the end
    auto part_end = [&](int c, bool normal) {
        const int64_t s0 = g[c] * a.seg;
        const int64_t s1 = (s0 + a.seg) < a.total ? (s0 + a.seg) : a.total;
        if (normal) {
            if (mode[c] == 1) {
                const int64_t hl = pend[c] - s0;
                if (hl & rmask) {
                    ckpt[(hl >> a.rshift) * a.segments + g[c]] =
                        static_cast<CkT>(Ck<T>::encode(e[c], mx[c] >= acc));
                    if (mx[c] >= acc) fm[c] |= 1ull << ((hl >> a.rshift) & 63);
                }
                if (masks) a.flagmask[g[c]] = fm[c];
                a.head_exit[g[c]] = e[c];
            } else if (a.off[in[c] + 1] > s1) {
                a.tail_exit[g[c]] = e[c];
            }
        }
        if (pend[c] >= s1) {
            pick(c);
            return;
        }
        int32_t i = in[c] + 1;
        while (a.off[i + 1] <= pend[c]) ++i;  // empty inputs starting here
        in[c] = i;
        q[c] = pend[c];
        pend[c] = a.off[i + 1] < s1 ? a.off[i + 1] : s1;
        mode[c] = 2;
        e[c] = start;
        mx[c] = 0;
        have[c] = false;
    };
    auto accept_at = [&](int c, int64_t pos) {
        const int32_t i = in[c];
        a.flags[i] = 1;
        a.lens[i] = static_cast<int32_t>(pos - a.off[i] + 1);
        __atomic_store_n(&a.done[i], 1, __ATOMIC_RELAXED);
    };

#pragma unroll
    for (int c = 0; c < K; ++c) pick(c);

    for (;;) {
        bool any = false, fast = true;
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (mode[c]) {
                any = true;
                if ((q[c] & (kR - 1)) || q[c] + kR > pend[c]) fast = false;
            }
        }
        if (!any) break;
        if (fast) {
            uint4 v[K][kV];
            uint32_t e0[K];
            int32_t dn[K];
            // this round's bytes: read ahead by the previous round, or loaded now (the first
            // round of a part)
#pragma unroll
            for (int c = 0; c < K; ++c) {
                if (mode[c] && !have[c]) {
#pragma unroll
                    for (int j = 0; j < kV; ++j) nv[c][j] = ld16(a.data + q[c] + 16 * j);
                }
            }
#pragma unroll
            for (int c = 0; c < K; ++c)
#pragma unroll
                for (int j = 0; j < kV; ++j) v[c][j] = mode[c] ? nv[c][j] : zero;
            // The next round's bytes and the input's done flag, loaded while this round is
            // walked. The loads are unconditional (a chain with no next round re-reads this
            // round's bytes; an idle chain reads the batch's first bytes, which exist since
            // another chain of the lane walks a full round), so the wait for this round's
            // bytes counts only the loads issued after them instead of draining every load.
#pragma unroll
            for (int c = 0; c < K; ++c) {
                e0[c] = e[c];
                dn[c] = __atomic_load_n(&a.done[mode[c] ? in[c] : 0], __ATOMIC_RELAXED);
                const bool next = mode[c] && q[c] + 2 * kR <= pend[c];
                const int64_t at = next ? q[c] + kR : (mode[c] ? q[c] : 0);
#pragma unroll
                for (int j = 0; j < kV; ++j) nv[c][j] = ld16(a.data + at + 16 * j);
                have[c] = next;
            }
            // keep the loads above ahead of the round (the scheduler would otherwise sink them
            // to where their registers free up, near the round's end)
            __builtin_amdgcn_sched_barrier(0);
            step_round<T, K, kR>(tab, col, e, mx, v);
#pragma unroll
            for (int c = 0; c < K; ++c) {
                if (!mode[c]) {
                    e[c] = start;
                    continue;
                }
                q[c] += kR;
                if (mode[c] == 2) {
                    if (mx[c] >= acc) {
                        accept_at(c, q[c] - kR + find_round<T, kR>(tab, col, e0[c], v[c], acc));
                        q[c] = pend[c];
                        part_end(c, false);
                        continue;
                    }
                    mx[c] = 0;
                } else {
                    const int64_t r = q[c] - g[c] * a.seg;
                    if ((r & rmask) == 0) {
                        ckpt[((r >> a.rshift) - 1) * a.segments + g[c]] =
                            static_cast<CkT>(Ck<T>::encode(e[c], mx[c] >= acc));
                        if (mx[c] >= acc) fm[c] |= 1ull << (((r >> a.rshift) - 1) & 63);
                        mx[c] = 0;
                    }
                    if ((r & 255) == 0 && dn[c]) {  // the flag as of this round's start
                        q[c] = pend[c];
                        part_end(c, false);
                        continue;
                    }
                }
                if (q[c] == pend[c]) part_end(c, true);
            }
        } else {
#pragma unroll
            for (int c = 0; c < K; ++c) {
                if (!mode[c] || !((q[c] & (kR - 1)) || q[c] + kR > pend[c])) continue;
                // byte by byte to the next 32-byte boundary (or the end of the part)
                const int64_t up = (q[c] | (kR - 1)) + 1;
                const int64_t stop = up < pend[c] ? up : pend[c];
                bool hit = false;
                for (; q[c] < stop; ++q[c]) {
                    e[c] = dstep<T>(tab, col, e[c], a.data[q[c]]);
                    if (mode[c] == 2) {
                        if (e[c] >= acc) {
                            accept_at(c, q[c]);
                            hit = true;
                            break;
                        }
                    } else {
                        mx[c] = e[c] > mx[c] ? e[c] : mx[c];
                    }
                }
                have[c] = false;
                if (hit) {
                    q[c] = pend[c];
                    part_end(c, false);
                } else if (q[c] == pend[c]) {
                    part_end(c, true);
                }
            }
        }
    }
}

// First interval after `merged` whose checkpoint is flagged, or nk when none is: one load
// of the head's flag mask when the segment has at most 64 intervals, else a scan.
template <typename T>
__device__ __forceinline__ int64_t split_next_flagged(const SplitArgs &a, int64_t g, CkView<T> ck,
                                                      int64_t merged, int64_t nk) {
    if (a.nck <= 64) {
        const int64_t from = merged + 1;
        const unsigned long long m = from < 64 ? (a.flagmask[g] >> from) : 0ull;
        return m ? from + __builtin_ctzll(m) : nk;
    }
    int64_t k = merged + 1;
    while (k < nk && !(ck[k] & 1u)) ++k;
    return k;
}

// Re-walk the head [s0, hend) of segment g from `entry` until it merges with the
// checkpoints `ck`. acc_out: first accept offset relative to s0 (or -1); exit_out: state at
// hend.
template <typename T>
__device__ __forceinline__ void split_merge_walk(const char *tab, const SplitArgs &a, int64_t g,
                                                 int64_t s0, int64_t hend, uint32_t entry,
                                                 CkView<T> ck, uint32_t spec_exit,
                                                 int32_t &acc_out, uint32_t &exit_out) {
    const uint32_t col = a.col, acc = a.acc;
    const int64_t rmask = (1ll << a.rshift) - 1;
    const uint8_t *data = a.data;
    uint32_t e = entry;
    int64_t q = s0;
    int64_t merged = -1;  // index of the checkpoint where the paths met
    acc_out = -1;
    while (q + 16 <= hend) {
        const uint4 v = ld16(data + q);
        const uint32_t e0 = e;
        uint32_t mx = 0;
        e = step16<T>(tab, col, e, v, mx);
        if (mx >= acc) {
            acc_out = static_cast<int32_t>(q - s0 + locate16<T>(tab, col, e0, v, acc));
            exit_out = e;
            return;
        }
        q += 16;
        const int64_t r = q - s0;
        if ((r & rmask) == 0 && ck.state((r >> a.rshift) - 1) == e) {
            merged = (r >> a.rshift) - 1;
            break;
        }
    }
    if (merged < 0) {
        for (; q < hend; ++q) {
            e = dstep<T>(tab, col, e, data[q]);
            if (e >= acc) {
                acc_out = static_cast<int32_t>(q - s0);
                exit_out = e;
                return;
            }
        }
        exit_out = e;
        const int64_t hl = hend - s0;
        if (!(hl & rmask) || ck.state(hl >> a.rshift) != e) return;
        merged = hl >> a.rshift;  // met at the final partial interval: nothing after it
    }
    // paths coincide after checkpoint `merged`: take the first flagged interval after it
    const int64_t nk = ((hend - s0) + rmask) >> a.rshift;
    const int64_t k = split_next_flagged<T>(a, g, ck, merged, nk);
    if (k < nk) {
        uint32_t c = ck.state(k - 1);
        const int64_t lo = s0 + (k << a.rshift);
        const int64_t hi = (lo + rmask + 1) < hend ? (lo + rmask + 1) : hend;
        const int64_t hit = walk_exact<T>(tab, col, acc, c, data, lo, hi);
        acc_out = static_cast<int32_t>(hit - s0);  // the flag guarantees a hit
    }
    exit_out = spec_exit;
}

// After a merge at checkpoint `merged` the pass-1 path is the true one: the first accept
// after it is in the first flagged interval, located by walking that interval.
template <typename T>
__device__ __forceinline__ int32_t split_accept_after(const char *tab, const SplitArgs &a,
                                                      int64_t g, int64_t s0, int64_t hend,
                                                      CkView<T> ck, int64_t merged) {
    const int64_t rmask = (1ll << a.rshift) - 1;
    const int64_t nk = ((hend - s0) + rmask) >> a.rshift;
    const int64_t k = split_next_flagged<T>(a, g, ck, merged, nk);
    if (k >= nk) return -1;
    uint32_t c = ck.state(k - 1);
    const int64_t lo = s0 + (k << a.rshift);
    const int64_t hi = (lo + rmask + 1) < hend ? (lo + rmask + 1) : hend;
    const int64_t hit = walk_exact<T>(tab, a.col, a.acc, c, a.data, lo, hi);
    return static_cast<int32_t>(hit - s0);  // the flag guarantees a hit
}

// Pass 2 with K interleaved merge walks per lane (see the walk kernel). repair = 0: every
// head, entry = the previous segment's pass-1 exit. repair = 1: only heads whose assumed
// entry differs from the previous segment's pass-2 exit, now walked from that exit. A head
// being repaired can race with the repair of its predecessor; the event pass re-checks every
// head afterwards, so the race only affects speed.
template <typename T, bool kLds, int K, int kFixRound>
__global__ __launch_bounds__(fix_threads_for(kFixRound) / K) void dfa_split_fix_kernel(
    const T *__restrict__ gtab, SplitArgs a, int32_t repair) {
    extern __shared__ uint4 lds_raw[];
    // Heads are handed out from a counter per block over the block's contiguous range, so a
    // lane that finishes a short merge walk takes the next head: the pass then lasts about
    // as long as its longest head, not as long as the unluckiest lane's fixed share.
    __shared__ unsigned int s_next;
    if (threadIdx.x == 0) s_next = 0u;
    __syncthreads();
    const char *tab = stage_table<T, kLds>(gtab, a.table_bytes, lds_raw);
    const uint32_t col = a.col, acc = a.acc;
    const int64_t rmask = (1ll << a.rshift) - 1;
    const typename Ck<T>::type *ckpt = static_cast<const typename Ck<T>::type *>(a.ckpt);
    const int64_t per_block = (a.segments + gridDim.x - 1) / gridDim.x;
    const int64_t g_lo = static_cast<int64_t>(blockIdx.x) * per_block;
    const int64_t g_hi = (g_lo + per_block) < a.segments ? (g_lo + per_block) : a.segments;
    const uint4 zero = make_uint4(0, 0, 0, 0);

    int64_t g[K], q[K], hend[K];
    uint32_t e[K], entry[K];
    bool act[K], have[K];
    constexpr int kFixVec = kFixRound / 16;
    uint4 nv[K][kFixVec];

    // one LDS atomic per head taken, per lane (no wave-wide step inside): the loop ends when
    // the block's range is used up, and a lane that finds it used up stays idle
    auto pick = [&](int c) {
        act[c] = false;
        have[c] = false;
        e[c] = a.start;
        for (;;) {
            const int64_t gg = g_lo + static_cast<int64_t>(atomicAdd(&s_next, 1u));
            if (gg >= g_hi) return;
            if (gg == 0) continue;
            const int64_t s0 = gg * a.seg;
            const int32_t i = a.seg_input[gg];
            if (a.off[i] >= s0 || a.done[i]) continue;
            const bool started_before = a.off[i] >= (gg - 1) * a.seg;
            uint32_t en;
            if (repair) {
                en = started_before ? a.tail_exit[gg - 1]
                                    : __atomic_load_n(&a.fixed_exit[gg - 1], __ATOMIC_RELAXED);
                if (en == a.fixed_entry[gg]) continue;
            } else {
                en = started_before ? a.tail_exit[gg - 1] : a.head_exit[gg - 1];
            }
            const int64_t s1 = (s0 + a.seg) < a.total ? (s0 + a.seg) : a.total;
            g[c] = gg;
            q[c] = s0;
            hend[c] = a.off[i + 1] < s1 ? a.off[i + 1] : s1;
            e[c] = en;
            entry[c] = en;
            act[c] = true;
            return;
        }
    };
    auto finish = [&](int c, int32_t acc_at, uint32_t ex) {
        a.fixed_entry[g[c]] = entry[c];
        a.fixed_acc[g[c]] = acc_at;
        __atomic_store_n(&a.fixed_exit[g[c]], ex, __ATOMIC_RELAXED);
        pick(c);
    };

#pragma unroll
    for (int c = 0; c < K; ++c) pick(c);

    for (;;) {
        bool any = false, fast = true;
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (act[c]) {
                any = true;
                if ((q[c] & (kFixRound - 1)) || q[c] + kFixRound > hend[c]) fast = false;
            }
        }
        if (!any) break;
        if (fast) {
            uint4 v[K][kFixVec];
            uint32_t e0[K], mx[K], ckr[K];
            bool at_ck[K];
            // this round's bytes: read ahead by the previous round, or loaded now (the first
            // round of a head)
#pragma unroll
            for (int c = 0; c < K; ++c) {
                if (act[c] && !have[c]) {
#pragma unroll
                    for (int j = 0; j < kFixVec; ++j) nv[c][j] = ld16(a.data + q[c] + 16 * j);
                }
            }
#pragma unroll
            for (int c = 0; c < K; ++c)
#pragma unroll
                for (int j = 0; j < kFixVec; ++j) v[c][j] = act[c] ? nv[c][j] : zero;
            // The checkpoint the round ends on and the next round's bytes, loaded while the
            // round is walked. Unconditional loads (clamped to valid addresses, see the walk
            // kernel) keep the wait for this round's bytes a count of the later loads.
#pragma unroll
            for (int c = 0; c < K; ++c) {
                e0[c] = e[c];
                mx[c] = 0;
                const int64_t r_end = act[c] ? q[c] + kFixRound - g[c] * a.seg : 0;
                at_ck[c] = act[c] && (r_end & rmask) == 0;
                ckr[c] = static_cast<uint32_t>(
                    ckpt[at_ck[c] ? ((r_end >> a.rshift) - 1) * a.segments + g[c] : 0]);
                const bool next = act[c] && q[c] + 2 * kFixRound <= hend[c];
                const int64_t at = next ? q[c] + kFixRound : (act[c] ? q[c] : 0);
#pragma unroll
                for (int j = 0; j < kFixVec; ++j) nv[c][j] = ld16(a.data + at + 16 * j);
                have[c] = next;
            }
            __builtin_amdgcn_sched_barrier(0);  // the loads stay ahead of the round
            step_round<T, K, kFixRound>(tab, col, e, mx, v);
#pragma unroll
            for (int c = 0; c < K; ++c) {
                if (!act[c]) {
                    e[c] = a.start;
                    continue;
                }
                // the state the head's speculative path had at the round's end, if the round
                // ends on a checkpoint
                const uint32_t ckn = at_ck[c] ? Ck<T>::state(ckr[c]) : ~0u;
                const int64_t s0 = g[c] * a.seg;
                q[c] += kFixRound;
                if (mx[c] >= acc) {
                    const int32_t k = find_round<T, kFixRound>(tab, col, e0[c], v[c], acc);
                    finish(c, static_cast<int32_t>(q[c] - kFixRound + k - s0), e[c]);
                    continue;
                }
                const int64_t r = q[c] - s0;
                const CkView<T> ck{ckpt + g[c], a.segments};
                if (ckn == e[c]) {
                    const int32_t after = split_accept_after<T>(tab, a, g[c], s0, hend[c], ck,
                                                                (r >> a.rshift) - 1);
                    finish(c, after, a.head_exit[g[c]]);
                    continue;
                }
                if (q[c] == hend[c]) finish(c, -1, e[c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < K; ++c) {
                if (!act[c] || !((q[c] & (kFixRound - 1)) || q[c] + kFixRound > hend[c])) continue;
                // the last partial round of a head: byte by byte to its end
                const int64_t s0 = g[c] * a.seg;
                int32_t hit = -1;
                for (; q[c] < hend[c]; ++q[c]) {
                    e[c] = dstep<T>(tab, col, e[c], a.data[q[c]]);
                    if (e[c] >= acc) {
                        hit = static_cast<int32_t>(q[c] - s0);
                        break;
                    }
                }
                have[c] = false;
                finish(c, hit, e[c]);
            }
        }
    }
}

// ---- full transition maps of unresolved heads ---------------------------------------------
//
// A head whose assumed entry was wrong, and whose repair would be a single-lane walk of up to
// a whole segment (the tail of the split walk), is resolved in parallel instead: its
// transition map for every entry state. The head is cut into 32 intervals, handled by 4
// blocks of 8 intervals each (one CU per block, so one head's lookups spread over 4 CUs).
// Each (interval, entry) pair is walked by one lane, two entries per lane walked together,
// recording the first accepting step; then one lane per entry composes the block's 8
// interval maps into a part map. The finish pass chains the 4 part maps of the head for its
// one true entry. That is n times the lookups of one walk, spread over 4096 lanes with at
// most interval-length dependent steps, which beats the single walk when few heads need it.
constexpr int kMapIntervals = 32;
constexpr int kMapParts = 4;
constexpr int kMapPartIntervals = kMapIntervals / kMapParts;  // 8
constexpr int kMapThreads = 1024;                              // 8 intervals x 128 lanes
constexpr uint16_t kNoHit = 0xffffu;

template <typename T>
__global__ __launch_bounds__(kMapThreads) void dfa_split_map_kernel(const T *__restrict__ gtab,
                                                                    SplitArgs a) {
    extern __shared__ uint4 lds_raw[];
    const unsigned int flagged = *a.map_count;
    const unsigned int cnt = flagged < static_cast<unsigned int>(a.map_cap)
                                 ? flagged : static_cast<unsigned int>(a.map_cap);
    // a block without a part to map leaves before staging the table (the whole block takes
    // the same branch): usually few heads need maps, and an idle block holding a CU's LDS
    // and 1024 lanes keeps other work (the next batch's walk) off that CU
    if (blockIdx.x >= cnt * kMapParts) return;
    const char *tab = stage_table<T, true>(gtab, a.table_bytes, lds_raw);
    const int n = a.n;  // at most 256: two entries per lane
    uint16_t *imap = reinterpret_cast<uint16_t *>(reinterpret_cast<char *>(lds_raw) + a.table_bytes);
    uint16_t *ihit = imap + kMapPartIntervals * n;
    const uint32_t col = a.col, acc = a.acc;
    const int64_t I = a.seg / kMapIntervals;  // a multiple of 16
    const int kk = threadIdx.x >> 7, sub = threadIdx.x & 127;
    for (unsigned int w = blockIdx.x; w < cnt * kMapParts; w += gridDim.x) {
        const unsigned int m = w / kMapParts, part = w % kMapParts;
        const int64_t g = a.map_list[m];
        const int64_t s0 = g * a.seg;
        const int32_t in = a.seg_input[g];
        int64_t hend = s0 + a.seg < a.total ? s0 + a.seg : a.total;
        if (a.off[in + 1] < hend) hend = a.off[in + 1];
        const int k = static_cast<int>(part) * kMapPartIntervals + kk;  // interval in the head
        const int64_t lo = s0 + k * I;
        const int64_t hi = (lo + I) < hend ? (lo + I) : hend;
        uint32_t e[2], hit[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int x = sub + 128 * j;  // entries past n walk state 0 and are not stored
            e[j] = static_cast<uint32_t>(x < n ? x : 0) * static_cast<uint32_t>(sizeof(T));
            hit[j] = kNoHit;
        }
        // the interval's bytes, one vector ahead: its loads are issued together and only the
        // vector being walked is waited for (unconditional loads, the last one re-read)
        int64_t q = lo;
        const int nvec = static_cast<int>((hi - lo) >> 4);
        uint4 nv = ld16(a.data + (nvec > 0 ? lo : 0));
        for (int k = 0; k < nvec; ++k, q += 16) {
            const uint4 v = nv;
            nv = ld16(a.data + (k + 1 < nvec ? q + 16 : q));
            __builtin_amdgcn_sched_barrier(0);
            const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int b = 0; b < 16; ++b) {
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    e[j] = dstepw<T>(tab, col, e[j], wv[b >> 2], b & 3);
                    if (hit[j] == kNoHit && e[j] >= acc) hit[j] = static_cast<uint32_t>(q - lo + b);
                }
            }
        }
        for (; q < hi; ++q) {
            const uint32_t c = a.data[q];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                e[j] = dstep<T>(tab, col, e[j], c);
                if (hit[j] == kNoHit && e[j] >= acc) hit[j] = static_cast<uint32_t>(q - lo);
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int x = sub + 128 * j;
            if (x < n) {
                imap[kk * n + x] = static_cast<uint16_t>(e[j]);
                ihit[kk * n + x] = static_cast<uint16_t>(hit[j]);
            }
        }
        __syncthreads();
        // compose the block's intervals: the part map, first accept as an offset in the head
        for (int x = threadIdx.x; x < n; x += blockDim.x) {
            uint32_t st = static_cast<uint32_t>(x), ev = static_cast<uint32_t>(x) * sizeof(T);
            int32_t first = -1;
            for (int j = 0; j < kMapPartIntervals; ++j) {
                if (first < 0) {
                    const uint32_t h = ihit[j * n + st];
                    if (h != kNoHit)
                        first = static_cast<int32_t>((part * kMapPartIntervals + j) * I + h);
                }
                ev = imap[j * n + st];
                st = ev / sizeof(T);
            }
            const int64_t at = static_cast<int64_t>(w) * n + x;  // w = m * kMapParts + part
            a.map_exit[at] = static_cast<uint16_t>(ev);
            a.map_acc[at] = first;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void dfa_split_event_kernel(SplitArgs a) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t g = 1 + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         g < a.segments; g += stride) {
        const int64_t s0 = g * a.seg;
        const int32_t i = a.seg_input[g];
        if (a.off[i] >= s0 || a.done[i]) continue;
        const uint32_t prev =
            a.off[i] >= (g - 1) * a.seg ? a.tail_exit[g - 1] : a.fixed_exit[g - 1];
        unsigned long long key = kNoEvent;
        const bool wrong = a.fixed_entry[g] != prev;
        if (wrong) key = 2ull * static_cast<unsigned long long>(g);
        else if (a.fixed_acc[g] >= 0) key = 2ull * static_cast<unsigned long long>(g) + 1ull;
        if (key != kNoEvent) atomicMin(&a.first_event[i], key);
        if (a.map_idx) {
            // a head whose assumed entry differs from its predecessor's exit gets its full
            // transition map, so that the finish pass resolves it with one lookup
            int32_t m = -1;
            if (wrong) {
                const unsigned int slot = atomicAdd(a.map_count, 1u);
                if (slot < static_cast<unsigned int>(a.map_cap)) {
                    m = static_cast<int32_t>(slot);
                    a.map_list[slot] = static_cast<int32_t>(g);
                }
            }
            a.map_idx[g] = m;
        }
    }
}

constexpr int kAhead = 8;          // heads whose fix results the finish pass loads together
constexpr int32_t kMapUnknown = -2;  // finish pass: head g's map not looked up yet

// Chain the part maps of map m for entry cur: cur becomes the head's exit; returns the
// first accept (offset in the head) or -1.
template <typename T>
__device__ __forceinline__ int32_t split_map_chain(const SplitArgs &a, int32_t m, uint32_t &cur) {
    const int64_t base = static_cast<int64_t>(m) * kMapParts;
    uint32_t st = cur / sizeof(T);
    int32_t acc_at = -1;
    for (int p = 0; p < kMapParts; ++p) {
        const int64_t at = (base + p) * a.n + st;
        if (acc_at < 0) acc_at = a.map_acc[at];
        cur = a.map_exit[at];
        st = cur / sizeof(T);
    }
    return acc_at;
}

// Finish-pass resolution of heads g..g1 of input i from the true entry `cur`, without the
// table. A head whose fix assumed `cur` keeps its fix result: those of kAhead heads are
// loaded together (with their map indexes), so a consistent run costs one memory round
// trip per kAhead heads. A head whose fix assumed another entry chains its part maps. `m`
// is head g's map index when the caller already knows that head g's entry was wrong (-1:
// it has none), else kMapUnknown. Writes the verdict and returns false at an accept or at
// the input's end; returns true, with g and cur set, at a head that needs a merge walk.
template <typename T>
__device__ bool split_resolve(const SplitArgs &a, int32_t i, int64_t ib, int64_t &g, int64_t g1,
                              uint32_t &cur, int32_t m) {
    while (g <= g1) {
        if (m == kMapUnknown) {
            uint32_t fe[kAhead], fx[kAhead];
            int32_t fa[kAhead], mi[kAhead];
#pragma unroll
            for (int k = 0; k < kAhead; ++k) {
                const int64_t h = (g + k <= g1) ? g + k : g1;  // only heads of this input
                fe[k] = a.fixed_entry[h];
                fa[k] = a.fixed_acc[h];
                fx[k] = a.fixed_exit[h];
                mi[k] = a.map_idx ? a.map_idx[h] : -1;
            }
            const int64_t w = (g1 - g + 1) < kAhead ? (g1 - g + 1) : kAhead;
            int64_t k = 0;  // heads of the window that were consistent so far
#pragma unroll
            for (int kk = 0; kk < kAhead; ++kk) {
                if (kk == k && kk < w) {
                    if (fe[kk] == cur) {
                        if (fa[kk] >= 0) {
                            a.flags[i] = 1;
                            a.lens[i] = static_cast<int32_t>((g + kk) * a.seg + fa[kk] - ib + 1);
                            return false;
                        }
                        cur = fx[kk];
                        k = kk + 1;
                    } else {
                        m = mi[kk];
                    }
                }
            }
            g += k;
            if (k == w) continue;
        }
        // head g: its fix assumed another entry
        if (m < 0) return true;
        const int32_t acc_at = split_map_chain<T>(a, m, cur);
        if (acc_at >= 0) {
            a.flags[i] = 1;
            a.lens[i] = static_cast<int32_t>(g * a.seg + acc_at - ib + 1);
            return false;
        }
        ++g;
        m = kMapUnknown;
    }
    return false;
}

template <typename T, bool kLds>
__global__ __launch_bounds__(256) void dfa_split_finish_kernel(const T *__restrict__ gtab,
                                                               SplitArgs a) {
    extern __shared__ uint4 lds_raw[];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ib = 0, ie = 0, g = 0, g1 = -1;
    uint32_t cur = 0;
    int32_t m = kMapUnknown;
    if (i < a.count) {
        const int32_t dn = a.done[i];
        const unsigned long long key = a.first_event[i];
        if (!dn && key != kNoEvent) {  // else: done, or consistent with no accept, (0, 0)
            g = static_cast<int64_t>(key >> 1);
            ib = a.off[i];
            ie = a.off[i + 1];
            if (key & 1ull) {
                const int32_t fa = a.fixed_acc[g];
                a.flags[i] = 1;
                a.lens[i] = static_cast<int32_t>(g * a.seg + fa - ib + 1);
            } else {
                // the entry the fix pass assumed for head g is wrong (that is the event):
                // resolve from the true entry, starting with head g's map
                const uint32_t te = a.tail_exit[g - 1], fxp = a.fixed_exit[g - 1];
                m = a.map_idx ? a.map_idx[g] : -1;
                cur = ib >= (g - 1) * a.seg ? te : fxp;
                g1 = (ie - 1) / a.seg;
            }
        }
    }
    bool walk = g <= g1 && split_resolve<T>(a, i, ib, g, g1, cur, m);
    const char *tab = reinterpret_cast<const char *>(gtab);
    if constexpr (kLds) {
        // the table is staged only when some input of the block needs a merge walk (the
        // condition is block-uniform, so every thread reaches the barrier inside)
        if (__syncthreads_or(walk ? 1 : 0)) tab = stage_table<T, kLds>(gtab, a.table_bytes, lds_raw);
    }
    const typename Ck<T>::type *ckpt = static_cast<const typename Ck<T>::type *>(a.ckpt);
    while (walk) {
        const int64_t s0 = g * a.seg;
        const int64_t hend = (s0 + a.seg) < ie ? (s0 + a.seg) : ie;
        int32_t acc_at;
        uint32_t ex;
        split_merge_walk<T>(tab, a, g, s0, hend, cur, CkView<T>{ckpt + g, a.segments},
                            a.head_exit[g], acc_at, ex);
        if (acc_at >= 0) {
            a.flags[i] = 1;
            a.lens[i] = static_cast<int32_t>(s0 + acc_at - ib + 1);
            break;
        }
        cur = ex;
        ++g;
        walk = g <= g1 && split_resolve<T>(a, i, ib, g, g1, cur, kMapUnknown);
    }
}

// ---------------------------------------------------------------- host side -------------

static size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Launch-shape facts cached per (device, kernel). Several host threads may launch at once
// (the serving loop does), so the caches are guarded by one mutex; the values they hold
// are the device's, so each device gets its own.
namespace {
std::mutex g_shape_mu;
std::map<std::pair<int, const void *>, size_t> g_lds_limit;    // raised dynamic-LDS limit
std::map<std::pair<int, const void *>, int> g_blocks_per_cu;  // occupancy at 256 threads
std::map<int, int32_t> g_cus;

int current_device() {
    int dev = 0;
    GFSM_HIP(hipGetDevice(&dev));
    return dev;
}
}  // namespace

int32_t device_cu_count() {
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_cus.find(dev);
    if (it != g_cus.end()) return it->second;
    int cus = 0;
    GFSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return g_cus[dev] = cus > 0 ? cus : 1;
}

void ensure_dynamic_lds(const void *kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;  // the default limit
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(g_shape_mu);
    size_t &have = g_lds_limit[{dev, kernel}];
    if (have >= bytes) return;
    GFSM_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(bytes)));
    have = bytes;
}

int persistent_blocks(const void *kernel, int threads) {
    const int32_t cus = device_cu_count();
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_blocks_per_cu.find({dev, kernel});
    int per_cu;
    if (it != g_blocks_per_cu.end()) {
        per_cu = it->second;
    } else {
        per_cu = 0;
        GFSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
        if (per_cu < 1) per_cu = 1;
        g_blocks_per_cu[{dev, kernel}] = per_cu;
    }
    return cus * per_cu;
}

size_t dfa_table_bytes(const DfaDev &d) { return static_cast<size_t>(d.col_bytes) * kAlphabet; }

static bool lds_fits(const DfaDev &d) { return !d.wide && dfa_table_bytes(d) <= 160 * 1024; }

// Tuning overrides for sweeps: GFSM_SPLIT_SEGMENT (bytes, power of two >= 256) and
// GFSM_SPLIT_CHAINS (1, 2 or 4 interleaved chains per lane).
int64_t env_int(const char *name, int64_t fallback) {
    const char *v = std::getenv(name);
    if (!v || !*v) return fallback;
    char *end = nullptr;
    const long long x = std::strtoll(v, &end, 10);
    return (end && *end == '\0' && x > 0) ? x : fallback;
}

DfaSplitPlan dfa_split_plan(const DfaDev &d, int64_t total_bytes, int32_t count,
                            int64_t seg_bytes) {
    DfaSplitPlan p;
    p.chains = static_cast<int32_t>(env_int("GFSM_SPLIT_CHAINS", 1));
    if (p.chains != 1 && p.chains != 2 && p.chains != 4) p.chains = 1;
    const bool valid = seg_bytes >= 256 && seg_bytes <= (1 << 20) && !(seg_bytes & (seg_bytes - 1));
    // the checkpoint interval is a multiple of the walk's round (split_walk_round)
    const int32_t round_shift = split_walk_round(p.chains, d.esz) == 128 ? 7 : 6;
    if (lds_fits(d)) {
        int64_t c = valid ? seg_bytes : env_int("GFSM_SPLIT_SEGMENT", 2048);
        if (c < 256 || c > (1 << 20) || (c & (c - 1))) c = 2048;
        p.seg_bytes = c;
        p.ckpt_shift = round_shift;
    } else {
        int64_t c = 1024;
        while (c < 4ll * d.n && c < 65536) c <<= 1;
        if (valid) c = seg_bytes;
        p.seg_bytes = c;
        p.ckpt_shift = round_shift;
        while ((1ll << p.ckpt_shift) * 16 < c) ++p.ckpt_shift;
    }
    p.segments = (total_bytes + p.seg_bytes - 1) / p.seg_bytes;
    // maps need the table and 32 interval maps in LDS (n <= 256) and 16-byte intervals
    p.maps = lds_fits(d) && d.n <= 256 && p.seg_bytes >= 512 && p.chains == 1;
    p.map_cap = p.maps ? 1024 : 0;
    const size_t S = static_cast<size_t>(p.segments > 0 ? p.segments : 1);
    const size_t nck = static_cast<size_t>(p.seg_bytes >> p.ckpt_shift);
    const size_t ent = d.wide ? 4 : 2;  // checkpoint entry (16 bits for byte and uint16 tables)
    const size_t cap = static_cast<size_t>(p.map_cap);
    const size_t nn = static_cast<size_t>(d.n);
    p.scratch_bytes = align256(static_cast<size_t>(count) * 4) +
                      align256(static_cast<size_t>(count) * 8) + 6 * align256(S * 4) +
                      align256(S * 8) + align256(S * nck * ent) +
                      (p.maps ? align256(S * 4) + align256(cap * 4) + 256 +
                                    align256(cap * kMapParts * nn * 2) +
                                    align256(cap * kMapParts * nn * 4)
                              : 0);
    return p;
}

// Raise a kernel's dynamic-LDS limit to `bytes` (needed above 64 KiB). The limit is what
// the launch uses, not the CU maximum: the kernel's own static LDS counts against the CU.
template <typename K>
static void allow_big_lds(K *kernel, size_t bytes) {
    ensure_dynamic_lds(reinterpret_cast<const void *>(kernel), bytes);
}

template <typename T, bool kLds, int kChains>
static void split_launch_passes(const DfaDev &d, SplitArgs &a, hipStream_t stream) {
    const T *tab = static_cast<const T *>(d.table);
    const size_t lds = kLds ? static_cast<size_t>(a.table_bytes) : 0;
    const int32_t cus = device_cu_count();
    if (kLds) {
        allow_big_lds(&dfa_split_walk_kernel<T, kLds, kChains>, lds);
        allow_big_lds(&dfa_split_fix_kernel<T, kLds, kChains, 64>, lds);
        allow_big_lds(&dfa_split_fix_kernel<T, kLds, kChains, 32>, lds);
    }
    // spread the segments over every CU, kChains per lane
    struct Shape {
        int32_t blocks, threads;
    };
    auto shape = [&](int32_t max_threads) {
        int64_t per_cu = (a.segments + static_cast<int64_t>(cus) * kChains - 1) /
                         (static_cast<int64_t>(cus) * kChains);
        per_cu = ((per_cu + 63) / 64) * 64;
        Shape sh;
        if (kLds) {  // the table fills the LDS: one block per CU
            const int64_t cap = max_threads / kChains;
            sh.threads = static_cast<int32_t>(per_cu < 64 ? 64 : (per_cu > cap ? cap : per_cu));
            const int64_t need = (a.segments + static_cast<int64_t>(sh.threads) * kChains - 1) /
                                 (static_cast<int64_t>(sh.threads) * kChains);
            sh.blocks = static_cast<int32_t>(need < cus ? (need > 0 ? need : 1) : cus);
        } else {
            sh.threads = per_cu >= 256 / kChains ? 256 / kChains : 64;
            const int64_t need = (a.segments + static_cast<int64_t>(sh.threads) * kChains - 1) /
                                 (static_cast<int64_t>(sh.threads) * kChains);
            const int64_t cap = static_cast<int64_t>(cus) * 32;
            sh.blocks = static_cast<int32_t>(need < cap ? (need > 0 ? need : 1) : cap);
        }
        return sh;
    };
    const Shape walk = shape(kWalkThreads);
    Shape fix = shape(kFixThreads);
    if (kLds && fix.threads > fix_threads_for(64) / kChains) {
        // more than 512 merge walks per CU: two blocks of half the size per CU (two tables
        // fit the LDS), so the fix pass keeps its 64-byte rounds (a 1024-thread block would
        // have to use 32-byte rounds)
        const int32_t half = fix.threads / 2;
        const int64_t need = (a.segments + static_cast<int64_t>(half) * kChains - 1) /
                             (static_cast<int64_t>(half) * kChains);
        const int64_t cap = 2 * static_cast<int64_t>(cus);
        fix.threads = half;
        fix.blocks = static_cast<int32_t>(need < cap ? need : cap);
    }
    dfa_split_walk_kernel<T, kLds, kChains><<<walk.blocks, walk.threads, lds, stream>>>(tab, a);
    GFSM_HIP(hipGetLastError());
    if (a.segments > 1) {
        // 64-byte rounds when the fix blocks fit their VGPR budget, else 32-byte rounds
        auto launch_fix = [&](int32_t repair) {
            if (fix.threads <= fix_threads_for(64) / kChains)
                dfa_split_fix_kernel<T, kLds, kChains, 64><<<fix.blocks, fix.threads, lds, stream>>>(tab, a, repair);
            else
                dfa_split_fix_kernel<T, kLds, kChains, 32><<<fix.blocks, fix.threads, lds, stream>>>(tab, a, repair);
            GFSM_HIP(hipGetLastError());
        };
        launch_fix(0);
        // The repair pass re-walks heads whose assumed entry was wrong, one lane each. With
        // maps and segments of 2 KiB or more there are few such heads (the speculative walk
        // of the previous segment failed to merge), and maps resolve them in parallel.
        if (!a.map_idx || a.seg < 2048) {
            launch_fix(1);
        }
        const int64_t eblocks64 = (a.segments + 255) / 256;
        const int32_t eblocks = static_cast<int32_t>(eblocks64 < 65536 ? eblocks64 : 65536);
        dfa_split_event_kernel<<<eblocks, 256, 0, stream>>>(a);
        GFSM_HIP(hipGetLastError());
        if constexpr (kLds && sizeof(T) <= 2) {
            if (a.map_idx) {
                const size_t mlds = static_cast<size_t>(a.table_bytes) +
                                    static_cast<size_t>(kMapPartIntervals) * a.n * 4;
                ensure_dynamic_lds(reinterpret_cast<const void *>(&dfa_split_map_kernel<T>), mlds);
                const int32_t mblocks = cus;
                dfa_split_map_kernel<T><<<mblocks, kMapThreads, mlds, stream>>>(tab, a);
                GFSM_HIP(hipGetLastError());
            }
        }
        // The finish pass reads the table from L2: it walks only the rare heads that have
        // neither a consistent fix nor a map, and a launch that reserved the table's LDS
        // would wait for a CU with that much free (behind other batches' walk and fix
        // blocks, in a stream of batches)
        dfa_split_finish_kernel<T, false><<<(a.count + 255) / 256, 256, 0, stream>>>(tab, a);
        GFSM_HIP(hipGetLastError());
    }
}

void launch_dfa_split(const DfaDev &d, const DfaSplitPlan &plan, const uint8_t *data,
                      const int64_t *off, int32_t count, int64_t total_bytes, int32_t *flags,
                      int32_t *lens, void *scratch, hipStream_t stream) {
    if (count <= 0) return;
    if (reinterpret_cast<uintptr_t>(data) & 15)
        throw Error(3, "split needs the input buffer 16-byte aligned");
    if (plan.segments >= (int64_t(1) << 31))  // segment numbers are stored as int32
        throw Error(1, "split: more than 2^31 segments; raise segment_bytes");
    SplitArgs a{};
    a.data = data;
    a.off = off;
    a.count = count;
    a.total = total_bytes;
    a.seg = plan.seg_bytes;
    a.rshift = plan.ckpt_shift;
    a.nck = plan.seg_bytes >> plan.ckpt_shift;
    a.segments = plan.segments;
    a.start = d.start_entry;
    a.acc = d.accept_from;
    a.col = d.col_bytes;
    a.table_bytes = d.col_bytes * kAlphabet;
    a.flags = flags;
    a.lens = lens;
    const size_t S = static_cast<size_t>(plan.segments > 0 ? plan.segments : 1);
    char *base = static_cast<char *>(scratch);
    a.done = reinterpret_cast<int32_t *>(base);
    base += align256(static_cast<size_t>(count) * 4);
    a.first_event = reinterpret_cast<unsigned long long *>(base);
    base += align256(static_cast<size_t>(count) * 8);
    a.head_exit = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.tail_exit = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.fixed_entry = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.fixed_exit = reinterpret_cast<uint32_t *>(base);
    base += align256(S * 4);
    a.fixed_acc = reinterpret_cast<int32_t *>(base);
    base += align256(S * 4);
    a.seg_input = reinterpret_cast<int32_t *>(base);
    base += align256(S * 4);
    a.flagmask = reinterpret_cast<unsigned long long *>(base);
    base += align256(S * 8);
    a.ckpt = base;
    base += align256(S * static_cast<size_t>(plan.seg_bytes >> plan.ckpt_shift) * (d.wide ? 4 : 2));
    a.n = d.n;
    if (plan.maps) {
        const size_t cap = static_cast<size_t>(plan.map_cap), nn = static_cast<size_t>(d.n);
        a.map_cap = plan.map_cap;
        a.map_idx = reinterpret_cast<int32_t *>(base);
        base += align256(S * 4);
        a.map_list = reinterpret_cast<int32_t *>(base);
        base += align256(cap * 4);
        a.map_count = reinterpret_cast<unsigned int *>(base);
        base += 256;
        a.map_exit = reinterpret_cast<uint16_t *>(base);
        base += align256(cap * kMapParts * nn * 2);
        a.map_acc = reinterpret_cast<int32_t *>(base);
    }
    split_init_kernel<<<(count + 255) / 256, 256, 0, stream>>>(flags, lens, a.done,
                                                                a.first_event, count,
                                                                a.map_count);
    GFSM_HIP(hipGetLastError());
    if (plan.segments <= 0) return;
    const bool lds = lds_fits(d);
    if (d.wide) {
        split_launch_passes<uint32_t, false, 1>(d, a, stream);
    } else if (d.esz == 1) {
        if (plan.chains == 4) split_launch_passes<uint8_t, true, 4>(d, a, stream);
        else if (plan.chains == 2) split_launch_passes<uint8_t, true, 2>(d, a, stream);
        else split_launch_passes<uint8_t, true, 1>(d, a, stream);
    } else if (lds) {
        if (plan.chains == 4) split_launch_passes<uint16_t, true, 4>(d, a, stream);
        else if (plan.chains == 1) split_launch_passes<uint16_t, true, 1>(d, a, stream);
        else if (plan.chains == 2) split_launch_passes<uint16_t, true, 2>(d, a, stream);
        else split_launch_passes<uint16_t, true, 1>(d, a, stream);
    } else {
        if (plan.chains == 4) split_launch_passes<uint16_t, false, 4>(d, a, stream);
        else if (plan.chains == 1) split_launch_passes<uint16_t, false, 1>(d, a, stream);
        else split_launch_passes<uint16_t, false, 2>(d, a, stream);
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
    const uint32_t tb = d.col_bytes * kAlphabet;
    if (d.wide)
        dfa_walk_kernel<uint32_t, false><<<grid, 256, 0, stream>>>(
            static_cast<const uint32_t *>(d.table), tb, d.col_bytes, d.start_entry,
            d.accept_from, data, off, count, flags, lens);
    else if (d.esz == 1)
        dfa_walk_kernel<uint8_t, false><<<grid, 256, 0, stream>>>(
            static_cast<const uint8_t *>(d.table), tb, d.col_bytes, d.start_entry,
            d.accept_from, data, off, count, flags, lens);
    else
        dfa_walk_kernel<uint16_t, false><<<grid, 256, 0, stream>>>(
            static_cast<const uint16_t *>(d.table), tb, d.col_bytes, d.start_entry,
            d.accept_from, data, off, count, flags, lens);
    GFSM_HIP(hipGetLastError());
}

void launch_dfa_lds(const DfaDev &d, const uint8_t *data, const int64_t *off, int32_t count,
                    int32_t *flags, int32_t *lens, hipStream_t stream) {
    if (count <= 0) return;
    if (!lds_fits(d)) throw Error(3, "lds: the table does not fit in LDS");
    const uint32_t tb = d.col_bytes * kAlphabet;
    allow_big_lds(&dfa_walk_kernel<uint16_t, true>, tb);
    allow_big_lds(&dfa_walk_kernel<uint8_t, true>, tb);
    // One block per CU holds the table, so the block is the lanes a CU gets: 256, or more
    // when the batch has that many inputs per CU (more reads in flight per CU; 131,072 x
    // 2 KiB inputs ran 1.7 TB/s at 256 lanes, latency-bound).
    const int cus = device_cu_count();
    const int64_t per_cu = (static_cast<int64_t>(count) + cus - 1) / cus;
    const int threads = per_cu >= 1024 ? 1024 : (per_cu >= 512 ? 512 : 256);
    int grid = (count + threads - 1) / threads;
    const int cap = 2 * cus;
    if (grid > cap) grid = cap;
    if (d.esz == 1)
        dfa_walk_kernel<uint8_t, true><<<grid, threads, tb, stream>>>(
            static_cast<const uint8_t *>(d.table), tb, d.col_bytes, d.start_entry,
            d.accept_from, data, off, count, flags, lens);
    else
        dfa_walk_kernel<uint16_t, true><<<grid, threads, tb, stream>>>(
            static_cast<const uint16_t *>(d.table), tb, d.col_bytes, d.start_entry,
            d.accept_from, data, off, count, flags, lens);
    GFSM_HIP(hipGetLastError());
}

}  // namespace gfsm
